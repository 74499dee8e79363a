/*
 * scess_cc.h - C++ method API of the closed-shell CCSD(T) path, shaped like the
 * reference's SCESS::CC namespace (src/Quantum/CC/CC.h:99-164) and implemented
 * over the C ABI of scess_gpu.h (scess_amd/csrc/cc_api.cpp, in libscess_gpu.so).
 *
 * Eigen is not a dependency: SCESS::Tensor<N> is a plain ColMajor (first index
 * fastest) owning tensor with the same element order as Eigen::Tensor<double, N>,
 * so data() pointers of Eigen tensors can be copied in without reordering.
 * Errors are reported as std::runtime_error, as the reference propagates C++
 * exceptions.
 */
#ifndef SCESS_CC_H
#define SCESS_CC_H

#include <array>
#include <cstdint>
#include <sstream>
#include <vector>

namespace SCESS {

template <int N>
class Tensor {
public:
    Tensor() { dims_.fill(0); }
    template <class... I>
    explicit Tensor(I... d) : dims_{{static_cast<int64_t>(d)...}} {
        static_assert(sizeof...(I) == N, "rank");
        data_.assign(size(), 0.0);
    }
    int64_t dimension(int k) const { return dims_[k]; }
    int64_t size() const {
        int64_t n = 1;
        for (auto d : dims_)
            n *= d;
        return n;
    }
    double *data() { return data_.data(); }
    const double *data() const { return data_.data(); }
    template <class... I>
    double &operator()(I... idx) {
        return data_[offset({static_cast<int64_t>(idx)...})];
    }
    template <class... I>
    double operator()(I... idx) const {
        return data_[offset({static_cast<int64_t>(idx)...})];
    }
    void setZero() { std::fill(data_.begin(), data_.end(), 0.0); }

private:
    int64_t offset(std::array<int64_t, N> ix) const {
        int64_t off = 0, s = 1;
        for (int k = 0; k < N; k++) {
            off += ix[k] * s;
            s *= dims_[k];
        }
        return off;
    }
    std::array<int64_t, N> dims_;
    std::vector<double> data_;
};

// The cancellation contract of DH::Thread used by the reference (ShouldStop() const).
class Thread {
public:
    virtual ~Thread() = default;
    virtual bool ShouldStop() const = 0;
};

namespace CC {

using T1Tensor = Tensor<2>;
using T2Tensor = Tensor<4>;
using TEITensor = Tensor<4>;

// src/Quantum/CC/CC.h:105-115
struct RCCSD_Wfn {
    double E_CCSD = 0.0;
    double E_F12_Coupling = 0.0;
    double E_T = 0.0;
    bool converged = false;
    T1Tensor t1;
    T2Tensor t2;
};

// src/Quantum/CC/CC.h:121-124
struct RCCSD_F12Coupling {
    Tensor<4> residual;
    Tensor<4> energy;
};

// src/Quantum/CC/CC.h:136-141: eri is <pq|rs> over all spatial MOs, occupied first.
RCCSD_Wfn SolveRCCSD_DIIS(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                          const RCCSD_F12Coupling *f12 = nullptr, const Thread *thread = nullptr,
                          std::stringstream *console = nullptr);

// src/Quantum/CC/CC.h:154-158
double RCCSD_TriplesCorrection(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                               const T1Tensor &t1, const T2Tensor &t2, const Thread *thread = nullptr);

// The CC part of RCCSD_T (src/Quantum/CC/RCCSD_T.cpp:170-195) from MO integrals and the HF energy:
// CCSD, then (T) only when CCSD converged, with the reference's console output.
RCCSD_Wfn RCCSD_T_FromMO(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                         double E_HF, const Thread *thread = nullptr, std::stringstream *console = nullptr);

}  // namespace CC

namespace HF {
// The fields of HF_Wfn (src/Quantum/HF/HF.h:26-67) that the closed-shell CC methods read.
// C_Alpha is nao x nmo ColMajor (column p = MO p), as Eigen::MatrixXd.
struct HF_Wfn {
    double E_HF = 0.0;
    int64_t nao = 0, nmo = 0;
    std::vector<double> C_Alpha;
    std::vector<double> Ep_Alpha;
    std::vector<int> Occ_Alpha;
};
}  // namespace HF

namespace Integrals {
using TEITensor = Tensor<4>;
// The AO two-electron tensor of MolInts (src/Quantum/Integrals/Integrals.h:28-69), chemists' (mu nu|kap lam).
struct MolInts {
    TEITensor eri;
};
// TransformTEI (src/Quantum/Integrals/IntegralTransforms.cpp:29-59) on the GPU.
TEITensor TransformTEI(const TEITensor &tei, const std::vector<double> &C, int64_t nmo, bool convertChemToPhys = false);
}  // namespace Integrals

namespace CC {
// RCCSD_T (src/Quantum/CC/CC.h:160-164, RCCSD_T.cpp:170-195): nOcc = leading occupied orbitals of
// Occ_Alpha, <pq|rs> = TransformTEI(mints.eri, C_Alpha, true), CCSD, then (T) when CCSD converged.
RCCSD_Wfn RCCSD_T(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints, const Thread *thread = nullptr,
                  std::stringstream *console = nullptr);

}  // namespace CC
}  // namespace SCESS

#endif /* SCESS_CC_H */
