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
#include <map>
#include <sstream>
#include <string>
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

// Gradients of E(T) (no counterpart in the reference): with respect to the three integral blocks the
// functional reads (the repack of RCCSD_T.cpp:37-51: ovvv(i,a,b,f) = (ia|bf) = eri(i, o+b, o+a, o+f),
// ovoo(i,a,m,j) = (ia|mj) = eri(i, m, o+a, j), ovov(i,a,j,b) = (ia|jb) = eri(i, j, o+a, o+b)), with
// respect to <pq|rs> itself (those three index patterns are disjoint, so eri holds their gradients
// scattered back and zeros elsewhere), and with respect to t1 and t2.
struct TriplesGradient {
    Tensor<4> ovvv, ovoo, ovov;
    TEITensor eri;
    T1Tensor t1;
    T2Tensor t2;
};

// Autograd-Function-style forward/backward pair for (T) (no counterpart in the reference).
// forward() is RCCSD_TriplesCorrection (same arguments, same E(T) in FP64, same cancellation: 0.0)
// and keeps the inputs resident on the GPU; its contractions can also run in FP32, split-BF16x3 or
// exact int8-digit products (OZAKI), with FP64 denominators and sums. backward(dE) returns dE times
// the gradients of E(T), W being recomputed per occupied triple in the precision given to the
// constructor (FP64, FP32, BF16X3 or OZAKI; FP64 by default), the back-contractions in FP64. backward()
// before a successful forward() throws; a cancelled backward() returns an empty TriplesGradient.
// One object holds one GPU context; it is not copyable.
class TriplesFunction {
public:
    enum class Precision { FP64, FP32, BF16X3, OZAKI };
    explicit TriplesFunction(Precision wRecompute = Precision::FP64, int device = 0);
    ~TriplesFunction();
    TriplesFunction(const TriplesFunction &) = delete;
    TriplesFunction &operator=(const TriplesFunction &) = delete;
    double forward(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                   const T1Tensor &t1, const T2Tensor &t2, const Thread *thread = nullptr,
                   Precision precision = Precision::FP64);
    TriplesGradient backward(double dE = 1.0, const Thread *thread = nullptr);

private:
    struct State;
    State *s_;
};

// The CC part of RCCSD_T (src/Quantum/CC/RCCSD_T.cpp:170-195) from MO integrals and the HF energy:
// CCSD, then (T) only when CCSD converged, with the reference's console output.
RCCSD_Wfn RCCSD_T_FromMO(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                         double E_HF, const Thread *thread = nullptr, std::stringstream *console = nullptr);

}  // namespace CC

// src/Quantum/Molecule.h: atoms (positions in bohr) and the reference's built-in molecules
// (src/Quantum/Molecule.cpp:194-231, 281-294).
struct Atom {
    std::string symbol;
    int Z = 0;
    double position[3] = {0.0, 0.0, 0.0};
};
class Molecule {
public:
    std::vector<Atom> atoms;
    int NumElectrons() const;
    double NuclearRepulsion() const;
    static Molecule GenerateWaterMolecule();
    static Molecule GenerateHydrogenMolecule();
    static Molecule GenerateLithiumHydrideMolecule();
    static Molecule GenerateAmmoniaMolecule();
    // "H2O", "H2", "LiH", "NH3": the names the reference's command line accepts (Application.cpp:105-110)
    static Molecule FromName(const std::string &name);
};

namespace Basis {
// One contracted shell of an element (src/Quantum/Basis/BasisSet.cpp:139-181): spherical unless the BSE
// shell is "gto_cartesian"; coefficients of the normalised primitives, as in the BSE file.
struct ShellData {
    int l = 0;
    bool pure = false;
    std::vector<double> exponents, coefficients;
};
struct BasisSet {
    std::string name;
    std::map<int, std::vector<ShellData>> elements;  // by atomic number
};
// A BSE JSON file: a path, or a basis set name looked up (case-insensitively, "6-31G*" = "6-31G(d)") as
// <name>.json in $SCESS_BASIS_DIR or its subdirectories (the layout of the reference's Assets/Basis).
BasisSet LoadBasisSet(const std::string &nameOrPath);
}  // namespace Basis

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
// MolInts (src/Quantum/Integrals/Integrals.h:28-69): overlap S, kinetic T, nuclear attraction V and the
// AO two-electron tensor, chemists' (mu nu|kap lam).
struct MolInts {
    Tensor<2> S, T, V;
    TEITensor eri;
};
// EvaluateMolecularIntegrals (src/Quantum/Integrals/Integrals.cpp:28-80) for shells up to f, on the host.
MolInts EvaluateMolecularIntegrals(const Molecule &molecule, const Basis::BasisSet &basis);
// TransformTEI (src/Quantum/Integrals/IntegralTransforms.cpp:29-59) on the GPU.
TEITensor TransformTEI(const TEITensor &tei, const std::vector<double> &C, int64_t nmo, bool convertChemToPhys = false);
}  // namespace Integrals

namespace HF {
// RHF (src/Quantum/HF/HF.h:80-86, RHF.cpp:34-211): closed-shell SCF from the core-Hamiltonian guess with
// symmetric orthogonalisation and Pulay DIIS (6 Fock matrices); converged when the energy changes by
// less than convergeLimit and the rms of X^T (FDS - SDF) X is below sqrt(convergeLimit). Throws when it
// does not converge; returns an empty HF_Wfn when the thread asks to stop.
HF_Wfn RHF(const Molecule &molecule, const Basis::BasisSet &basis, double convergeLimit = 1e-5,
           size_t scfMaxIterations = 100, Integrals::MolInts *mintsOut = nullptr, const Thread *thread = nullptr,
           std::stringstream *console = nullptr);
}  // namespace HF

namespace CC {
// RCCSD_T (src/Quantum/CC/CC.h:160-164, RCCSD_T.cpp:170-195): nOcc = leading occupied orbitals of
// Occ_Alpha, <pq|rs> = TransformTEI(mints.eri, C_Alpha, true), CCSD, then (T) when CCSD converged.
RCCSD_Wfn RCCSD_T(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints, const Thread *thread = nullptr,
                  std::stringstream *console = nullptr);

// RCCSD_DIIS (src/Quantum/CC/CC.h:143-146, RCCSD.cpp:396-405): nOcc from Occ_Alpha, <pq|rs> =
// TransformTEI(mints.eri, C_Alpha, true), SolveRCCSD_DIIS, and the total-energy line when converged.
RCCSD_Wfn RCCSD_DIIS(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints, const Thread *thread = nullptr,
                     std::stringstream *console = nullptr);

// Spin-orbital CCSD and CCSD(T) (src/Quantum/CC/CC.h:27-33,57-73; CCSD.cpp, CCSD_T.cpp): the
// reference's independent implementations in (2n)^4 antisymmetrised integrals, kept to test the
// closed-shell solvers. Host code, tiny systems only (the regression runner's H2 and water STO-3G).
struct CCSD_Wfn {
    double E_CCSD = 0.0;
    double E_T = 0.0;
    bool converged = false;
};
CCSD_Wfn CCSD_DIIS(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints);
CCSD_Wfn CCSD_T(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints);
// The same from MO integrals: eri is <pq|rs> (spatial, occupied first), E_T only when withTriples.
CCSD_Wfn CCSD_DIIS_FromMO(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                          bool withTriples);

}  // namespace CC

namespace HF {
// MP2_RHF_ClosedShellAO (src/Quantum/HF/HF.h:148-151): the closed-shell MP2 correlation energy from the
// AO integrals (transformed with TransformTEI on the GPU).
double MP2_RHF_ClosedShellAO(const HF_Wfn &wfn, const Integrals::TEITensor &aoERI);
}  // namespace HF

namespace Tests {
// RunRegressionTests (src/SCESS/Tests.cpp:494-737, `SCESS test`): the closed-shell CC rows, each printed
// as [PASS] / [FAIL] with value, reference, |diff| and tolerance; returns the number of failures.
// Basis sets are looked up as Basis::LoadBasisSet does ($SCESS_BASIS_DIR).
int RunRegressionTests();
}  // namespace Tests

namespace Driver {
// src/SCESS/Driver.h:13-40, Driver.cpp:51-65,173-231: RHF on the basis set (a BSE JSON name or path, as
// Basis::LoadBasisSet), then the method, printing the reference's blocks to stdout ("--- Performing ...
// calculation on <moleculeName> ---", the SCF and CC logs, the single-point energy summary). The SCF is
// converged to 1e-10 (the reference's drivers use 1e-8) so that the CC energies carry no SCF residue.
// The reference's drivers return void; these return whether the method converged (RHF: true when it
// returns). Errors propagate as exceptions.
bool RHF(const char *basisSetName, const Molecule &mol, const char *moleculeName = nullptr);
bool CCSD(const char *basisSetName, const Molecule &mol, const char *moleculeName = nullptr);
bool CCSD_T(const char *basisSetName, const Molecule &mol, const char *moleculeName = nullptr);
}  // namespace Driver
}  // namespace SCESS

#endif /* SCESS_CC_H */
