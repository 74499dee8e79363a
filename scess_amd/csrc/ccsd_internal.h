// Device tensors, the einsum engine and the closed-shell CCSD state.
#pragma once

#include "triples_internal.h"

#include <cstdint>
#include <string>
#include <vector>

namespace scess {

// Row-major FP64 device tensor owning its allocation.
struct DTensor {
    double *p = nullptr;
    std::vector<int64_t> dims;
    hipStream_t s = nullptr;  // stream-ordered allocation (hipMallocAsync / hipFreeAsync)
    DTensor() = default;
    DTensor(const DTensor &) = delete;
    DTensor &operator=(const DTensor &) = delete;
    DTensor(DTensor &&o) noexcept;
    DTensor &operator=(DTensor &&o) noexcept;
    ~DTensor();
    void release();
    int64_t size() const;
};

// Stream-ordered tensor operations; errors are thrown as std::runtime_error.
struct Engine {
    hipStream_t stream = nullptr;
    double *dotbuf = nullptr;
    int prec = 0;  // product precision of the einsum contractions (P64, P32, PB3), see GemmArgs::prec
    ~Engine();
    DTensor alloc(std::vector<int64_t> dims, bool zero = false);
    void upload(DTensor &t, const double *host);
    void download(const DTensor &t, double *host);
    void copy(DTensor &dst, const DTensor &src);
    void axpby(double *out, const double *x, const double *y, double alpha, double beta, int64_t n);
    void permute(const std::string &li, const double *in, const std::vector<int64_t> &idims, const std::string &lo,
                 double *out, const std::vector<int64_t> &odims, double alpha, double beta);
    // the permutation of `in` (labels li -> lo, output dims odims) written as pre-split BF16x3 rows of its
    // last K elements (rows padded to a multiple of 32 k); returned as rows x Kp/2 eight-byte slots
    DTensor permute_split(const std::string &li, const double *in, const std::vector<int64_t> &idims,
                          const std::string &lo, const std::vector<int64_t> &odims, int64_t K);
    // C = beta C + alpha einsum(spec, A, B); B == nullptr for a pure permutation
    void einsum(const std::string &spec, const DTensor &A, const DTensor *B, DTensor &C, double alpha, double beta);
    double dot(const double *x, const double *y, int64_t n);  // compensated, deterministic
};

// Resident integral blocks of SolveRCCSD_DIIS (src/Quantum/CC/RCCSD.cpp:31-44), row-major with
// the reference's index order, plus the optional Cholesky factor of (ac|bd).
struct CcsdState {
    bool have = false;
    int64_t o = 0, v = 0, naux = 0;
    DTensor ovov, oovv, ovvo, ovvv, ovoo, oooo_phys, vvvv_phys, oovv_phys, L_oovv_phys, L_ovov, Lvv, e;
    DTensor t2driver, f12energy;
    void make_tau(Engine &E, const DTensor &t1, const DTensor &t2, DTensor &tau);
    void ladder_from_factor(Engine &E, const DTensor &tau, DTensor &R);
    void update(Engine &E, DTensor &t1, DTensor &t2);  // one Jacobi update (UpdateAmplitudes)
    double energy(Engine &E, const DTensor &t1, const DTensor &t2);
};

// t1 /= e_i - e_a and t2 /= e_i + e_j - e_a - e_b in place (row-major)
hipError_t ccsd_jacobi(double *t1, double *t2, const double *e, int o, int v, hipStream_t s);

}  // namespace scess
