/*
 * scess_gpu.h - C ABI of libscess_gpu.so, the MI355X (gfx950) closed-shell
 * CCSD(T) hot path of SCESS.
 *
 * The reference has no C ABI: its boundary is the C++ method API of
 * libQuantumSCESS, src/Quantum/CC/CC.h:136-164. The entry points below are what
 * a binding of that API to this library calls; the C++ mirror of CC.h in
 * include/scess_cc.h (scess_amd/csrc/cc_api.cpp) and the Python mirror in
 * scess_amd/triples.py sit on top of them.
 *
 * Conventions
 *   - All host arrays are FP64 in Eigen's default ColMajor layout (first index
 *     fastest), the layout of the reference's Eigen::Tensor<double, N>.
 *   - Integral blocks are in chemists' notation, named as in
 *     src/Quantum/CC/RCCSD.cpp:31-44:  ovvv(i,a,b,f) = (ia|bf),
 *     ovoo(i,a,m,j) = (ia|mj), ovov(i,a,j,b) = (ia|jb).
 *   - Return codes: SCESS_OK (0) or a negative SCESS_ERR_*; no exception
 *     crosses the ABI. scess_last_error() describes the last failure.
 *   - One context per host thread. Calls are synchronous (they return after the
 *     device work has finished) and run on the context's own HIP stream.
 *   - Cancellation: an optional should_stop callback is polled between batches
 *     of occupied triples (the reference polls DH::Thread::ShouldStop() per
 *     triple, src/Quantum/CC/RCCSD_T.cpp:123,161-162); a cancelled call
 *     returns SCESS_CANCELLED and writes 0.0 as the energy, as the reference
 *     returns 0.0.
 */
#ifndef SCESS_GPU_H
#define SCESS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scess_ctx scess_ctx;

enum {
    SCESS_OK = 0,
    SCESS_ERR_ARG = -1,    /* invalid argument or unsupported size */
    SCESS_ERR_HIP = -2,    /* a HIP runtime call failed */
    SCESS_ERR_OOM = -3,    /* device memory exhausted */
    SCESS_ERR_STATE = -4,  /* call out of order (e.g. forward before inputs) */
    SCESS_CANCELLED = -5   /* should_stop returned non-zero */
};

/* Precision of the W contractions; denominators, V, r3 and every energy
 * reduction are FP64 in all modes. */
enum {
    SCESS_PREC_FP64 = 0,   /* FP64 operands, FP64 MFMA */
    SCESS_PREC_FP32 = 1,   /* FP32 operands, FP32 MFMA, FP32 accumulation */
    SCESS_PREC_BF16X3 = 2, /* x = hi + lo in bf16; hi*hi + hi*lo + lo*hi on bf16 MFMA, FP32 accumulation */
    SCESS_PREC_OZAKI = 3   /* each tensor scaled by a power of two and split into S signed 8-bit
                              digits (Ozaki; S = 5, 6 or 7, scess_t_set_int8_digits); the S(S+1)/2 leading
                              digit products summed exactly on int8 MFMA, FP64 accumulation; product error
                              <= ~(S+1) 2^-8S of max|A| max|B| */
};

/* Inputs of the (T) functional. Replaces the (eri, orbitalEnergies, nOcc, t1,
 * t2) arguments of RCCSD_TriplesCorrection (src/Quantum/CC/CC.h:154-158) by
 * the three integral blocks the functional reads (RCCSD_T.cpp:27-56). */
typedef struct scess_t_in {
    int64_t o;            /* occupied orbitals (nOcc) */
    int64_t v;            /* virtual orbitals */
    const double *e;      /* [o+v] orbital energies, occupied first */
    const double *t1;     /* (i,a)      o x v */
    const double *t2;     /* (i,j,a,b)  o x o x v x v, t_ij^ab = t_ji^ba */
    const double *ovvv;   /* (i,a,b,f)  (ia|bf) */
    const double *ovoo;   /* (i,a,m,j)  (ia|mj) */
    const double *ovov;   /* (i,a,j,b)  (ia|jb) */
} scess_t_in;

typedef struct scess_t_stats {
    double seconds;            /* device time of the whole call (HIP events) */
    double w_build_ms;         /* summed device time of the W-build launches */
    double epilogue_ms;        /* summed span of the V-piece, epilogue and partial-sum launches on the
                                  epilogue stream; they overlap the next batch's W-build, so this is
                                  not additive with w_build_ms */
    int64_t w_build_launches;
    int64_t epilogue_launches;
    int64_t n_ijk;             /* unique occupied triples i>=j>=k evaluated */
    double flops_algorithmic;  /* 2 o^3 v^3 (o+v) */
    double flops_executed;     /* MFMA work actually issued, padding included */
} scess_t_stats;

/* Gradients of E(T), same layouts and sizes as the inputs (host, FP64). */
typedef struct scess_t_grad {
    double *ovvv;
    double *ovoo;
    double *ovov;
    double *t1;
    double *t2;
} scess_t_grad;

/* -- context ---------------------------------------------------------------- */
int scess_create(scess_ctx **ctx, int device);
void scess_destroy(scess_ctx *ctx);
const char *scess_last_error(const scess_ctx *ctx);
int scess_set_should_stop(scess_ctx *ctx, int (*should_stop)(void *), void *user);
/* Occupied-triple batch size (0 = automatic). */
int scess_set_batch(scess_ctx *ctx, int64_t nbatch);

/* -- (T) ----------------------------------------------------------------------
 * scess_t_set_inputs uploads the inputs once and repacks them on the device
 * into the layouts of every precision in prec_mask (bit 1 << SCESS_PREC_*);
 * they stay resident in HBM for any number of forward/backward calls. This is
 * the device counterpart of the serial repack of RCCSD_T.cpp:27-56. */
int scess_t_set_inputs(scess_ctx *ctx, const scess_t_in *in, int prec_mask);
/* Digits per element of SCESS_PREC_OZAKI (5, 6 or 7; default 6), applied by the next
 * scess_t_set_inputs (no counterpart in the reference). */
int scess_t_set_int8_digits(scess_ctx *ctx, int digits);

/* E(T) (RCCSD_TriplesCorrection, src/Quantum/CC/RCCSD_T.cpp:17-168) from the
 * resident inputs. stats may be NULL. */
int scess_t_forward(scess_ctx *ctx, int prec, double *e_t, scess_t_stats *stats);

/* As scess_t_forward, and optionally also:
 *   F_ijk     [scess_t_num_ijk] per unique occupied triple (enumeration i outer,
 *             j <= i, k <= j), F = sum_abc W r3(Y)/D before the 2/deg weight;
 *   partials  [scess_t_num_ijk * scess_t_num_partials] pieces of F per ijk (each
 *             the sum over a run of virtual-block orbits), for local finite
 *             differences. */
int scess_t_forward_ex(scess_ctx *ctx, int prec, double *e_t, double *F_ijk, double *partials,
                       scess_t_stats *stats);
int64_t scess_t_num_ijk(const scess_ctx *ctx);
/* E(T) from the resident inputs without allocating device memory (no counterpart in the reference).
 * For callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation): W, the V pieces, partials, F_ijk and E(T) live in a 64 MiB device array of
 * the library (one per GPU, held until the process ends), one occupied triple per launch on one
 * stream. Sizes whose scratch does not fit (mostly W: 2 vR^3 bytes-per-element, vR = v rounded up to
 * 16; FP64 and OZAKI up to v = 160, FP32 and BF16X3 up to v = 192) are rejected with SCESS_ERR_ARG. The result equals scess_t_forward's bit for bit.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU race. */
int scess_t_forward_noalloc(scess_ctx *ctx, int prec, double *e_t);
/* Restrict scess_t_forward / scess_t_backward to one of nshards contiguous, equal-count ranges of the
 * occupied-triple work list (no counterpart in the reference; default 0 of 1 = everything). The
 * result is that shard's share: the sum over shards of E(T), F_ijk and the gradients is the full
 * result, which is how the work would be split over several GPUs. */
int scess_t_set_shard(scess_ctx *ctx, int64_t shard, int64_t nshards);
int64_t scess_t_num_partials(const scess_ctx *ctx);

/* Adjoint of E(T) (no counterpart in the reference): writes dE * dE(T)/d(input)
 * for every input entry, in the input layouts, into the caller's host arrays
 * (overwritten). Every array entry is an independent variable read exactly as
 * the forward reads it. W is recomputed per occupied triple in precision prec
 * (FP64, FP32, BF16X3 or OZAKI); the back-contractions run in FP64 and need SCESS_PREC_FP64 packed by
 * scess_t_set_inputs. Deterministic: triples are processed in a fixed order. */
int scess_t_backward(scess_ctx *ctx, int prec, double dE, scess_t_grad *grad);

/* -- closed-shell CCSD ---------------------------------------------------------
 * Integral blocks of SolveRCCSD_DIIS (src/Quantum/CC/RCCSD.cpp:31-44), chemists'
 * notation, ColMajor with the index order as named. Either vvvv or chol_vv must
 * be given; with chol_vv the particle-particle ladder regenerates (ac|bd) tiles
 * and no v^4 tensor is formed. */
typedef struct scess_ccsd_in {
    int64_t o, v;
    const double *e;          /* [o+v] canonical orbital energies, occupied first */
    const double *ovov;       /* (k,c,l,d) = (kc|ld) */
    const double *oovv;       /* (k,i,a,c) = (ki|ac) */
    const double *ovvo;       /* (k,c,a,i) = (kc|ai) */
    const double *ovvv;       /* (k,d,a,c) = (kd|ac) */
    const double *ovoo;       /* (l,c,k,i) = (lc|ki) */
    const double *oooo;       /* (k,i,l,j) = (ki|lj) */
    const double *vvvv;       /* (a,c,b,d) = (ac|bd), or NULL */
    const double *chol_vv;    /* (a,c,P): (ac|bd) = sum_P L(a,c,P) L(b,d,P), or NULL */
    int64_t naux;             /* P extent of chol_vv */
    const double *t2_driver;  /* optional (i,j,a,b): RCCSD_F12Coupling::residual (CC.h:121-124) */
    const double *f12_energy; /* optional (i,j,a,b): RCCSD_F12Coupling::energy */
} scess_ccsd_in;

typedef struct scess_ccsd_result {
    double E_CCSD;            /* correlation energy (coupling included), as RCCSD_Wfn::E_CCSD */
    double E_F12_Coupling;    /* RCCSD_Wfn::E_F12_Coupling */
    int converged;
    int iterations;
    double last_dE, last_rms;
} scess_ccsd_result;

/* Per-iteration progress (the reference prints "CCSD(n): E(corr) dE rms(dt)", RCCSD.cpp:366-368). */
typedef void (*scess_ccsd_log)(void *user, int iteration, double e_corr, double dE, double rms);

int scess_ccsd_set_inputs(scess_ctx *ctx, const scess_ccsd_in *in);
/* One Jacobi update t <- R(t)/D (UpdateAmplitudes, RCCSD.cpp:101-237); t1 (i,a), t2 (i,j,a,b) ColMajor. */
int scess_ccsd_update(scess_ctx *ctx, const double *t1, const double *t2, double *t1_new, double *t2_new);
/* E = sum (2<ij|ab> - <ij|ba>) tau_ij^ab (+ F12 coupling) (RCCSD.cpp:85-98). */
int scess_ccsd_energy(scess_ctx *ctx, const double *t1, const double *t2, double *e_corr);
/* SolveRCCSD_DIIS (RCCSD.cpp:295-394): MP2 guess, <= 100 iterations, DIIS (8 vectors) from iteration 2,
 * converged when dE < 1e-10 and rms(dt) < 1e-8. Amplitudes, DIIS vectors and dot products stay on the
 * GPU in FP64; the bordered DIIS system is solved on the host with full pivoting. t1_out / t2_out may
 * be NULL. */
int scess_ccsd_solve(scess_ctx *ctx, double *t1_out, double *t2_out, scess_ccsd_result *res, scess_ccsd_log log,
                     void *user);

/* Product precision of the residual contractions of scess_ccsd_update / scess_ccsd_solve (no counterpart
 * in the reference, which is FP64 throughout): SCESS_PREC_FP64 (default), SCESS_PREC_FP32 or
 * SCESS_PREC_BF16X3. Inputs are rounded in registers, each 16-deep K chunk is summed on the MFMA in FP32
 * and accumulated in FP64; the amplitude update, energies and DIIS stay FP64. */
int scess_ccsd_set_precision(scess_ctx *ctx, int prec);

/* TransformTEI (IntegralTransforms.cpp:29-59): (pq|rs) = sum C(mu,p) C(nu,q) C(kap,r) C(lam,s) (mu nu|kap lam),
 * one index at a time on the GPU (four FP64 GEMMs). tei is the nao^4 AO tensor in chemists' notation, C is
 * nao x nmo, out is nmo^4; with chem_to_phys != 0, out(p,r,q,s) = (pq|rs) = <pr|qs>. ColMajor throughout. */
int scess_transform_tei(scess_ctx *ctx, int64_t nao, int64_t nmo, const double *tei, const double *C,
                        int chem_to_phys, double *out);

#ifdef __cplusplus
}
#endif
#endif /* SCESS_GPU_H */
