// Internal declarations shared by the (T) kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace scess {

enum Prec : int { P64 = 0, P32 = 1, PB3 = 2, PI8 = 3 };
constexpr int kNumPrec = 4;
// GemmArgs::prec only: BF16x3 products of operands stored pre-split (split_rows_kernel layout: per 8 k,
// 8 hi then 8 lo bf16; rows padded to a multiple of 32 k with zeros). Such an operand is passed as
// FP64 rows of K/2 slots (the same bytes), so K2 and the strides count 8-byte slots.
constexpr int PB3S = 4;

// K-dimension chunk (elements) of each precision's W-build kernel; the K
// extents of the packed operands are padded to a multiple of it with zeros.
// PI8 pads vK and oK to 32 (the packed row chunk of the int8 W-build).
constexpr int kChunk[kNumPrec] = {16, 16, 32, 32};
constexpr int kChunkO[kNumPrec] = {16, 16, 32, 32};
// Bytes of one packed element: FP64 8, FP32 4, BF16X3 4 (hi and lo bf16); PI8 one per digit (the
// digit count is chosen per context, Operands::slices).
constexpr int kElemBytes[kNumPrec] = {8, 4, 4, 1};
// PI8 digits per element: 5, 6 or 7 (15, 21 or 28 int8 digit products per FP64 product)
constexpr int kI8MinDigits = 5, kI8MaxDigits = 7, kI8DefaultDigits = 6;

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// LDS pointer type of global_load_lds (direct-to-LDS loads)
typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt vmcnt(N) leaving the other counters alone (gfx9 encoding)
template <int N> __device__ inline void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Work-group id -> logical id so that consecutive logical ids run on the same
// XCD: the dispatcher hands work-groups to the 8 XCDs of the MI355X round-robin,
// and neighbouring work items share operand rows (one L2 per XCD).
__device__ inline int xcd_remap(int w, int n) {
    const int xcd = w & 7, idx = w >> 3;
    const int per = n >> 3, rem = n & 7;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Device-resident operands of the W contractions in one precision.
//   X [p][x][y][f]  = (px|yf)         o * v * v * vK
//   T [q][r][z][f]  = t_qr^zf         o * o * v * vK
//   Rn[p][q][x][m]  = -(px|mq)        o * o * v * oK
//   TM[r][y][z][m]  = t_mr^yz         o * v * v * oK
// FP64 and FP32 store plain rows; BF16X3 stores each row as groups of 8 k's:
// 8 hi bf16 followed by 8 lo bf16 (32 bytes per group).
// PI8 (Ozaki splitting): every element of a tensor is scaled by the tensor's
// power of two 2^-e (|x| 2^-e <= 0.498), rounded to the integer N = rint(x 2^-e 256^S)
// and written as S balanced base-256 digits d_1..d_S in [-128, 127], N = sum_t d_t
// 256^(S-t), so x = 2^e sum_t d_t 256^-t up to 2^e 256^-S / 2. Every row (vK or oK
// long) is stored as chunks of 32 k's, each chunk slice-major ([slice][32 bytes]).
struct Operands {
    void *X = nullptr, *T = nullptr, *Rn = nullptr, *TM = nullptr;
    int64_t vK = 0, oK = 0;   // padded K extents (elements)
    int64_t rbv = 0, rbo = 0; // bytes per packed row of K extent vK / oK
    int eX = 0, eT = 0, eR = 0;  // PI8: exponents of the (ia|bf), t2 and (ia|mj) scalings (eR == eX)
    int slices = 0;              // PI8: digits per element
};

// Which source tensor a pack launch reads.
enum PackMode : int { PACK_X = 0, PACK_T = 1, PACK_RN = 2, PACK_TM = 3, PACK_S = 4 };

hipError_t launch_pack(int mode, int prec, const double *src, void *dst, int o, int v, int64_t kpad,
                       hipStream_t s);
// PI8 pack: the elements of mode's tensor times `scale` (a power of two, |x| scale <= 0.498) as
// `slices` digits, in chunks of `gran` (32) k's
hipError_t launch_pack_i8(int mode, const double *src, void *dst, int o, int v, int64_t kpad, int gran,
                          double scale, int slices, hipStream_t s);
// max |x| over n doubles into *out (device; *out must be zero before the call), as a bit pattern
hipError_t launch_absmax(const double *x, int64_t n, unsigned long long *out, hipStream_t s);

struct WBuildArgs {
    const char *X, *T, *Rn, *TM;
    int o, v, vR;
    int64_t rbv, rbo;  // bytes per row (vK, oK rows)
    int nchv, ncho;    // K chunks of a vK row and of an oK row
    int nt;            // tiles per row / column dimension
    int ntile1;        // work-groups of the fixed-a family
    const int *ijk;    // 3 ints per batch entry
    void *W1, *W2;     // accumulator type, [nbatch] x vR^3 in 8x8x8 blocks
    int64_t wstride;   // elements per batch entry
    // 1: an entry with i == j builds only H = the family-1 segments {0, 1, 3, 4, 5, 6} into W1
    // (W = H(a,b,c) + H(b,a,c) for i == j) and its family-2 work-groups do nothing
    int halve;
    // PI8: W = scale * (digit products of the f- and the m-terms; the two integral blocks share one
    // power-of-two scale and T2 serves both terms)
    double scale;
    int slices;    // PI8 digits per element
};

// Tile configuration: 2x2 waves, each owning blk x blk MFMA 16x16 blocks.
int wbuild_tile(int prec, int v);  // returns blk for this precision and size
// dma: the direct-to-LDS kernel (blk 2 or 4); otherwise the register-staged one (blk 1, 2, 4, 5)
hipError_t launch_wbuild(int prec, int blk, bool dma, const WBuildArgs &a, int nbatch, hipStream_t s);

struct EpiArgs {
    const void *W1, *W2;  // accumulator type, [nbatch] x vR^3 in 8x8x8 blocks
    int64_t wstride;
    int o, v, vR, nb;     // nb = vR / 8 blocks per dimension
    const int *ijk;       // 3 ints per batch entry
    const int *orbits;    // 3 ints per orbit (A >= B >= C block indices)
    const uint64_t *orbtab;  // 6 packed geometry entries per orbit (orbit_table)
    int norb;
    const double *U;      // [nbatch][3][v][vR] V pieces (launch_vpieces), rows zero-padded
    const double *t1;     // [o][vR], rows zero-padded
    const double *e;      // [o+v]
    const double *ev;     // [vR] virtual orbital energies, zero-padded
    int ngrp;             // work-groups per ijk (runs of consecutive orbits)
    double *partials;     // [nbatch][ngrp]
    int halve;            // entries with i == j read W1 in place of W2 (see WBuildArgs::halve)
};
int epilogue_groups(int norb);
// Packed geometry of each orbit {(A,B,C) and its permutations} for the epilogue: per distinct
// block d (in the order of the permutations (ABC, BCA, CAB, CBA, ACB, BAC)) its (P,Q,R), the
// count of distinct blocks and where the five permuted reads of r3 find their Y block.
// False when a block index does not fit the table's 10-bit fields (nb > kMaxOrbitBlocks).
constexpr int kMaxOrbitBlocks = 1024;
bool orbit_table(const std::vector<int> &orbits, std::vector<uint64_t> &tab);
// slices: int8 digits of a PI8 W-build (chooses the epilogue width that runs beside it)
hipError_t launch_epilogue(int prec, const EpiArgs &a, int nbatch, hipStream_t s, int slices = 0);
// fixed indices per work-group of the int8-digit W-build with `slices` digits
int i8_fixed_per_group(int slices);
// V pieces of every batch entry from S[p][q][x][y] = (px|qy) into U[nbatch][3][v][vR]
hipError_t launch_vpieces(const double *S, const int *ijk, int o, int v, int vR, double *U, int nbatch, hipStream_t s);

// F[ijk_index[b]] = deterministic compensated sum of partials[b][0..norb)
hipError_t launch_reduce_partials(const double *partials, int norb, const int *ijk_index, double *F, int nbatch,
                                  hipStream_t s);
// out[0] = sum_t 2/deg_t F[t] over the unique occupied triples (deterministic, compensated)
hipError_t launch_reduce_energy(const double *F, const int *ijk_all, int nall, double *out, hipStream_t s);
// p[0..n) = 0 by a kernel (usable where a hipMemsetAsync node is not wanted)
hipError_t launch_fill_zero(double *p, int64_t n, hipStream_t s);

// ----------------------------------------------------------------- adjoint
// FP64 operand of the generic GEMM: element (r, k1, k2) of batch b at
// base + b*sb + r*sr + k1*s1 + k2*s2.
struct GemmOp {
    const double *base;
    int64_t sr, s1, s2, sb;
    // Elements readable along the unit-stride dimension from index 0 (the padded row length of the
    // array; 0 = unknown), and whether entries with k2 in [K2, K2 rounded up to 16) read as zero. With
    // these known, the back-contraction GEMMs use direct-to-LDS loads (gemm_kn_kernel).
    int64_t cext;
    bool kzero;
};
struct GemmSeg {
    GemmOp A, B;  // A rows m, B rows n
    int K1, K2;   // K = K1 x K2, K2 the inner (chunked) extent
};
// C[b][m][n] += alpha sum_seg sum_k A(m,k) B(n,k); splits > 1 (batch 1) sums K-range partials in order.
struct GemmArgs {
    GemmSeg seg[2];
    int nseg;
    int M, N;
    double *C;
    int64_t sCm, sCn, sCb;
    double alpha;
    int splits;
    double *scratch;  // splits * M * N
    // products in P64 (default), P32 or PB3 on the K-contiguous path; FP64 inputs are rounded (split
    // into hi + lo bf16 for PB3) in registers, each 16-deep K chunk is summed on the MFMA in FP32
    // and added to FP64 accumulators. Other layouts always use FP64.
    int prec;
    // Set by launch_gemm_group only (gemm_kn_kernel): > 0 when the batch is folded into the rows, M
    // then being mfold x batch and row m' of A and C row m' % mfold of batch entry m' / mfold.
    int mfold;
};
hipError_t launch_gemm(const GemmArgs &g, int batch, hipStream_t s);
// Up to kMaxGemmGroup GEMMs of one shape (M, N, batch, splits) writing disjoint outputs, launched
// together (direct-to-LDS kernel: grid z = split x problem; grouped by operand layout) or, where that
// kernel does not apply, one after another in the given order. Each problem's splits > 1 needs its own
// scratch. The sums are those of launch_gemm on each problem.
constexpr int kMaxGemmGroup = 6;
struct GemmGroup {
    GemmArgs g[kMaxGemmGroup];
    int n;
};
hipError_t launch_gemm_group(const GemmGroup &gg, int batch, hipStream_t s);
// G_W, G_V ([vR]^3 row-major, FP64) of the single occupied triple a.ijk; W in a.W1/W2, V pieces in a.U
hipError_t launch_adjoint_g(int prec, const EpiArgs &a, double wgt, double *GW, double *GV, hipStream_t s);
// V backward of one occupied triple into dS [o][o][v][v] and dt1 [o][vR]; dU (3 v^2), dt (3 v) and H (v^2)
// are scratch
hipError_t launch_vback(const double *GV, const double *U, const double *t1, int i, int j, int k, int o, int v, int vR,
                        double *dU, double *dt, double *H, double *dS, double *dt1, hipStream_t s);
// ColMajor gradients: mode 0 ovvv <- dX, 1 ovoo <- -dRn, 2 ovov <- dS, 3 t2 <- dT + dTM, 4 t1 <- dt1
hipError_t launch_unpack(int mode, const double *A, const double *B, double *out, int o, int v, int64_t vK, int64_t oK,
                         int64_t vR, int64_t n, hipStream_t s);

}  // namespace scess
