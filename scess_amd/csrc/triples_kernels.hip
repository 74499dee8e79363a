// gfx950 kernels of the closed-shell (T) triples correction.
//
// Algorithm (ijk-driven form of RCCSD_TriplesCorrection,
// src/Quantum/CC/RCCSD_T.cpp:17-168): for every unique occupied triple
// i >= j >= k (not all equal) the symmetrised connected triples
//   W^abc = sum over the six simultaneous permutations of the pairs (ia),
//           (jb), (kc) of  w(x,y,z; p,q,r) = sum_f (px|yf) t_rq^zf
//                                          - sum_m (px|mq) t_mr^yz
// are built for all a, b, c on MFMA, then
//   F_ijk = sum_abc W r3(W + V) / D,  r3(Y) = 4Y^abc + Y^bca + Y^cab
//           - 2Y^cba - 2Y^acb - 2Y^bac,  D = e_i+e_j+e_k-e_a-e_b-e_c
// and E(T) = sum_ijk 2/deg_ijk F_ijk (deg = 2 if two of i,j,k are equal).
//
// W-build: a term with free virtual indices (x, y, z) is a GEMM at fixed x
// or y with z as a row or column index, so no single output layout holds all
// six. Four f-terms and four m-terms share the layout "fixed a, rows b,
// cols c" (family 1, W1[a][b][c]) and are concatenated along K; the other two
// of each share "fixed b, rows a, cols c" (family 2, W2[b][a][c]). Every
// segment is an NT product C[r][c] += sum_k A[r][k] B[c][k] with K contiguous.
//
// Epilogue: W = W1[a][b][c] + W2[b][a][c] is read once per ijk in 8^3 blocks;
// a work-group owns one orbit {(A,B,C) and its permutations} of virtual
// blocks so r3 needs no other block, builds V on the fly in FP64 and writes
// one compensated FP64 partial per (ijk, orbit).
#include "triples_internal.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>

namespace scess {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ helpers
__device__ inline uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_float(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }


// Neumaier compensated accumulation
__device__ inline void nadd(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

__device__ inline void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// Deterministic fixed-tree reduction of 256 (s, c) pairs; returns s + c.
__device__ inline double block_reduce_256(double s, double c, double *red_s, double *red_c) {
    const int tid = threadIdx.x;
    red_s[tid] = s;
    red_c[tid] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            double ss, ee;
            two_sum(red_s[tid], red_s[tid + w], ss, ee);
            red_c[tid] = red_c[tid] + red_c[tid + w] + ee;
            red_s[tid] = ss;
        }
        __syncthreads();
    }
    return red_s[0] + red_c[0];
}

// -------------------------------------------------------------------- pack
// One thread per packed element (row, k); gathers from the ColMajor FP64
// input. Rows past the real K extent are zero.
// Element (row, k) of a packed operand, gathered from the ColMajor FP64 input (0 past the real K).
__device__ inline double pack_value(int mode, const double *__restrict__ src, int64_t o, int64_t v, int64_t row,
                                    int64_t k) {
    double val = 0.0;
    switch (mode) {
    case PACK_X: {  // [p][x][y] x f : (px|yf) = ovvv(p,x,y,f)
        const int64_t y = row % v, t = row / v, x = t % v, p = t / v;
        if (k < v)
            val = src[p + o * (x + v * (y + v * k))];
    } break;
    case PACK_T: {  // [q][r][z] x f : t2(q,r,z,f)
        const int64_t z = row % v, t = row / v, r = t % o, q = t / o;
        if (k < v)
            val = src[q + o * (r + o * (z + v * k))];
    } break;
    case PACK_RN: {  // [p][q][x] x m : -(px|mq) = -ovoo(p,x,m,q)
        const int64_t x = row % v, t = row / v, q = t % o, p = t / o;
        if (k < o)
            val = -src[p + o * (x + v * (k + o * q))];
    } break;
    case PACK_TM: {  // [r][y][z] x m : t_mr^yz = t2(m,r,y,z)
        const int64_t z = row % v, t = row / v, y = t % v, r = t / v;
        if (k < o)
            val = src[k + o * (r + o * (y + v * z))];
    } break;
    case PACK_S: {  // [p][q][x] x y : (px|qy) = ovov(p,x,q,y)
        const int64_t x = row % v, t = row / v, q = t % o, p = t / o;
        val = src[p + o * (x + v * (q + o * k))];
    } break;
    }
    return val;
}

template <int PREC>
__global__ void pack_kernel(int mode, const double *__restrict__ src, void *__restrict__ dst, int o_, int v_,
                            int64_t kpad, int64_t nrows) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nrows * kpad)
        return;
    const int64_t row = gid / kpad, k = gid - row * kpad;
    const double val = pack_value(mode, src, o_, v_, row, k);
    if constexpr (PREC == P64) {
        ((double *)dst)[gid] = val;
    } else if constexpr (PREC == P32) {
        ((float *)dst)[gid] = (float)val;
    } else {
        uint16_t *d = (uint16_t *)dst;
        const uint16_t hi = bf16_rne((float)val);
        const uint16_t lo = bf16_rne((float)(val - (double)bf16_to_float(hi)));
        const int64_t base = row * 2 * kpad + (k >> 3) * 16 + (k & 7);
        d[base] = hi;
        d[base + 8] = lo;
    }
}

// Ozaki splitting of one element: m = x * scale (exact, |m| <= 0.498), N = rint(m 256^S) (|N| <
// 0.4981 256^S fits S balanced base-256 digits), then from the least significant digit up
// d = ((N + 128) mod 256) - 128 in [-128, 127] and N = (N - d) / 256, all exact in int64.
__global__ void pack_i8_kernel(int mode, const double *__restrict__ src, int8_t *__restrict__ dst, int o_, int v_,
                               int64_t kpad, int gran, double scale, int S, int64_t nrows) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nrows * kpad)
        return;
    const int64_t row = gid / kpad, k = gid - row * kpad;
    const double m = pack_value(mode, src, o_, v_, row, k) * scale;
    long long N = llrint(ldexp(m, 8 * S));
    int8_t *d = dst + row * kpad * S + (k / gran) * gran * S + (k % gran);
    for (int t = S - 1; t >= 0; t--) {
        const long long dig = ((N + 128) & 255) - 128;
        d[t * gran] = (int8_t)dig;
        N = (N - dig) / 256;
    }
}

static int64_t pack_rows(int mode, int64_t o, int64_t v) {
    switch (mode) {
    case PACK_X: return o * v * v;
    case PACK_T: return o * o * v;
    case PACK_RN: return o * o * v;
    case PACK_TM: return o * v * v;
    case PACK_S: return o * o * v;
    default: return -1;
    }
}

hipError_t launch_pack_i8(int mode, const double *src, void *dst, int o, int v, int64_t kpad, int gran,
                          double scale, int slices, hipStream_t s) {
    const int64_t nrows = pack_rows(mode, o, v);
    if (nrows < 0 || mode == PACK_S || (gran != 32 && gran != 64) || kpad % gran || slices < kI8MinDigits ||
        slices > kI8MaxDigits)
        return hipErrorInvalidValue;
    const int64_t blocks = (nrows * kpad + 255) / 256;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_i8_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mode, src, (int8_t *)dst, o, v,
                       kpad, gran, scale, slices, nrows);
    return hipGetLastError();
}

// |x| of non-negative doubles orders like their bit patterns, so an integer atomic max finds the max
__global__ __launch_bounds__(256) void absmax_kernel(const double *__restrict__ x, int64_t n,
                                                     unsigned long long *out) {
    __shared__ unsigned long long red[256];
    unsigned long long m = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(x[i]));
        m = b > m ? b : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out, red[0]);
}

hipError_t launch_absmax(const double *x, int64_t n, unsigned long long *out, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n, out);
    return hipGetLastError();
}

hipError_t launch_pack(int mode, int prec, const double *src, void *dst, int o, int v, int64_t kpad,
                       hipStream_t s) {
    int64_t nrows = 0;
    switch (mode) {
    case PACK_X: nrows = (int64_t)o * v * v; break;
    case PACK_T: nrows = (int64_t)o * o * v; break;
    case PACK_RN: nrows = (int64_t)o * o * v; break;
    case PACK_TM: nrows = (int64_t)o * v * v; break;
    case PACK_S: nrows = (int64_t)o * o * v; break;
    default: return hipErrorInvalidValue;
    }
    const int64_t n = nrows * kpad;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL)
        return hipErrorInvalidValue;
    if (prec == P64)
        hipLaunchKernelGGL(pack_kernel<P64>, dim3((unsigned)blocks), dim3(256), 0, s, mode, src, dst, o, v, kpad, nrows);
    else if (prec == P32)
        hipLaunchKernelGGL(pack_kernel<P32>, dim3((unsigned)blocks), dim3(256), 0, s, mode, src, dst, o, v, kpad, nrows);
    else
        hipLaunchKernelGGL(pack_kernel<PB3>, dim3((unsigned)blocks), dim3(256), 0, s, mode, src, dst, o, v, kpad, nrows);
    return hipGetLastError();
}

// ----------------------------------------------------------------- W build
template <int PREC> struct Tr;
template <> struct Tr<P64> {
    using Acc = double;
    using AccV = d4;
    static constexpr int RB = 128;  // bytes of one K chunk of a row (16 doubles)
};
template <> struct Tr<P32> {
    using Acc = float;
    using AccV = f4;
    static constexpr int RB = 64;  // 16 floats
};
template <> struct Tr<PB3> {
    using Acc = float;
    using AccV = f4;
    static constexpr int RB = 128;  // 32 k's as 4 groups of (8 hi, 8 lo) bf16
};

struct Seg {
    const char *A, *B;
    int64_t lda, ldb;  // bytes between rows
    int nch;           // K chunks
};

// Family-1 segments of H for an occupied triple with i == j: one of each pair of permutation
// terms exchanged by (i,a) <-> (j,b), so that W(a,b,c) = H(a,b,c) + H(b,a,c).
__device__ inline int halved_seg(int s) { return s < 2 ? s : s + 1; }  // 0 1 3 4 5 6

// Segment s of family 1 (fixed a = fx, rows b, cols c) or family 2 (fixed b = fx,
// rows a, cols c) for the occupied triple (i, j, k).
__device__ inline Seg get_seg(const WBuildArgs &a, bool fam1, int s, int64_t i, int64_t j, int64_t k, int64_t fx) {
    const int64_t o = a.o, v = a.v, rbv = a.rbv, rbo = a.rbo;
    Seg g;
    if (fam1) {
        switch (s) {
        case 0:  // sum_f (ia|bf) t_kj^cf
            g = {a.X + ((i * v + fx) * v) * rbv, a.T + ((k * o + j) * v) * rbv, rbv, rbv, a.nchv};
            break;
        case 1:  // sum_f (ia|cf) t_jk^bf
            g = {a.T + ((j * o + k) * v) * rbv, a.X + ((i * v + fx) * v) * rbv, rbv, rbv, a.nchv};
            break;
        case 2:  // sum_f (jb|af) t_ki^cf
            g = {a.X + ((j * v) * v + fx) * rbv, a.T + ((k * o + i) * v) * rbv, v * rbv, rbv, a.nchv};
            break;
        case 3:  // sum_f (kc|af) t_ji^bf
            g = {a.T + ((j * o + i) * v) * rbv, a.X + ((k * v) * v + fx) * rbv, rbv, v * rbv, a.nchv};
            break;
        case 4:  // -sum_m (jb|mi) t_mk^ac
            g = {a.Rn + ((j * o + i) * v) * rbo, a.TM + ((k * v + fx) * v) * rbo, rbo, rbo, a.ncho};
            break;
        case 5:  // -sum_m (jb|mk) t_mi^ca
            g = {a.Rn + ((j * o + k) * v) * rbo, a.TM + ((i * v) * v + fx) * rbo, rbo, v * rbo, a.ncho};
            break;
        case 6:  // -sum_m (kc|mi) t_mj^ab
            g = {a.TM + ((j * v + fx) * v) * rbo, a.Rn + ((k * o + i) * v) * rbo, rbo, rbo, a.ncho};
            break;
        default:  // -sum_m (kc|mj) t_mi^ba
            g = {a.TM + ((i * v) * v + fx) * rbo, a.Rn + ((k * o + j) * v) * rbo, v * rbo, rbo, a.ncho};
            break;
        }
    } else {
        switch (s) {
        case 0:  // sum_f (jb|cf) t_ik^af
            g = {a.T + ((i * o + k) * v) * rbv, a.X + ((j * v + fx) * v) * rbv, rbv, rbv, a.nchv};
            break;
        case 1:  // sum_f (kc|bf) t_ij^af
            g = {a.T + ((i * o + j) * v) * rbv, a.X + ((k * v) * v + fx) * rbv, rbv, v * rbv, a.nchv};
            break;
        case 2:  // -sum_m (ia|mj) t_mk^bc
            g = {a.Rn + ((i * o + j) * v) * rbo, a.TM + ((k * v + fx) * v) * rbo, rbo, rbo, a.ncho};
            break;
        default:  // -sum_m (ia|mk) t_mj^cb
            g = {a.Rn + ((i * o + k) * v) * rbo, a.TM + ((j * v) * v + fx) * rbo, rbo, v * rbo, a.ncho};
            break;
        }
    }
    return g;
}

template <int PREC, int BLK>
__global__ __launch_bounds__(256) void wbuild_kernel(WBuildArgs a) {
    using AccV = typename Tr<PREC>::AccV;
    using Acc = typename Tr<PREC>::Acc;
    constexpr int RB = Tr<PREC>::RB;
    constexpr int PITCH = RB + 16;
    constexpr int VPR = RB / 16;
    constexpr int BM = 2 * BLK * 16;  // rows (and columns) of the work-group tile
    constexpr int NV = BM * VPR;      // 16-byte vectors per operand per chunk
    constexpr int NVT = (NV + 255) / 256;
    __shared__ __attribute__((aligned(16))) char lds[2][2 * BM * PITCH];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lq = lane >> 4;

    const int nt2 = a.nt * a.nt;
    const int gx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
    const int bat = lin / gx;
    int bid = lin - bat * gx;
    const bool fam1 = bid < a.ntile1;
    if (!fam1)
        bid -= a.ntile1;
    const int64_t fx = bid / nt2;
    const int tt = bid - (int)fx * nt2;
    const int tr = tt / a.nt, tc = tt - tr * a.nt;
    const int *q = a.ijk + 3 * bat;
    const int64_t i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int r0 = tr * BM, c0 = tc * BM;

    const bool deg = a.halve && i == j;
    if (deg && !fam1)
        return;  // the whole work-group: W2 of this entry is not used
    const int nseg = fam1 ? (deg ? 6 : 8) : 4;
    const int total = fam1 ? (deg ? 3 : 4) * (a.nchv + a.ncho) : 2 * (a.nchv + a.ncho);

    // which MFMA blocks of this wave hold rows / columns below v
    bool vr[BLK], vc[BLK];
#pragma unroll
    for (int b = 0; b < BLK; b++) {
        vr[b] = r0 + wm * BLK * 16 + b * 16 < v;
        vc[b] = c0 + wn * BLK * 16 + b * 16 < v;
    }

    AccV acc[BLK][BLK];
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLK; y++)
            acc[x][y] = AccV{};

    int s = 0, kc = 0;
    Seg sg = get_seg(a, fam1, 0, i, j, k, fx);
    // two register stages: the global loads of chunk c+2 are issued before chunk c is computed
    uint4 ra0[NVT], rb0[NVT], ra1[NVT], rb1[NVT];

    auto load = [&](uint4 *ra, uint4 *rb) {
#pragma unroll
        for (int u = 0; u < NVT; u++) {
            const int idx = tid + 256 * u;
            if (idx < NV) {
                const int row = idx / VPR, vec = idx - row * VPR;
                const int gr = r0 + row, gc = c0 + row;
                const int64_t off = (int64_t)kc * RB + vec * 16;
                ra[u] = gr < v ? *(const uint4 *)(sg.A + gr * sg.lda + off) : uint4{0, 0, 0, 0};
                rb[u] = gc < v ? *(const uint4 *)(sg.B + gc * sg.ldb + off) : uint4{0, 0, 0, 0};
            }
        }
        if (++kc == sg.nch) {
            kc = 0;
            if (++s < nseg)
                sg = get_seg(a, fam1, deg ? halved_seg(s) : s, i, j, k, fx);
        }
    };
    auto store = [&](int buf, const uint4 *ra, const uint4 *rb) {
        char *As = lds[buf];
        char *Bs = As + BM * PITCH;
#pragma unroll
        for (int u = 0; u < NVT; u++) {
            const int idx = tid + 256 * u;
            if (idx < NV) {
                const int row = idx / VPR, vec = idx - row * VPR;
                *(uint4 *)(As + row * PITCH + vec * 16) = ra[u];
                *(uint4 *)(Bs + row * PITCH + vec * 16) = rb[u];
            }
        }
    };
    auto compute = [&](int buf) {
        const char *As = lds[buf];
        const char *Bs = As + BM * PITCH;
        if constexpr (PREC == P64) {
            double af[BLK][4], bf[BLK][4];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                const double *pa = (const double *)(As + (wm * BLK * 16 + b * 16 + lr) * PITCH + lq * 32);
                const double *pb = (const double *)(Bs + (wn * BLK * 16 + b * 16 + lr) * PITCH + lq * 32);
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    af[b][t] = pa[t];
                    bf[b][t] = pb[t];
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int x = 0; x < BLK; x++)
                    if (vr[x])
#pragma unroll
                        for (int y = 0; y < BLK; y++)
                            if (vc[y])
                                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
        } else if constexpr (PREC == P32) {
            f4 af[BLK], bf[BLK];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                af[b] = *(const f4 *)(As + (wm * BLK * 16 + b * 16 + lr) * PITCH + lq * 16);
                bf[b] = *(const f4 *)(Bs + (wn * BLK * 16 + b * 16 + lr) * PITCH + lq * 16);
            }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int x = 0; x < BLK; x++)
                    if (vr[x])
#pragma unroll
                        for (int y = 0; y < BLK; y++)
                            if (vc[y])
                                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
        } else {
            bf8 ah[BLK], al[BLK], bh[BLK], bl[BLK];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                const char *pa = As + (wm * BLK * 16 + b * 16 + lr) * PITCH + lq * 32;
                const char *pb = Bs + (wn * BLK * 16 + b * 16 + lr) * PITCH + lq * 32;
                ah[b] = *(const bf8 *)pa;
                al[b] = *(const bf8 *)(pa + 16);
                bh[b] = *(const bf8 *)pb;
                bl[b] = *(const bf8 *)(pb + 16);
            }
#pragma unroll
            for (int x = 0; x < BLK; x++)
                if (vr[x])
#pragma unroll
                    for (int y = 0; y < BLK; y++)
                        if (vc[y]) {
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                        }
        }
    };

    load(ra0, rb0);
    store(0, ra0, rb0);
    if (total > 1)
        load(ra1, rb1);
    __syncthreads();
    for (int c = 0; c < total; c += 2) {
        if (c + 2 < total)
            load(ra0, rb0);
        compute(0);
        if (c + 1 < total)
            store(1, ra1, rb1);
        __syncthreads();
        if (c + 1 >= total)
            break;
        if (c + 3 < total)
            load(ra1, rb1);
        compute(1);
        if (c + 2 < total)
            store(0, ra0, rb0);
        __syncthreads();
    }

    // W in 8x8x8 blocks: element (fixed, row, col) at wblk(fixed, row, col)
    Acc *W = (Acc *)(fam1 ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int nbR = a.vR >> 3;
    const int64_t fbase = ((int64_t)(fx >> 3) * nbR) * nbR * 512 + (fx & 7) * 64;
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLK; y++) {
            const int col = c0 + wn * BLK * 16 + y * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = r0 + wm * BLK * 16 + x * 16 + (PREC == P64 ? lq + 4 * r : 4 * lq + r);
                if (row < v && col < v)
                    W[fbase + ((int64_t)(row >> 3) * nbR + (col >> 3)) * 512 + (row & 7) * 8 + (col & 7)] = acc[x][y][r];
            }
        }
}

// ------------------------------------------------ W build, direct-to-LDS loads
// Same tiles, segments and MFMA core as wbuild_kernel, but the operand chunks go
// from global memory straight into LDS (global_load_lds_dwordx4, 16 bytes per
// lane, one contiguous KiB per wave instruction) through a ring of STAGES
// buffers, so no registers hold staged data and no ds_write is issued. Rows are
// stored unpadded; the 16-byte vectors of row r are permuted by swz(r) so that
// the MFMA fragment reads (16 rows x 16 bytes per pass) hit distinct banks.

template <int VPR> __device__ inline int lds_swz(int r) {
    if constexpr (VPR == 8)
        return (r ^ (r >> 3)) & 7;
    else
        return ((r >> 1) ^ (r >> 3)) & 3;
}


template <int PREC, int BLK, int STAGES>
__global__ __launch_bounds__(256) void wbuild_dma_kernel(WBuildArgs a) {
    using AccV = typename Tr<PREC>::AccV;
    using Acc = typename Tr<PREC>::Acc;
    constexpr int RB = Tr<PREC>::RB;
    constexpr int VPR = RB / 16;
    constexpr int BM = 2 * BLK * 16;
    constexpr int OPB = BM * RB;      // bytes of one operand chunk
    constexpr int NIW = OPB / 4096;   // wave instructions per operand per chunk (four waves, 1 KiB each)
    static_assert(OPB % 4096 == 0, "operand chunk must split into whole wave loads");
    static_assert(STAGES >= 2 && STAGES <= 4, "stages");
    __shared__ __attribute__((aligned(1024))) char lds[STAGES][2 * OPB];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lq = lane >> 4;

    const int nt2 = a.nt * a.nt;
    const int gx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
    const int bat = lin / gx;
    int bid = lin - bat * gx;
    const bool fam1 = bid < a.ntile1;
    if (!fam1)
        bid -= a.ntile1;
    const int64_t fx = bid / nt2;
    const int tt = bid - (int)fx * nt2;
    const int tr = tt / a.nt, tc = tt - tr * a.nt;
    const int *q = a.ijk + 3 * bat;
    const int64_t i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int r0 = tr * BM, c0 = tc * BM;

    const bool deg = a.halve && i == j;
    if (deg && !fam1)
        return;  // the whole work-group: W2 of this entry is not used
    const int nseg = fam1 ? (deg ? 6 : 8) : 4;
    const int total = fam1 ? (deg ? 3 : 4) * (a.nchv + a.ncho) : 2 * (a.nchv + a.ncho);

    bool vr[BLK], vc[BLK];
#pragma unroll
    for (int b = 0; b < BLK; b++) {
        vr[b] = r0 + wm * BLK * 16 + b * 16 < v;
        vc[b] = c0 + wn * BLK * 16 + b * 16 < v;
    }

    AccV acc[BLK][BLK];
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLK; y++)
            acc[x][y] = AccV{};

    // per-lane part of the load addresses (fixed for the whole kernel): row in the tile and the
    // byte offset of the logical vector this lane fetches for LDS unit n * 64 + lane
    int lrow[NIW], lvoff[NIW];
#pragma unroll
    for (int u = 0; u < NIW; u++) {
        const int qq = (wave * NIW + u) * 64 + lane;
        const int r = qq / VPR, p = qq - r * VPR;
        lrow[u] = r;
        lvoff[u] = (p ^ lds_swz<VPR>(r)) * 16;
    }

    int s = 0, kc = 0;
    Seg sg = get_seg(a, fam1, 0, i, j, k, fx);
    // this lane's load addresses at k-chunk 0 of the current segment (rows past v load row v - 1,
    // valid memory whose outputs are never stored); a chunk adds kc * RB
    const char *pa[NIW], *pb[NIW];
    auto lane_rows = [&]() {
#pragma unroll
        for (int u = 0; u < NIW; u++) {
            const int gr = min(r0 + lrow[u], v - 1), gc = min(c0 + lrow[u], v - 1);
            pa[u] = sg.A + gr * sg.lda + lvoff[u];
            pb[u] = sg.B + gc * sg.ldb + lvoff[u];
        }
    };
    lane_rows();
    auto issue = [&](int st) {
        char *base = lds[st];
        const int64_t off = (int64_t)kc * RB;
#pragma unroll
        for (int u = 0; u < NIW; u++) {
            __builtin_amdgcn_global_load_lds((const void *)(pa[u] + off),
                                             (lds_void *)(base + (wave * NIW + u) * 1024), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(pb[u] + off),
                                             (lds_void *)(base + OPB + (wave * NIW + u) * 1024), 16, 0, 0);
        }
        if (++kc == sg.nch) {
            kc = 0;
            if (++s < nseg) {
                sg = get_seg(a, fam1, deg ? halved_seg(s) : s, i, j, k, fx);
                lane_rows();
            }
        }
    };
    // fragment address of logical 16-byte vector lv of tile row rr
    auto frag = [&](const char *op, int rr, int lv) { return op + rr * RB + ((lv ^ lds_swz<VPR>(rr)) * 16); };
    auto compute = [&](int st) {
        const char *As = lds[st];
        const char *Bs = As + OPB;
        if constexpr (PREC == P64) {
#pragma unroll
            for (int h = 0; h < 2; h++) {  // two k-pairs per 16-byte vector: half the fragment registers
                d2 af[BLK], bf[BLK];
#pragma unroll
                for (int b = 0; b < BLK; b++) {
                    af[b] = *(const d2 *)frag(As, wm * BLK * 16 + b * 16 + lr, 2 * lq + h);
                    bf[b] = *(const d2 *)frag(Bs, wn * BLK * 16 + b * 16 + lr, 2 * lq + h);
                }
#pragma unroll
                for (int t = 0; t < 2; t++)
#pragma unroll
                    for (int x = 0; x < BLK; x++)
                        if (vr[x])
#pragma unroll
                            for (int y = 0; y < BLK; y++)
                                if (vc[y])
                                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
            }
        } else if constexpr (PREC == P32) {
            f4 af[BLK], bf[BLK];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                af[b] = *(const f4 *)frag(As, wm * BLK * 16 + b * 16 + lr, lq);
                bf[b] = *(const f4 *)frag(Bs, wn * BLK * 16 + b * 16 + lr, lq);
            }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int x = 0; x < BLK; x++)
                    if (vr[x])
#pragma unroll
                        for (int y = 0; y < BLK; y++)
                            if (vc[y])
                                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
        } else {
            bf8 ah[BLK], al[BLK], bh[BLK], bl[BLK];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                const int ra = wm * BLK * 16 + b * 16 + lr, rb = wn * BLK * 16 + b * 16 + lr;
                ah[b] = *(const bf8 *)frag(As, ra, 2 * lq);
                al[b] = *(const bf8 *)frag(As, ra, 2 * lq + 1);
                bh[b] = *(const bf8 *)frag(Bs, rb, 2 * lq);
                bl[b] = *(const bf8 *)frag(Bs, rb, 2 * lq + 1);
            }
#pragma unroll
            for (int x = 0; x < BLK; x++)
                if (vr[x])
#pragma unroll
                    for (int y = 0; y < BLK; y++)
                        if (vc[y]) {
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                        }
        }
    };

#pragma unroll
    for (int st = 0; st < STAGES - 1; st++)
        if (st < total)
            issue(st);
    for (int c = 0; c < total; c++) {
        // this wave's loads of chunk c have landed once at most `ahead` later chunks are outstanding
        const int ahead = min(total - 1 - c, STAGES - 2);
        if (STAGES > 3 && ahead >= 2)
            wait_vm<(STAGES > 3 ? 2 : 0) * 2 * NIW>();
        else if (STAGES > 2 && ahead == 1)
            wait_vm<(STAGES > 2 ? 1 : 0) * 2 * NIW>();
        else
            wait_vm<0>();
        __syncthreads();  // every wave's chunk c is in LDS; stage (c - 1) % STAGES is free again
        if (c + STAGES - 1 < total)
            issue((c + STAGES - 1) % STAGES);
        compute(c % STAGES);
    }

    Acc *W = (Acc *)(fam1 ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int nbR = a.vR >> 3;
    const int64_t fbase = ((int64_t)(fx >> 3) * nbR) * nbR * 512 + (fx & 7) * 64;
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLK; y++) {
            const int col = c0 + wn * BLK * 16 + y * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = r0 + wm * BLK * 16 + x * 16 + (PREC == P64 ? lq + 4 * r : 4 * lq + r);
                if (row < v && col < v)
                    W[fbase + ((int64_t)(row >> 3) * nbR + (col >> 3)) * 512 + (row & 7) * 8 + (col & 7)] = acc[x][y][r];
            }
        }
}

// Whether the fixed index of segment s (get_seg numbering) sits on operand A (else on B).
__device__ inline bool seg_fx_on_a(bool fam1, int s) { return fam1 && (s == 0 || s == 2 || s == 6 || s == 7); }

// ------------------------------- W build, FP64, F fixed indices per work-group
// wbuild_dma_kernel<P64, 2, STAGES> with F (2 or 4) outputs per work-group: the 64 x 64 tile (tr, tc)
// of W for the fixed indices fx0 = F g .. fx0 + F - 1. In every segment exactly one operand depends on
// the fixed index (seg_fx_on_a), so a stage holds the other operand's tile once (tile 0) and the
// fixed-index operand's tile of each fixed index (tiles 1 .. F): F + 1 tiles of 8 KiB for F outputs,
// where F one-index work-groups load 2F. 2F waves; wave w computes fixed index w >> 1, columns
// 32 (w & 1) .. + 31 of the tile, all 64 rows (eight 16 x 16 blocks, 32 MFMAs per barrier against 16).
// The fragments and the order of the MFMAs into each accumulator are those of wbuild_dma_kernel, so W
// is bitwise the same. F = 4: 512 threads, 80 KiB of LDS and at most 128 registers (four waves per
// SIMD), so two work-groups fill a CU.
template <int F, int STAGES>
__global__ __launch_bounds__(64 * 2 * F) __attribute__((amdgpu_waves_per_eu(4))) void wbuild_fx_kernel(WBuildArgs a) {
    constexpr int RB = 128, VPR = 8;  // 16 doubles per row chunk, as eight 16-byte vectors
    constexpr int OPB = 64 * RB;      // bytes of one operand tile chunk (8 KiB)
    constexpr int NW = 2 * F;                        // waves
    constexpr int NLW = (F + 1) * OPB / 1024 / NW;  // 1 KiB wave loads per wave per stage
    static_assert((F + 1) * OPB / 1024 % NW == 0, "loads split evenly over the waves");
    static_assert(STAGES == 2 || STAGES == 3, "stages");
    __shared__ __attribute__((aligned(1024))) char lds[STAGES][(F + 1) * OPB];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wf = wave >> 1, wp = wave & 1;  // fixed index (0 .. F-1) and column half of this wave
    const int lr = lane & 15, lq = lane >> 4;

    const int nt2 = a.nt * a.nt;
    const int nfg = (a.v + F - 1) / F;  // fixed-index groups per family
    const int gx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
    const int bat = lin / gx;
    int bid = lin - bat * gx;
    const bool fam1 = bid < nfg * nt2;
    if (!fam1)
        bid -= nfg * nt2;
    const int fg = bid / nt2;
    const int tt = bid - fg * nt2;
    const int tr = tt / a.nt, tc = tt - tr * a.nt;
    const int *q = a.ijk + 3 * bat;
    const int64_t i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int r0 = tr * 64, c0 = tc * 64;
    const int64_t fx0 = (int64_t)F * fg;
    const int64_t fx1 = min(fx0 + 1, (int64_t)v - 1);  // gives the address step per fixed index
    // fixed indices past v load the operand rows of index v - 1; their waves compute and store nothing
    const bool live = fx0 + wf < v;

    const bool deg = a.halve && i == j;
    if (deg && !fam1)
        return;  // the whole work-group: W2 of this entry is not used
    const int nseg = fam1 ? (deg ? 6 : 8) : 4;
    const int total = fam1 ? (deg ? 3 : 4) * (a.nchv + a.ncho) : 2 * (a.nchv + a.ncho);

    bool vr[4], vc[2];
#pragma unroll
    for (int x = 0; x < 4; x++)
        vr[x] = r0 + x * 16 < v;
#pragma unroll
    for (int y = 0; y < 2; y++)
        vc[y] = c0 + wp * 32 + y * 16 < v;

    d4 acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            acc[x][y] = d4{};

    // wave load L = u * NW + wave fills LDS KiB L of a stage: tile L / 8 (a constant of u, since NW is 4
    // or 8), rows 8 (L % 8) .. + 7; the lane's row in its tile and the byte offset of the logical vector
    // it fetches (fixed per lane)
    constexpr auto tile_of = [](int u) { return u * NW / 8; };
    int lrow[NLW], lvoff[NLW];
#pragma unroll
    for (int u = 0; u < NLW; u++) {
        const int L = u * NW + wave;
        const int r = (L & 7) * 8 + (lane >> 3), p = lane & 7;
        lrow[u] = r;
        lvoff[u] = (p ^ lds_swz<VPR>(r)) * 16;
    }

    // this lane's load addresses at k-chunk 0 of segment sgi (rows past v load row v - 1, valid memory
    // whose outputs are never stored); a chunk adds kc * RB
    const char *pl[NLW];
    int nch = 0;
    bool fa = false;
    auto set_seg = [&](int sgi) {
        const Seg s0 = get_seg(a, fam1, sgi, i, j, k, fx0);
        const Seg s1 = get_seg(a, fam1, sgi, i, j, k, fx1);  // address arithmetic only: the step per index
        fa = seg_fx_on_a(fam1, sgi);
        nch = s0.nch;
        const int64_t dfx = fa ? s1.A - s0.A : s1.B - s0.B;  // bytes per fixed index (0 if fx1 == fx0)
#pragma unroll
        for (int u = 0; u < NLW; u++) {
            const int tl = tile_of(u);
            // tile 0: the operand without the fixed index; tile 1 + t: the one with it, index fx0 + t
            const bool use_a = (tl == 0) != fa;
            const int64_t t = tl == 0 ? 0 : min(fx0 + tl - 1, (int64_t)v - 1) - fx0;
            const char *base = (use_a ? s0.A : s0.B) + t * dfx;
            const int64_t ld = use_a ? s0.lda : s0.ldb;
            const int row = min((use_a ? r0 : c0) + lrow[u], v - 1);
            pl[u] = base + row * ld + lvoff[u];
        }
    };
    set_seg(0);
    int s = 0, kc = 0, famask = 0;  // famask bit st: the chunk in stage st has the fixed index on A
    auto issue = [&](int st) {
        char *base = lds[st];
        const int64_t off = (int64_t)kc * RB;
#pragma unroll
        for (int u = 0; u < NLW; u++)
            __builtin_amdgcn_global_load_lds((const void *)(pl[u] + off),
                                             (lds_void *)(base + (u * NW + wave) * 1024), 16, 0, 0);
        famask = fa ? famask | (1 << st) : famask & ~(1 << st);
        if (++kc == nch) {
            kc = 0;
            if (++s < nseg)
                set_seg(deg ? halved_seg(s) : s);
        }
    };
    auto frag = [&](const char *op, int rr, int lv) { return op + rr * RB + ((lv ^ lds_swz<VPR>(rr)) * 16); };
    auto compute = [&](int st) {
        const bool sa = (famask >> st) & 1;
        const char *T0 = lds[st], *Tf = lds[st] + (1 + wf) * OPB;
        const char *As = sa ? Tf : T0, *Bs = sa ? T0 : Tf;
#pragma unroll
        for (int h = 0; h < 2; h++) {  // two k-pairs per 16-byte vector
            d2 af[4], bf[2];
#pragma unroll
            for (int x = 0; x < 4; x++)
                af[x] = *(const d2 *)frag(As, x * 16 + lr, 2 * lq + h);
#pragma unroll
            for (int y = 0; y < 2; y++)
                bf[y] = *(const d2 *)frag(Bs, wp * 32 + y * 16 + lr, 2 * lq + h);
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int x = 0; x < 4; x++)
                    if (vr[x])
#pragma unroll
                        for (int y = 0; y < 2; y++)
                            if (vc[y])
                                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
        }
    };

#pragma unroll
    for (int st = 0; st < STAGES - 1; st++)
        if (st < total)
            issue(st);
    for (int c = 0; c < total; c++) {
        // this wave's loads of chunk c have landed once at most the next chunk's are outstanding
        if (STAGES > 2 && c + 1 < total)
            wait_vm<(STAGES > 2 ? NLW : 0)>();
        else
            wait_vm<0>();
        __syncthreads();  // every wave's chunk c is in LDS; stage (c - 1) % STAGES is free again
        if (c + STAGES - 1 < total)
            issue((c + STAGES - 1) % STAGES);
        if (live)
            compute(c % STAGES);
    }
    if (!live)
        return;

    double *W = (double *)(fam1 ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int nbR = a.vR >> 3;
    const int64_t fx = fx0 + wf;
    const int64_t fbase = ((int64_t)(fx >> 3) * nbR) * nbR * 512 + (fx & 7) * 64;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            const int col = c0 + wp * 32 + y * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = r0 + x * 16 + lq + 4 * r;
                if (row < v && col < v)
                    W[fbase + ((int64_t)(row >> 3) * nbR + (col >> 3)) * 512 + (row & 7) * 8 + (col & 7)] = acc[x][y][r];
            }
        }
}

// ------------------------------------------------------ W build, int8 digits
// Exact-product mode (PI8). Every operand element is S signed 8-bit digits (pack_i8_kernel); the
// product of two elements is sum_{t,u} d_t d'_u 256^-(t+u) times the two tensor scales. The leading
// pairs t + u <= S + 1 (S (S + 1) / 2 of them) are summed exactly on v_mfma_i32_32x32x32_i8 into one
// int32 accumulator per g = t + u over the whole K of the f- and the m-terms (the two integral blocks
// share one power-of-two scale, and T2 serves both terms), and the groups are converted to FP64 once
// at the end. The digits past S and the dropped pairs bound the product error by about
// (S + 1) 256^-S of max|A| max|B|.
//
// Tiles of 64 x 64 outputs of one triple and segments as the floating-point W-build; the K loop runs
// in chunks of 32 k's. Every row of every operand is packed in 32-k chunks ([slice][32 bytes]); the
// m-terms (K = o) are whole chunks. A 32-k row chunk is 2 S sixteen-byte vectors (digit t, half h at
// vector 2 t + h); vector v of tile row r sits at LDS slot r * 2S + (v ^ i8_swz(r)), which puts the
// rows of every 8- and 16-lane fragment read on distinct banks for S = 5, 6, 7.
//
// A work-group serves F fixed indices fx0 .. fx0 + F - 1. In every segment exactly one operand depends
// on the fixed index (family 1: A in segments 0, 2, 6, 7 and B in 1, 3, 4, 5; family 2: B; see
// get_seg), so a stage holds the other operand's 64-row tile once (tile 0, shared) and one 64-row tile
// of the fixed-index operand per fx (tiles 1 .. F): F + 1 tiles for F outputs. F = 1: four waves, each
// one 32 x 32 block (S int32 accumulators per lane; 139-189 registers, so the 256-thread epilogue fits
// beside it); F = 2 (four waves) and F = 4 (eight waves, 512 threads): two waves per fixed index, each
// owning 64 x 32 of its output (2 S accumulators), so that per chunk a wave reads 3 S fragments for
// 2 S (S + 1) / 2 MFMAs instead of 2 S for S (S + 1) / 2.
typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));

template <int S> __device__ inline int i8_swz(int r) {
    if constexpr (S == 6)
        return ((r >> 1) ^ (r >> 2)) & 3;
    else
        return ((r >> 2) ^ (r >> 3)) & 1;
}

template <int F> constexpr int i8f_waves() { return F == 4 ? 8 : 4; }

template <int S, int F, int NST>
__global__ __launch_bounds__(F == 4 ? 512 : 256, (F == 1 || (F == 2 && S == 5)) ? 2 : 1) void wbuild_i8f_kernel(WBuildArgs a) {
    static_assert(F == 1 || F == 2 || F == 4, "fixed indices per work-group");
    static_assert(NST == 2 || NST == 3, "stages");
    constexpr int NW = i8f_waves<F>();
    constexpr int NR = F == 1 ? 1 : 2, NC = 1;               // 32 x 32 blocks per wave: rows x columns
    constexpr int WPF = NW / F;                              // waves per fixed index
    constexpr int NV = 2 * S;                                // 16-byte vectors of a 32-k row chunk
    constexpr int RB = NV * 16;                              // bytes of a row chunk
    constexpr int OPB = 64 * RB;                             // bytes of a 64-row tile chunk (NV KiB)
    constexpr int NT = F + 1;                                // tiles per stage
    constexpr int NL = NT * NV;                              // 1 KiB wave loads per stage
    constexpr int NLW = (NL + NW - 1) / NW;                  // the most any wave issues
    __shared__ __attribute__((aligned(1024))) char lds[NST][NT * OPB];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wf = wave / WPF, wp = wave - wf * WPF;  // this wave's fixed index (0 .. F-1), position
    const int rb0 = F == 1 ? wp >> 1 : 0;
    const int cb0 = F == 1 ? wp & 1 : wp;

    const int nt2 = a.nt * a.nt;
    const int nfg = (a.v + F - 1) / F;  // fixed-index groups per family
    const int gx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
    const int bat = lin / gx;
    int bid = lin - bat * gx;
    const bool fam1 = bid < nfg * nt2;
    if (!fam1)
        bid -= nfg * nt2;
    const int fg = bid / nt2;
    const int tt = bid - fg * nt2;
    const int tr = tt / a.nt, tc = tt - tr * a.nt;
    const int *q = a.ijk + 3 * bat;
    const int64_t i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int r0 = tr * 64, c0 = tc * 64;
    const int64_t fx0 = (int64_t)fg * F;

    const bool deg = a.halve && i == j;
    if (deg && !fam1)
        return;  // the whole work-group: W2 of this entry is not used
    const int nfs = fam1 ? (deg ? 3 : 4) : 2;  // f segments (and as many m segments)
    const int total = nfs * (a.nchv + a.ncho);
    // segment (get_seg numbering) at position p of the chunk sequence: the f segments, then the m ones
    auto seg_at = [&](int p) { return p < nfs ? (fam1 && deg ? halved_seg(p) : p) : (fam1 ? 4 : 2) + (p - nfs); };

    i16v acc[NR][NC][S];
#pragma unroll
    for (int x = 0; x < NR; x++)
#pragma unroll
        for (int y = 0; y < NC; y++)
#pragma unroll
            for (int g = 0; g < S; g++)
                acc[x][y][g] = i16v{};

    // load units: wave load L = wave + NW n of a stage fills LDS bytes [L KiB, L KiB + 1 KiB): tile L / NV,
    // slot (L % NV) * 64 + lane = row * NV + physical vector. Each unit's source at chunk 0 of the current
    // segment is formed when the segment changes; a chunk adds kc * RB.
    const char *ptr[NLW];
    int cur_seg = -1;
    auto set_ptrs = [&](int sgi) {
        const Seg s0 = get_seg(a, fam1, sgi, i, j, k, fx0);
        const Seg s1 = get_seg(a, fam1, sgi, i, j, k, fx0 + 1);  // address arithmetic only: the step per fx
        const bool fa = seg_fx_on_a(fam1, sgi);
        const char *shb = fa ? s0.B : s0.A;
        const int64_t shld = fa ? s0.ldb : s0.lda;
        const int shr0 = fa ? c0 : r0;
        const char *fxb = fa ? s0.A : s0.B;
        const int64_t fxld = fa ? s0.lda : s0.ldb;
        const int fxr0 = fa ? r0 : c0;
        const int64_t dfx = fa ? s1.A - s0.A : s1.B - s0.B;
#pragma unroll
        for (int n = 0; n < NLW; n++) {
            const int L = wave + NW * n;
            if (L >= NL)
                continue;
            const int t = L / NV;
            const int slot = (L - t * NV) * 64 + lane;
            const int r = slot / NV, vv = slot - r * NV;
            const int lo = (vv ^ i8_swz<S>(r)) * 16;
            if (t == 0) {
                ptr[n] = shb + (int64_t)min(shr0 + r, v - 1) * shld + lo;
            } else {  // fixed indices past v load row fx = v - 1 (valid memory, outputs never stored)
                const int64_t f = min(fx0 + t - 1, (int64_t)v - 1) - fx0;
                ptr[n] = fxb + f * dfx + (int64_t)min(fxr0 + r, v - 1) * fxld + lo;
            }
        }
    };
    // issue cursor (segment position sp, k-chunk kc) of the next chunk to load; bit st of famask: the
    // chunk in stage st has its fixed-index operand on the A side (no divisions in the chunk loop)
    int sp = 0, kc = 0;
    unsigned famask = 0;
    auto issue = [&](int st) {
        const int sgi = seg_at(sp);
        if (sgi != cur_seg) {
            cur_seg = sgi;
            set_ptrs(sgi);
        }
        famask = seg_fx_on_a(fam1, sgi) ? famask | (1u << st) : famask & ~(1u << st);
        const int64_t off = (int64_t)kc * RB;
        char *base = lds[st];
#pragma unroll
        for (int n = 0; n < NLW; n++)
            if (wave + NW * n < NL)
                __builtin_amdgcn_global_load_lds((const void *)(ptr[n] + off),
                                                 (lds_void *)(base + (wave + NW * n) * 1024), 16, 0, 0);
        if (++kc == (sp < nfs ? a.nchv : a.ncho)) {
            kc = 0;
            ++sp;
        }
    };
    // per-lane byte offset of a fragment row in a tile: digit t of row r is logical vector 2 t + hh at
    // physical vector (2 t + hh) ^ swz(r) = 4 (t >> 1) + ((2 (t & 1) + hh) ^ swz(r)) (swz < 4), so it sits at
    // base[t & 1] + 64 (t >> 1) with two bases per row
    const int rr = lane & 31, hh = lane >> 5;
    int aof[NR][2], bof[NC][2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
        for (int x = 0; x < NR; x++) {
            const int row = (rb0 + x) * 32 + rr;
            aof[x][p] = row * RB + ((2 * p + hh) ^ i8_swz<S>(row)) * 16;
        }
#pragma unroll
        for (int y = 0; y < NC; y++) {
            const int row = (cb0 + y) * 32 + rr;
            bof[y][p] = row * RB + ((2 * p + hh) ^ i8_swz<S>(row)) * 16;
        }
    }
    const bool live = fx0 + wf < v;  // a wave whose fixed index is past v computes and stores nothing
    // the S digits of the wave's NC B fragments stay in registers; the A digits are loaded one at a time
    auto compute = [&](int st, bool fa) {
        if (!live)
            return;
        const char *As = lds[st] + (fa ? 1 + wf : 0) * OPB;
        const char *Bs = lds[st] + (fa ? 0 : 1 + wf) * OPB;
        i4v bf[NC][S];
#pragma unroll
        for (int y = 0; y < NC; y++)
#pragma unroll
            for (int u = 0; u < S; u++)
                bf[y][u] = *(const i4v *)(Bs + bof[y][u & 1] + 64 * (u >> 1));
#pragma unroll
        for (int t = 1; t <= S; t++) {
            i4v af[NR];
#pragma unroll
            for (int x = 0; x < NR; x++)
                af[x] = *(const i4v *)(As + aof[x][(t - 1) & 1] + 64 * ((t - 1) >> 1));
#pragma unroll
            for (int u = 1; t + u <= S + 1; u++)
#pragma unroll
                for (int x = 0; x < NR; x++)
#pragma unroll
                    for (int y = 0; y < NC; y++)
                        acc[x][y][t + u - 2] =
                            __builtin_amdgcn_mfma_i32_32x32x32_i8(af[x], bf[y][u - 1], acc[x][y][t + u - 2], 0, 0, 0);
        }
    };

    const int nlw = (NL - wave + NW - 1) / NW;  // this wave's loads per stage: NLW or NLW - 1
#pragma unroll
    for (int st = 0; st < NST - 1; st++)
        if (st < total)
            issue(st);
    int cst = 0;  // stage of chunk c
    for (int c = 0; c < total; c++) {
        // this wave's loads of chunk c have landed once at most the next chunk's are outstanding
        if (NST > 2 && c + 1 < total) {
            if (nlw == NLW)
                wait_vm<(NST > 2 ? NLW : 0)>();
            else
                wait_vm<(NST > 2 ? NLW - 1 : 0)>();
        } else {
            wait_vm<0>();
        }
        __syncthreads();  // every wave's chunk c is in LDS; stage (c - 1) % NST is free again
        if (c + NST - 1 < total)
            issue(cst == 0 ? NST - 1 : cst - 1);  // stage (c + NST - 1) % NST
        compute(cst, (famask >> cst) & 1);
        cst = cst == NST - 1 ? 0 : cst + 1;
    }
    if (!live)
        return;

    // W = scale * sum_g acc_g 256^-g; register e of lane l of a 32 x 32 block holds row
    // (e & 3) + 8 (e >> 2) + 4 (l >> 5), column l & 31
    double *W = (double *)(fam1 ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int nbR = a.vR >> 3;
    const int64_t fx = fx0 + wf;
    const int64_t fbase = ((int64_t)(fx >> 3) * nbR) * nbR * 512 + (fx & 7) * 64;
#pragma unroll
    for (int x = 0; x < NR; x++)
#pragma unroll
        for (int y = 0; y < NC; y++) {
            const int col = c0 + (cb0 + y) * 32 + rr;
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = r0 + (rb0 + x) * 32 + (e & 3) + 8 * (e >> 2) + 4 * hh;
                double sum = 0.0;
#pragma unroll
                for (int g = S + 1; g >= 2; g--)
                    sum += (double)acc[x][y][g - 2][e] * __builtin_ldexp(1.0, -8 * g);
                if (row < v && col < v)
                    W[fbase + ((int64_t)(row >> 3) * nbR + (col >> 3)) * 512 + (row & 7) * 8 + (col & 7)] =
                        a.scale * sum;
            }
        }
}

// Fixed indices per work-group of the int8-digit W-build, from timings at o=30, v=300 (W-build s per
// step, F = 1 / 2 / 4): S = 5 5.49 / 5.25 / 5.48, S = 6 6.42 / 12.4 / 6.97, S = 7 8.62 / 14.2 / -.
int i8_fixed_per_group(int slices) { return slices == 5 ? 2 : 1; }

// MFMA blocks per wave edge, from timing sweeps at v = 300: 2 (64x64 work-group tiles) for FP64
// and FP32; 4 for BF16x3, whose three MFMAs per product need the larger register tile to keep
// LDS reads per MFMA low.
int wbuild_tile(int prec, int v) {
    if (prec == PI8)
        return 2;
    if (prec == PB3 && v >= 240)
        return 4;
    if (v > 32)
        return 2;
    return 1;
}

template <int PREC>
static hipError_t launch_wbuild_p(int blk, bool dma, const WBuildArgs &a, int nbatch, hipStream_t s) {
    const dim3 grid((unsigned)(2 * a.ntile1), (unsigned)nbatch);
    if constexpr (PREC == P64) {
        // FP64 64 x 64 tiles: four fixed indices per work-group (SCESS_WBUILD_F2=0: one, tuning override)
        const char *e = std::getenv("SCESS_WBUILD_F2");
        if (dma && blk == 2 && !(e && std::atoi(e) == 0)) {
            const dim3 g2((unsigned)(2 * ((a.v + 1) / 2) * a.nt * a.nt), (unsigned)nbatch);
            const char *es = std::getenv("SCESS_WBUILD_STAGES");  // tuning override: 3 (two indices)
            const char *ef = std::getenv("SCESS_WBUILD_FX");      // tuning override: 2 fixed indices
            if ((ef && std::atoi(ef) == 2) || (es && std::atoi(es) == 3)) {
                if (es && std::atoi(es) == 3)
                    hipLaunchKernelGGL((wbuild_fx_kernel<2, 3>), g2, dim3(256), 0, s, a);
                else
                    hipLaunchKernelGGL((wbuild_fx_kernel<2, 2>), g2, dim3(256), 0, s, a);
            } else {
                const dim3 g4((unsigned)(2 * ((a.v + 3) / 4) * a.nt * a.nt), (unsigned)nbatch);
                hipLaunchKernelGGL((wbuild_fx_kernel<4, 2>), g4, dim3(512), 0, s, a);
            }
            return hipGetLastError();
        }
    }
    if (dma && (blk == 2 || blk == 4 || (blk == 5 && PREC != P32))) {
        // two stages: more resident work-groups per CU beat deeper prefetch (sweeps at v = 300)
        int st = 2;
        if (const char *e = std::getenv("SCESS_WBUILD_STAGES"))  // tuning override: 2, 3 or 4 (3 at most for blk 5)
            st = std::atoi(e) >= 2 && std::atoi(e) <= (blk == 5 ? 3 : 4) ? std::atoi(e) : st;
#define SCESS_DMA_CASE(B, S)                                                                     \
    if (blk == B && st == S) {                                                                   \
        hipLaunchKernelGGL((wbuild_dma_kernel<PREC, B, S>), grid, dim3(256), 0, s, a);           \
        return hipGetLastError();                                                                \
    }
        SCESS_DMA_CASE(2, 2) SCESS_DMA_CASE(2, 3) SCESS_DMA_CASE(2, 4)
        SCESS_DMA_CASE(4, 2) SCESS_DMA_CASE(4, 3) SCESS_DMA_CASE(4, 4)
        if constexpr (PREC != P32) {  // 160-row FP32 chunks do not split into whole wave loads
            SCESS_DMA_CASE(5, 2) SCESS_DMA_CASE(5, 3)
        }
#undef SCESS_DMA_CASE
        return hipErrorInvalidValue;
    }
    switch (blk) {
    case 1: hipLaunchKernelGGL((wbuild_kernel<PREC, 1>), grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((wbuild_kernel<PREC, 2>), grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((wbuild_kernel<PREC, 4>), grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL((wbuild_kernel<PREC, 5>), grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_wbuild(int prec, int blk, bool dma, const WBuildArgs &a, int nbatch, hipStream_t s) {
    if (prec == PI8) {  // 64 x 64 tiles of F fixed indices per work-group
        int nst = 2, f = i8_fixed_per_group(a.slices);
        if (const char *e = std::getenv("SCESS_I8_STAGES"))  // tuning override: 2 or 3
            nst = std::atoi(e) == 3 ? 3 : 2;
        if (const char *e = std::getenv("SCESS_I8_FX"))  // tuning override: 1, 2 or 4
            f = std::atoi(e);
        // overrides without an instantiation below fall back to the defaults
        const bool built = (f == 1 || f == 2 || (f == 4 && a.slices <= 6)) &&
                           (nst == 2 || (nst == 3 && f == 4 && a.slices == 5));
        if (!built) {
            f = i8_fixed_per_group(a.slices);
            nst = 2;
        }
        const int nfg = (a.v + f - 1) / f;
        const dim3 grid((unsigned)(2 * nfg * a.nt * a.nt), (unsigned)nbatch);
#define SCESS_I8F_CASE(S_, F_, N_)                                                               \
    if (a.slices == S_ && f == F_ && nst == N_) {                                                \
        hipLaunchKernelGGL((wbuild_i8f_kernel<S_, F_, N_>), grid, dim3(64 * i8f_waves<F_>()), 0, s, a); \
        return hipGetLastError();                                                                \
    }
        SCESS_I8F_CASE(5, 1, 2) SCESS_I8F_CASE(5, 2, 2) SCESS_I8F_CASE(5, 4, 2) SCESS_I8F_CASE(5, 4, 3)
        SCESS_I8F_CASE(6, 1, 2) SCESS_I8F_CASE(6, 2, 2) SCESS_I8F_CASE(6, 4, 2)
        SCESS_I8F_CASE(7, 1, 2) SCESS_I8F_CASE(7, 2, 2)
#undef SCESS_I8F_CASE
        return hipErrorInvalidValue;
    }
    if (prec == P64)
        return launch_wbuild_p<P64>(blk, dma, a, nbatch, s);
    if (prec == P32)
        return launch_wbuild_p<P32>(blk, dma, a, nbatch, s);
    return launch_wbuild_p<PB3>(blk, dma, a, nbatch, s);
}

// ---------------------------------------------------------------- epilogue
// 1/D to full FP64 accuracy: hardware reciprocal and two Newton steps
__device__ inline double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

constexpr int EPI_THREADS = 384;  // one thread per (block, a', b') of an orbit's six 8^3 blocks

// U[b][0] = 1/2 (S_ij + S_ji^T), U[b][1] = 1/2 (S_ik + S_ki^T), U[b][2] = 1/2 (S_jk + S_kj^T)
// (v x vR each, rows zero-padded to vR) for every batch entry b, so that the six GetV terms of Symmetrise read
// V(x,y,z) = U0[x][y] t_k^z + U1[x][z] t_j^y + U2[y][z] t_i^x.
__global__ __launch_bounds__(256) void vpieces_kernel(const double *__restrict__ S, const int *__restrict__ ijk, int o,
                                                       int v, int vR, double *__restrict__ U) {
    const int64_t vv = (int64_t)v * v, vr = (int64_t)v * vR;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= 3 * vr)
        return;
    const int b = blockIdx.y;
    const int *q = ijk + 3 * b;
    const int w = (int)(gid / vr);
    const int64_t xy = gid - w * vr;
    const int x = (int)(xy / vR), y = (int)(xy - (int64_t)x * vR);
    if (y >= v) {  // zero padding of each row to vR
        U[(int64_t)b * 3 * vr + gid] = 0.0;
        return;
    }
    int p1, p2;
    if (w == 0) {
        p1 = q[0];
        p2 = q[1];
    } else if (w == 1) {
        p1 = q[0];
        p2 = q[2];
    } else {
        p1 = q[1];
        p2 = q[2];
    }
    U[(int64_t)b * 3 * vr + gid] =
        0.5 * (S[((int64_t)p1 * o + p2) * vv + (int64_t)x * v + y] + S[((int64_t)p2 * o + p1) * vv + (int64_t)y * v + x]);
}

hipError_t launch_vpieces(const double *S, const int *ijk, int o, int v, int vR, double *U, int nbatch, hipStream_t s) {
    const int64_t n = 3 * (int64_t)v * vR;
    hipLaunchKernelGGL(vpieces_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nbatch), dim3(256), 0, s, S, ijk,
                       o, v, vR, U);
    return hipGetLastError();
}

// Deterministic (s, c) reduction within a wave: fixed xor-butterfly of TwoSum steps.
__device__ inline void wave_reduce(double &s, double &c) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double so = __shfl_xor(s, m, 64), co = __shfl_xor(c, m, 64);
        double ss, ee;
        // order the operands by lane so both partners compute the same bits
        const bool lo = (threadIdx.x & m) == 0;
        two_sum(lo ? s : so, lo ? so : s, ss, ee);
        c = (lo ? c + co : co + c) + ee;
        s = ss;
    }
}

// Geometry of one orbit {(A,B,C) and its permutations} of 8-wide virtual blocks
// as seen by one thread owning row (d, a', b') of the distinct block d.
struct OrbitRow {
    bool own;       // d < number of distinct blocks
    int ga, gb, gc0;  // global a, b and first c of the row
    int nd;
    int off[5];     // Yl offsets (block, x, y) of the five permuted reads at c' = 0
    int step[5];    // Yl offset increment per c'
};

// The geometry from the host-built table (orbit_table: one 64-bit entry per orbit and distinct block d):
// P, Q, R in bits 0-29 (10 bits each), the number of distinct blocks in 30-32, and the distinct-block
// indices of the five permuted reads of r3 in 33-47 (3 bits each): Y(lb, lc, la), Y(lc, la, lb),
// Y(lc, lb, la), Y(la, lc, lb), Y(lb, la, lc) with lc = (sR, c').
template <int PB, int PX, int PY>
__device__ inline OrbitRow orbit_row_tab(const uint64_t *tab, int orb, int tid) {
    constexpr int EB = 8;
    OrbitRow r;
    const int d = tid >> 6, ap = (tid >> 3) & 7, bp = tid & 7;
    const uint64_t e = tab[(int64_t)orb * 6 + d];
    const int P = (int)(e & 1023u), Q = (int)((e >> 10) & 1023u), R = (int)((e >> 20) & 1023u);
    r.nd = (int)((e >> 30) & 7u);
    r.own = d < r.nd;
    r.ga = P * EB + ap;
    r.gb = Q * EB + bp;
    r.gc0 = R * EB;
    auto di = [&](int n) { return (int)((e >> (33 + 3 * n)) & 7u); };
    r.off[0] = di(0) * PB + bp * PX + ap;
    r.step[0] = PY;
    r.off[1] = di(1) * PB + ap * PY + bp;
    r.step[1] = PX;
    r.off[2] = di(2) * PB + bp * PY + ap;
    r.step[2] = PX;
    r.off[3] = di(3) * PB + ap * PX + bp;
    r.step[3] = PY;
    r.off[4] = di(4) * PB + bp * PX + ap * PY;
    r.step[4] = 1;
    return r;
}

// One work-group per (ijk, run of EPI_RUN consecutive orbits).
constexpr int EPI_RUN = 8;

template <typename AccT>
__global__ __launch_bounds__(EPI_THREADS) void epilogue_kernel(EpiArgs a) {
    constexpr int EB = 8;
    // Y of the orbit's distinct blocks, padded (pitches 9 and 73) against bank conflicts
    constexpr int PY = 9, PX = EB * PY + 1, PB = EB * PX;
    __shared__ double Yl[6 * PB];
    __shared__ double red[2 * (EPI_THREADS / 64)];

    const int tid = threadIdx.x;
    const int nx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * nx + blockIdx.x, nx * gridDim.y);
    const int grp = lin % nx, bat = lin / nx;
    const int *q = a.ijk + 3 * bat;
    const int i = q[0], j = q[1], k = q[2];
    const int o = a.o, v = a.v;
    const int vR = a.vR, nbR = a.vR >> 3;
    const int64_t vr = (int64_t)v * vR;
    const double eocc = a.e[i] + a.e[j] + a.e[k];
    const double *U = a.U + (int64_t)bat * 3 * vr;
    const AccT *W1b = (const AccT *)a.W1 + (int64_t)bat * a.wstride;
    const AccT *W2b = (const AccT *)(a.halve && i == j ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int ap = (tid >> 3) & 7;
    const int o0 = grp * EPI_RUN, o1 = min(o0 + EPI_RUN, a.norb);

    // register stage: W (FP64), Y, virtual orbital energies of the row's 8 c'. Every row
    // array is padded to vR, so the 8-wide rows load as whole vectors; columns past v are
    // masked where they are used.
    double Wn[EB], Yn[EB], En[EB];
    OrbitRow rn;
    auto fetch = [&](int orb) {
        rn = orbit_row_tab<PB, PX, PY>(a.orbtab, orb, tid);
        if (rn.own && rn.ga < v && rn.gb < v) {
            // W1 block (P,Q,R) row (a', b') and W2 block (Q,P,R) row (b', a'), 8 contiguous c'
            const int P = rn.ga >> 3, Q = rn.gb >> 3, R = rn.gc0 >> 3;
            const AccT *W1 = W1b + (((int64_t)P * nbR + Q) * nbR + R) * 512 + (rn.ga & 7) * 64 + (rn.gb & 7) * 8;
            const AccT *W2 = W2b + (((int64_t)Q * nbR + P) * nbR + R) * 512 + (rn.gb & 7) * 64 + (rn.ga & 7) * 8;
            const double u0 = U[(int64_t)rn.ga * vR + rn.gb];
            const double *u1 = U + vr + (int64_t)rn.ga * vR + rn.gc0;
            const double *u2 = U + 2 * vr + (int64_t)rn.gb * vR + rn.gc0;
            const double *tk = a.t1 + (int64_t)k * vR + rn.gc0;
            const double *evp = a.ev + rn.gc0;
            const double ti = a.t1[(int64_t)i * vR + rn.ga], tj = a.t1[(int64_t)j * vR + rn.gb];
            const double eab = eocc - a.ev[rn.ga] - a.ev[rn.gb];
            AccT w1[EB], w2[EB];
            double x1[EB], x2[EB], xk[EB], xe[EB];
#pragma unroll
            for (int cp = 0; cp < EB; cp++) {
                w1[cp] = W1[cp];
                w2[cp] = W2[cp];
                x1[cp] = u1[cp];
                x2[cp] = u2[cp];
                xk[cp] = tk[cp];
                xe[cp] = evp[cp];
            }
#pragma unroll
            for (int cp = 0; cp < EB; cp++) {
                const double w = (double)w1[cp] + (double)w2[cp];
                Wn[cp] = w;
                Yn[cp] = w + u0 * xk[cp] + x1[cp] * tj + x2[cp] * ti;
                En[cp] = eab - xe[cp];
            }
        } else {
#pragma unroll
            for (int cp = 0; cp < EB; cp++) {
                Wn[cp] = Yn[cp] = 0.0;
                En[cp] = 1.0;
            }
        }
    };

    double s = 0.0, c = 0.0;
    for (int orb = o0; orb < o1; orb++) {
        fetch(orb);
        const OrbitRow &r = rn;
        const int d = tid >> 6, bp = tid & 7;
        if (r.own)
#pragma unroll
            for (int cp = 0; cp < EB; cp++)
                Yl[d * PB + ap * PX + bp * PY + cp] = Yn[cp];
        __syncthreads();
        if (r.own && r.ga < v && r.gb < v) {
            // the row's (at most 8) terms in plain FP64, the rows compensated (Neumaier)
            double rs = 0.0;
#pragma unroll
            for (int cp = 0; cp < EB; cp++) {
                if (r.gc0 + cp < v) {  // Yl of masked columns is finite: padding is zero
                    const double r3 = 4.0 * Yn[cp] + Yl[r.off[0] + cp * r.step[0]] + Yl[r.off[1] + cp * r.step[1]] -
                                      2.0 * (Yl[r.off[2] + cp * r.step[2]] + Yl[r.off[3] + cp * r.step[3]] +
                                             Yl[r.off[4] + cp * r.step[4]]);
                    rs += Wn[cp] * r3 * recip(En[cp]);
                }
            }
            nadd(s, c, rs);
        }
        __syncthreads();
    }
    // fixed-order reduction: xor butterfly in each wave, then the waves in order
    wave_reduce(s, c);
    const int wv = tid >> 6;
    if ((tid & 63) == 0) {
        red[2 * wv] = s;
        red[2 * wv + 1] = c;
    }
    __syncthreads();
    if (tid == 0) {
        double ts = red[0], tc = red[1];
        for (int w = 1; w < EPI_THREADS / 64; w++) {
            double ss, ee;
            two_sum(ts, red[2 * w], ss, ee);
            tc = tc + red[2 * w + 1] + ee;
            ts = ss;
        }
        a.partials[(int64_t)bat * a.ngrp + grp] = ts + tc;
    }
}

// The same epilogue on 256 threads (4 waves), so that its work-group fits the register, LDS and wave
// slot of a W-build work-group: row-task rt (distinct block rt >> 6, a' = (rt >> 3) & 7, b' = rt & 7)
// of an orbit's 384 is taken by thread rt and, for rt >= 256, by thread rt - 256. Each thread keeps W/D
// of its rows in registers; Y goes through LDS (its own row included).
constexpr int EPI2_THREADS = 256;

template <typename AccT>
__device__ inline void epilogue256_body(const EpiArgs &a, int grp, int bat, double *Yl, double *red) {
    constexpr int EB = 8;
    constexpr int PY = 9, PX = EB * PY + 1, PB = EB * PX;
    const int tid = threadIdx.x;
    const int *q = a.ijk + 3 * bat;
    const int i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int vR = a.vR, nbR = a.vR >> 3;
    const int64_t vr = (int64_t)v * vR;
    const double eocc = a.e[i] + a.e[j] + a.e[k];
    const double *U = a.U + (int64_t)bat * 3 * vr;
    const AccT *W1b = (const AccT *)a.W1 + (int64_t)bat * a.wstride;
    const AccT *W2b = (const AccT *)(a.halve && i == j ? a.W1 : a.W2) + (int64_t)bat * a.wstride;
    const int o0 = grp * EPI_RUN, o1 = min(o0 + EPI_RUN, a.norb);
    const int nrow = tid < 384 - EPI2_THREADS ? 2 : 1;

    double s = 0.0, c = 0.0;
    for (int orb = o0; orb < o1; orb++) {
        double WD[2][EB];
        bool ok[2] = {false, false};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (h >= nrow)
                break;
            const int rt = tid + EPI2_THREADS * h;
            const OrbitRow rn = orbit_row_tab<PB, PX, PY>(a.orbtab, orb, rt);
            if (!rn.own)
                continue;
            const int d = rt >> 6, ap = (rt >> 3) & 7, bp = rt & 7;
            double Y[EB];
            if (rn.ga < v && rn.gb < v) {
                // W1 block (P,Q,R) row (a', b') and W2 block (Q,P,R) row (b', a'), 8 contiguous c'
                const int P = rn.ga >> 3, Q = rn.gb >> 3, R = rn.gc0 >> 3;
                const AccT *W1 = W1b + (((int64_t)P * nbR + Q) * nbR + R) * 512 + (rn.ga & 7) * 64 + (rn.gb & 7) * 8;
                const AccT *W2 = W2b + (((int64_t)Q * nbR + P) * nbR + R) * 512 + (rn.gb & 7) * 64 + (rn.ga & 7) * 8;
                const double u0 = U[(int64_t)rn.ga * vR + rn.gb];
                const double *u1 = U + vr + (int64_t)rn.ga * vR + rn.gc0;
                const double *u2 = U + 2 * vr + (int64_t)rn.gb * vR + rn.gc0;
                const double *tk = a.t1 + (int64_t)k * vR + rn.gc0;
                const double *evp = a.ev + rn.gc0;
                const double ti = a.t1[(int64_t)i * vR + rn.ga], tj = a.t1[(int64_t)j * vR + rn.gb];
                const double eab = eocc - a.ev[rn.ga] - a.ev[rn.gb];
#pragma unroll
                for (int cp = 0; cp < EB; cp++) {
                    const double w = (double)W1[cp] + (double)W2[cp];
                    Y[cp] = w + u0 * tk[cp] + u1[cp] * tj + u2[cp] * ti;
                    WD[h][cp] = w * recip(eab - evp[cp]);
                }
                ok[h] = true;
            } else {
#pragma unroll
                for (int cp = 0; cp < EB; cp++)
                    Y[cp] = 0.0;
            }
#pragma unroll
            for (int cp = 0; cp < EB; cp++)
                Yl[d * PB + ap * PX + bp * PY + cp] = Y[cp];
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (h >= nrow || !ok[h])
                continue;
            const int rt = tid + EPI2_THREADS * h;
            const OrbitRow r = orbit_row_tab<PB, PX, PY>(a.orbtab, orb, rt);  // reloaded: fewer live registers
            const int self = (rt >> 6) * PB + ((rt >> 3) & 7) * PX + (rt & 7) * PY;
            // the row's (at most 8) terms in plain FP64, the rows compensated (Neumaier)
            double rs = 0.0;
#pragma unroll
            for (int cp = 0; cp < EB; cp++) {
                if (r.gc0 + cp < v) {  // Yl of masked columns is finite: padding is zero
                    const double r3 = 4.0 * Yl[self + cp] + Yl[r.off[0] + cp * r.step[0]] +
                                      Yl[r.off[1] + cp * r.step[1]] -
                                      2.0 * (Yl[r.off[2] + cp * r.step[2]] + Yl[r.off[3] + cp * r.step[3]] +
                                             Yl[r.off[4] + cp * r.step[4]]);
                    rs += WD[h][cp] * r3;
                }
            }
            nadd(s, c, rs);
        }
        __syncthreads();
    }
    // fixed-order reduction: xor butterfly in each wave, then the waves in order
    wave_reduce(s, c);
    const int wv = tid >> 6;
    if ((tid & 63) == 0) {
        red[2 * wv] = s;
        red[2 * wv + 1] = c;
    }
    __syncthreads();
    if (tid == 0) {
        double ts = red[0], tc = red[1];
        for (int w = 1; w < EPI2_THREADS / 64; w++) {
            double ss, ee;
            two_sum(ts, red[2 * w], ss, ee);
            tc = tc + red[2 * w + 1] + ee;
            ts = ss;
        }
        a.partials[(int64_t)bat * a.ngrp + grp] = ts + tc;
    }
}

template <typename AccT>
__global__ __launch_bounds__(EPI2_THREADS, 3) void epilogue256_kernel(EpiArgs a) {
    constexpr int EB = 8, PY = 9, PX = EB * PY + 1, PB = EB * PX;
    __shared__ double Yl[6 * PB];
    __shared__ double red[2 * (EPI2_THREADS / 64)];
    const int nx = gridDim.x;
    const int lin = xcd_remap(blockIdx.y * nx + blockIdx.x, nx * gridDim.y);
    epilogue256_body<AccT>(a, lin % nx, lin / nx, Yl, red);
}

int epilogue_groups(int norb) { return (norb + EPI_RUN - 1) / EPI_RUN; }

bool orbit_table(const std::vector<int> &orbits, std::vector<uint64_t> &tab) {
    const size_t norb = orbits.size() / 3;
    tab.assign(norb * 6, 0);
    for (int b : orbits)
        if (b < 0 || b >= kMaxOrbitBlocks)  // block indices are packed into 10 bits
            return false;
    for (size_t orb = 0; orb < norb; orb++) {
        const int A = orbits[3 * orb], B = orbits[3 * orb + 1], C = orbits[3 * orb + 2];
        const int bl1 = B != A ? B : C;
        auto slot = [&](int blk) { return blk == A ? 0 : (blk == bl1 ? 1 : 2); };
        const int sB = slot(B), sC = slot(C);
        // distinct blocks in permutation order, keyed by their slot pattern
        int key_of[27];
        for (int &x : key_of)
            x = -1;
        int blk[6][3], sl[6][3], nd = 0;
        const int perms[6][3] = {{A, B, C}, {B, C, A}, {C, A, B}, {C, B, A}, {A, C, B}, {B, A, C}};
        const int pslots[6][3] = {{0, sB, sC}, {sB, sC, 0}, {sC, 0, sB}, {sC, sB, 0}, {0, sC, sB}, {sB, 0, sC}};
        for (int p = 0; p < 6; p++) {
            const int key = pslots[p][0] * 9 + pslots[p][1] * 3 + pslots[p][2];
            if (key_of[key] < 0) {
                key_of[key] = nd;
                for (int x = 0; x < 3; x++) {
                    blk[nd][x] = perms[p][x];
                    sl[nd][x] = pslots[p][x];
                }
                nd++;
            }
        }
        auto dindex = [&](int s0, int s1, int s2) { return key_of[s0 * 9 + s1 * 3 + s2]; };
        for (int d = 0; d < 6; d++) {
            uint64_t e = (uint64_t)nd << 30;
            if (d < nd) {
                const int sP = sl[d][0], sQ = sl[d][1], sR = sl[d][2];
                e |= (uint64_t)blk[d][0] | ((uint64_t)blk[d][1] << 10) | ((uint64_t)blk[d][2] << 20);
                // the five permuted reads of r3: Y(lb,lc,la), Y(lc,la,lb), Y(lc,lb,la), Y(la,lc,lb), Y(lb,la,lc)
                const int di[5] = {dindex(sQ, sR, sP), dindex(sR, sP, sQ), dindex(sR, sQ, sP), dindex(sP, sR, sQ),
                                   dindex(sQ, sP, sR)};
                for (int n = 0; n < 5; n++)
                    e |= (uint64_t)di[n] << (33 + 3 * n);
            }
            tab[orb * 6 + d] = e;
        }
    }
    return true;
}

hipError_t launch_epilogue(int prec, const EpiArgs &a, int nbatch, hipStream_t s, int slices) {
    const dim3 grid((unsigned)a.ngrp, (unsigned)nbatch);
    // 256 threads beside an int8-digit W-build of one fixed index per work-group, which leaves a
    // work-group slot per CU that this epilogue fits (5.31 s -> 5.17 s per step at o=30, v=300, S = 5);
    // 384 beside the others (FP64 9.77 s against 9.91 s; S = 5 with two fixed indices 5.18 s against
    // 5.25 s). A/B override: 256 or 384.
    bool narrow = prec == PI8 && i8_fixed_per_group(slices) == 1;
    if (const char *e = std::getenv("SCESS_EPI_THREADS"))
        narrow = std::atoi(e) == 256 ? true : std::atoi(e) == 384 ? false : narrow;
    if (narrow) {
        if (prec == P64 || prec == PI8)
            hipLaunchKernelGGL(epilogue256_kernel<double>, grid, dim3(EPI2_THREADS), 0, s, a);
        else
            hipLaunchKernelGGL(epilogue256_kernel<float>, grid, dim3(EPI2_THREADS), 0, s, a);
        return hipGetLastError();
    }
    if (prec == P64 || prec == PI8)  // both build W in FP64
        hipLaunchKernelGGL(epilogue_kernel<double>, grid, dim3(EPI_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL(epilogue_kernel<float>, grid, dim3(EPI_THREADS), 0, s, a);
    return hipGetLastError();
}

// -------------------------------------------------------------- reductions
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *partials, int norb, const int *ijk_index,
                                                              double *F) {
    __shared__ double red_s[256], red_c[256];
    const double *p = partials + (int64_t)blockIdx.x * norb;
    double s = 0.0, c = 0.0;
    for (int t = threadIdx.x; t < norb; t += 256)
        nadd(s, c, p[t]);
    const double tot = block_reduce_256(s, c, red_s, red_c);
    if (threadIdx.x == 0)
        F[ijk_index[blockIdx.x]] = tot;
}

hipError_t launch_reduce_partials(const double *partials, int norb, const int *ijk_index, double *F, int nbatch,
                                  hipStream_t s) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)nbatch), dim3(256), 0, s, partials, norb, ijk_index, F);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void reduce_energy_kernel(const double *F, const int *ijk_all, int nall, double *out) {
    __shared__ double red_s[256], red_c[256];
    double s = 0.0, c = 0.0;
    for (int t = threadIdx.x; t < nall; t += 256) {
        const int *q = ijk_all + 3 * t;
        const double w = (q[0] == q[2]) ? 1.0 / 3.0 : ((q[0] == q[1] || q[1] == q[2]) ? 1.0 : 2.0);
        nadd(s, c, w * F[t]);
    }
    const double tot = block_reduce_256(s, c, red_s, red_c);
    if (threadIdx.x == 0)
        out[0] = tot;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *p, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        p[t] = 0.0;
}

hipError_t launch_fill_zero(double *p, int64_t n, hipStream_t s) {
    if (n <= 0)
        return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_reduce_energy(const double *F, const int *ijk_all, int nall, double *out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_energy_kernel, dim3(1), dim3(256), 0, s, F, ijk_all, nall, out);
    return hipGetLastError();
}

}  // namespace scess
