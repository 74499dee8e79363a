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

#include <cstdint>

namespace scess {

typedef double d4 __attribute__((ext_vector_type(4)));
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
template <int PREC>
__global__ void pack_kernel(int mode, const double *__restrict__ src, void *__restrict__ dst, int o_, int v_,
                            int64_t kpad, int64_t nrows) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nrows * kpad)
        return;
    const int64_t o = o_, v = v_;
    const int64_t row = gid / kpad, k = gid - row * kpad;
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
    int bid = blockIdx.x;
    const bool fam1 = bid < a.ntile1;
    if (!fam1)
        bid -= a.ntile1;
    const int64_t fx = bid / nt2;
    const int tt = bid - (int)fx * nt2;
    const int tr = tt / a.nt, tc = tt - tr * a.nt;
    const int *q = a.ijk + 3 * blockIdx.y;
    const int64_t i = q[0], j = q[1], k = q[2];
    const int v = a.v;
    const int r0 = tr * BM, c0 = tc * BM;

    const int nseg = fam1 ? 8 : 4;
    const int total = fam1 ? 4 * (a.nchv + a.ncho) : 2 * (a.nchv + a.ncho);

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
    uint4 ra[NVT], rb[NVT];

    auto load = [&]() {
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
                sg = get_seg(a, fam1, s, i, j, k, fx);
        }
    };
    auto store = [&](int buf) {
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

    load();
    store(0);
    __syncthreads();
    for (int c = 0; c < total; ++c) {
        const bool more = c + 1 < total;
        if (more)
            load();
        compute(c & 1);
        if (more)
            store((c + 1) & 1);
        __syncthreads();
    }

    Acc *W = (Acc *)(fam1 ? a.W1 : a.W2) + (int64_t)blockIdx.y * a.wstride + fx * (int64_t)a.vR * a.vR;
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLK; y++) {
            const int col = c0 + wn * BLK * 16 + y * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = r0 + wm * BLK * 16 + x * 16 + (PREC == P64 ? lq + 4 * r : 4 * lq + r);
                if (row < v && col < v)
                    W[(int64_t)row * a.vR + col] = acc[x][y][r];
            }
        }
}

int wbuild_tile(int prec, int v) {
    (void)prec;
    if (v >= 240)
        return 5;
    if (v > 32)
        return 2;
    return 1;
}

template <int PREC>
static hipError_t launch_wbuild_p(int blk, const WBuildArgs &a, int nbatch, hipStream_t s) {
    const dim3 grid((unsigned)(2 * a.ntile1), (unsigned)nbatch);
    switch (blk) {
    case 1: hipLaunchKernelGGL((wbuild_kernel<PREC, 1>), grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((wbuild_kernel<PREC, 2>), grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL((wbuild_kernel<PREC, 5>), grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_wbuild(int prec, int blk, const WBuildArgs &a, int nbatch, hipStream_t s) {
    if (prec == P64)
        return launch_wbuild_p<P64>(blk, a, nbatch, s);
    if (prec == P32)
        return launch_wbuild_p<P32>(blk, a, nbatch, s);
    return launch_wbuild_p<PB3>(blk, a, nbatch, s);
}

// ---------------------------------------------------------------- epilogue
template <typename AccT>
__global__ __launch_bounds__(256) void epilogue_kernel(EpiArgs a) {
    constexpr int EB = 8;
    constexpr int LM = 3 * EB;  // largest local extent
    __shared__ double Wl[6 * EB * EB * EB];
    __shared__ double U[3][LM * LM];
    __shared__ double tl[3][LM];
    __shared__ double evl[LM];
    __shared__ double red_s[256], red_c[256];
    __shared__ int dist[6][3];
    __shared__ int smap[27];
    __shared__ int bl[3];
    __shared__ int nd_s, nl_s;

    const int tid = threadIdx.x;
    const int bat = blockIdx.y;
    const int *ob = a.orbits + 3 * blockIdx.x;
    const int *q = a.ijk + 3 * bat;
    const int i = q[0], j = q[1], k = q[2];
    const int o = a.o, v = a.v;
    const int64_t vR = a.vR;

    if (tid == 0) {
        const int A = ob[0], B = ob[1], C = ob[2];
        int nl = 0;
        bl[nl++] = A;
        if (B != A)
            bl[nl++] = B;
        if (C != B)
            bl[nl++] = C;
        for (int t = nl; t < 3; t++)
            bl[t] = -1;
        for (int t = 0; t < 27; t++)
            smap[t] = -1;
        const int perm[6][3] = {{A, B, C}, {B, C, A}, {C, A, B}, {C, B, A}, {A, C, B}, {B, A, C}};
        int nd = 0;
        for (int p = 0; p < 6; p++) {
            int key = 0;
            for (int d = 0; d < 3; d++) {
                const int sl = perm[p][d] == bl[0] ? 0 : (perm[p][d] == bl[1] ? 1 : 2);
                key = key * 3 + sl;
            }
            if (smap[key] < 0) {
                smap[key] = nd;
                dist[nd][0] = perm[p][0];
                dist[nd][1] = perm[p][1];
                dist[nd][2] = perm[p][2];
                nd++;
            }
        }
        nd_s = nd;
        nl_s = nl;
    }
    __syncthreads();
    const int nd = nd_s, nl = nl_s, L = nl * EB;

    // W of the distinct blocks of the orbit
    const AccT *W1 = (const AccT *)a.W1 + (int64_t)bat * a.wstride;
    const AccT *W2 = (const AccT *)a.W2 + (int64_t)bat * a.wstride;
    for (int idx = tid; idx < nd * 512; idx += 256) {
        const int d = idx >> 9, e = idx & 511;
        const int ga = dist[d][0] * EB + (e >> 6), gb = dist[d][1] * EB + ((e >> 3) & 7),
                  gc = dist[d][2] * EB + (e & 7);
        double w = 0.0;
        if (ga < v && gb < v && gc < v)
            w = (double)W1[(ga * vR + gb) * vR + gc] + (double)W2[(gb * vR + ga) * vR + gc];
        Wl[idx] = w;
    }
    // V pieces: V(x,y,z) = Uk[x][y] t_k[z] + Uj[x][z] t_j[y] + Ui[y][z] t_i[x]
    const double *S = a.S;
    const int64_t vv = (int64_t)v * v;
    for (int idx = tid; idx < L * L; idx += 256) {
        const int u1 = idx / L, u2 = idx - u1 * L;
        const int x = bl[u1 >> 3] * EB + (u1 & 7), y = bl[u2 >> 3] * EB + (u2 & 7);
        double uk = 0.0, uj = 0.0, ui = 0.0;
        if (x < v && y < v) {
            uk = 0.5 * (S[((int64_t)i * o + j) * vv + (int64_t)x * v + y] + S[((int64_t)j * o + i) * vv + (int64_t)y * v + x]);
            uj = 0.5 * (S[((int64_t)i * o + k) * vv + (int64_t)x * v + y] + S[((int64_t)k * o + i) * vv + (int64_t)y * v + x]);
            ui = 0.5 * (S[((int64_t)j * o + k) * vv + (int64_t)x * v + y] + S[((int64_t)k * o + j) * vv + (int64_t)y * v + x]);
        }
        U[0][u1 * LM + u2] = uk;
        U[1][u1 * LM + u2] = uj;
        U[2][u1 * LM + u2] = ui;
    }
    if (tid < L) {
        const int g = bl[tid >> 3] * EB + (tid & 7);
        const bool ok = g < v;
        tl[0][tid] = ok ? a.t1[(int64_t)i * v + g] : 0.0;
        tl[1][tid] = ok ? a.t1[(int64_t)j * v + g] : 0.0;
        tl[2][tid] = ok ? a.t1[(int64_t)k * v + g] : 0.0;
        evl[tid] = ok ? a.e[o + g] : 0.0;
    }
    __syncthreads();

    auto Yat = [&](int x, int y, int z) -> double {
        const int d = smap[(x >> 3) * 9 + (y >> 3) * 3 + (z >> 3)];
        return Wl[(d << 9) + ((x & 7) << 6) + ((y & 7) << 3) + (z & 7)] + U[0][x * LM + y] * tl[2][z] +
               U[1][x * LM + z] * tl[1][y] + U[2][y * LM + z] * tl[0][x];
    };

    const double eocc = a.e[i] + a.e[j] + a.e[k];
    double s = 0.0, c = 0.0;
    for (int idx = tid; idx < nd * 512; idx += 256) {
        const int d = idx >> 9, e = idx & 511;
        const int P = dist[d][0], Q = dist[d][1], R = dist[d][2];
        const int ap = e >> 6, bp = (e >> 3) & 7, cp = e & 7;
        if (P * EB + ap >= v || Q * EB + bp >= v || R * EB + cp >= v)
            continue;
        const int sP = P == bl[0] ? 0 : (P == bl[1] ? 1 : 2);
        const int sQ = Q == bl[0] ? 0 : (Q == bl[1] ? 1 : 2);
        const int sR = R == bl[0] ? 0 : (R == bl[1] ? 1 : 2);
        const int la = sP * EB + ap, lb = sQ * EB + bp, lc = sR * EB + cp;
        const double r3 = 4.0 * Yat(la, lb, lc) + Yat(lb, lc, la) + Yat(lc, la, lb) -
                          2.0 * (Yat(lc, lb, la) + Yat(la, lc, lb) + Yat(lb, la, lc));
        const double D = eocc - evl[la] - evl[lb] - evl[lc];
        nadd(s, c, Wl[idx] * r3 / D);
    }
    const double tot = block_reduce_256(s, c, red_s, red_c);
    if (tid == 0)
        a.partials[(int64_t)bat * a.norb + blockIdx.x] = tot;
}

hipError_t launch_epilogue(int prec, const EpiArgs &a, int nbatch, hipStream_t s) {
    const dim3 grid((unsigned)a.norb, (unsigned)nbatch);
    if (prec == P64)
        hipLaunchKernelGGL(epilogue_kernel<double>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(epilogue_kernel<float>, grid, dim3(256), 0, s, a);
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

hipError_t launch_reduce_energy(const double *F, const int *ijk_all, int nall, double *out, hipStream_t s) {
    hipLaunchKernelGGL(reduce_energy_kernel, dim3(1), dim3(256), 0, s, F, ijk_all, nall, out);
    return hipGetLastError();
}

}  // namespace scess
