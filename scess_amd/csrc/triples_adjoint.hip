// gfx950 kernels of the (T) adjoint: dE(T)/d(ovvv), d(ovoo), d(ovov), dt1, dt2.
//
// E = sum_ijk w_ijk sum_abc W r3(Y)/D with Y = W + V and w_ijk = 2/deg_ijk. Since
// r3 is symmetric,  G_W = dE/dW = w (r3(Y)/D + r3(W/D))  and  G_V = dE/dV = w r3(W/D).
// W is recomputed per ijk (the forward W-build), G is formed by an orbit kernel
// like the forward epilogue, and the six permutation terms
//   w(x,y,z; p,q,r) = sum_f X[p][x][y][f] T[r][q][z][f] + sum_m Rn[p][q][x][m] TM[r][y][z][m]
// are back-contracted by one generic FP64 MFMA GEMM:
//   dX[p][x][y][f]  += sum_z     G(x,y,z) T[r][q][z][f]
//   dT[r][q][z][f]  += sum_{x,y} G(x,y,z) X[p][x][y][f]
//   dRn[p][q][x][m] += sum_{y,z} G(x,y,z) TM[r][y][z][m]
//   dTM[r][y][z][m] += sum_x     Rn[p][q][x][m] G(x,y,z)
// Occupied triples are processed one after another and every kernel writes each
// output element from one thread in a fixed order (split-K partials are summed
// in split order), so the gradients are deterministic.
#include <type_traits>
#include "triples_internal.h"

#include <cstdint>
#include <cstdlib>

namespace scess {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

__device__ inline void nadd_a(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// ------------------------------------------------------------------ G kernel
// Orbit geometry as in the forward epilogue (8-wide virtual blocks).
struct OrbitRowA {
    bool own;
    int ga, gb, gc0;
    int off[5], step[5];
};

template <int PB, int PX, int PY>
__device__ inline OrbitRowA orbit_row_a(const int *orbits, int orb, int tid) {
    constexpr int EB = 8;
    OrbitRowA r;
    const int A = orbits[3 * orb], B = orbits[3 * orb + 1], C = orbits[3 * orb + 2];
    const int bl1 = B != A ? B : C;
    auto slot = [&](int blk) { return blk == A ? 0 : (blk == bl1 ? 1 : 2); };
    const int sB = slot(B), sC = slot(C);
    uint64_t map_lo = 0, map_hi = 0;
    uint32_t seen = 0;
    int nd = 0;
    const int d = tid >> 6, ap = (tid >> 3) & 7, bp = tid & 7;
    int P = 0, Q = 0, R = 0, sP = 0, sQ = 0, sR = 0;
    auto add = [&](int p0, int p1, int p2, int s0, int s1, int s2) {
        const int key = s0 * 9 + s1 * 3 + s2;
        if (!(seen >> key & 1u)) {
            seen |= 1u << key;
            if (key < 21)
                map_lo |= (uint64_t)nd << (3 * key);
            else
                map_hi |= (uint64_t)nd << (3 * (key - 21));
            if (nd == d) {
                P = p0;
                Q = p1;
                R = p2;
                sP = s0;
                sQ = s1;
                sR = s2;
            }
            nd++;
        }
    };
    add(A, B, C, 0, sB, sC);
    add(B, C, A, sB, sC, 0);
    add(C, A, B, sC, 0, sB);
    add(C, B, A, sC, sB, 0);
    add(A, C, B, 0, sC, sB);
    add(B, A, C, sB, 0, sC);
    auto dindex = [&](int s0, int s1, int s2) -> int {
        const int key = s0 * 9 + s1 * 3 + s2;
        return (int)((key < 21 ? map_lo >> (3 * key) : map_hi >> (3 * (key - 21))) & 7u);
    };
    r.own = d < nd;
    r.ga = P * EB + ap;
    r.gb = Q * EB + bp;
    r.gc0 = R * EB;
    r.off[0] = dindex(sQ, sR, sP) * PB + bp * PX + ap;  // (b, c, a)
    r.step[0] = PY;
    r.off[1] = dindex(sR, sP, sQ) * PB + ap * PY + bp;  // (c, a, b)
    r.step[1] = PX;
    r.off[2] = dindex(sR, sQ, sP) * PB + bp * PY + ap;  // (c, b, a)
    r.step[2] = PX;
    r.off[3] = dindex(sP, sR, sQ) * PB + ap * PX + bp;  // (a, c, b)
    r.step[3] = PY;
    r.off[4] = dindex(sQ, sP, sR) * PB + bp * PX + ap * PY;  // (b, a, c)
    r.step[4] = 1;
    return r;
}

constexpr int GK_THREADS = 384;

// G_W and G_V (row-major [vR][vR][vR], FP64) of one occupied triple.
template <typename AccT>
__global__ __launch_bounds__(GK_THREADS) void adjoint_g_kernel(EpiArgs a, double wgt, double *GW, double *GV) {
    constexpr int EB = 8;
    constexpr int PY = 9, PX = EB * PY + 1, PB = EB * PX;
    __shared__ double Yl[6 * PB];
    __shared__ double Zl[6 * PB];
    const int tid = threadIdx.x;
    const int orb = blockIdx.x;
    const int *q = a.ijk;
    const int i = q[0], j = q[1], k = q[2];
    const int v = a.v, vR = a.vR, nbR = a.vR >> 3;
    const int64_t vr = (int64_t)v * vR;
    const double eocc = a.e[i] + a.e[j] + a.e[k];
    const double *U = a.U;
    const AccT *W1b = (const AccT *)a.W1;
    const AccT *W2b = (const AccT *)a.W2;
    const int ap = (tid >> 3) & 7, bp = tid & 7, d = tid >> 6;

    const OrbitRowA r = orbit_row_a<PB, PX, PY>(a.orbits, orb, tid);
    double Y0[EB], Z0[EB], D0[EB];
    const bool rowok = r.own && r.ga < v && r.gb < v;
    if (rowok) {
        const int P = r.ga >> 3, Q = r.gb >> 3, R = r.gc0 >> 3;
        const AccT *W1 = W1b + (((int64_t)P * nbR + Q) * nbR + R) * 512 + (r.ga & 7) * 64 + (r.gb & 7) * 8;
        const AccT *W2 = W2b + (((int64_t)Q * nbR + P) * nbR + R) * 512 + (r.gb & 7) * 64 + (r.ga & 7) * 8;
        const double u0 = U[(int64_t)r.ga * vR + r.gb];
        const double *u1 = U + vr + (int64_t)r.ga * vR + r.gc0;
        const double *u2 = U + 2 * vr + (int64_t)r.gb * vR + r.gc0;
        const double *tk = a.t1 + (int64_t)k * vR + r.gc0;
        const double ti = a.t1[(int64_t)i * vR + r.ga], tj = a.t1[(int64_t)j * vR + r.gb];
        const double eab = eocc - a.ev[r.ga] - a.ev[r.gb];
#pragma unroll
        for (int cp = 0; cp < EB; cp++) {
            const bool ok = r.gc0 + cp < v;
            const double w = (double)W1[cp] + (double)W2[cp];
            const double Dv = eab - a.ev[r.gc0 + cp];
            Y0[cp] = ok ? w + u0 * tk[cp] + u1[cp] * tj + u2[cp] * ti : 0.0;
            Z0[cp] = ok ? w / Dv : 0.0;
            D0[cp] = ok ? Dv : 1.0;
        }
    } else {
#pragma unroll
        for (int cp = 0; cp < EB; cp++) {
            Y0[cp] = Z0[cp] = 0.0;
            D0[cp] = 1.0;
        }
    }
    if (r.own)
#pragma unroll
        for (int cp = 0; cp < EB; cp++) {
            Yl[d * PB + ap * PX + bp * PY + cp] = Y0[cp];
            Zl[d * PB + ap * PX + bp * PY + cp] = Z0[cp];
        }
    __syncthreads();
    if (rowok) {
        double *gw = GW + ((int64_t)r.ga * vR + r.gb) * vR + r.gc0;
        double *gv = GV + ((int64_t)r.ga * vR + r.gb) * vR + r.gc0;
#pragma unroll
        for (int cp = 0; cp < EB; cp++) {
            double y3 = 0.0, z3 = 0.0;
            if (r.gc0 + cp < v) {
                y3 = 4.0 * Y0[cp] + Yl[r.off[0] + cp * r.step[0]] + Yl[r.off[1] + cp * r.step[1]] -
                     2.0 * (Yl[r.off[2] + cp * r.step[2]] + Yl[r.off[3] + cp * r.step[3]] + Yl[r.off[4] + cp * r.step[4]]);
                z3 = 4.0 * Z0[cp] + Zl[r.off[0] + cp * r.step[0]] + Zl[r.off[1] + cp * r.step[1]] -
                     2.0 * (Zl[r.off[2] + cp * r.step[2]] + Zl[r.off[3] + cp * r.step[3]] + Zl[r.off[4] + cp * r.step[4]]);
            }
            gw[cp] = wgt * (y3 / D0[cp] + z3);
            gv[cp] = wgt * z3;
        }
    }
}

hipError_t launch_adjoint_g(int prec, const EpiArgs &a, double wgt, double *GW, double *GV, hipStream_t s) {
    if (prec == P64 || prec == PI8)  // both build W in FP64
        hipLaunchKernelGGL(adjoint_g_kernel<double>, dim3((unsigned)a.norb), dim3(GK_THREADS), 0, s, a, wgt, GW, GV);
    else
        hipLaunchKernelGGL(adjoint_g_kernel<float>, dim3((unsigned)a.norb), dim3(GK_THREADS), 0, s, a, wgt, GW, GV);
    return hipGetLastError();
}

// ------------------------------------------------------------ generic GEMM
// C[beta][m][n] += alpha * sum_seg sum_{k1,k2} A(m, k1, k2) B(n, k1, k2), FP64 MFMA 16x16x4,
// 64x64 tiles (2x2 waves of 2x2 blocks), K chunks of 16 along k2 staged through LDS.
// With splits > 1 (no batch), split p handles a fixed contiguous range of the K chunks
// and writes its partial tile to scratch[p]; gemm_reduce_kernel adds them in split order.
// pitch 80: the four 16-lane groups of a fragment read land 32 banks apart (conflict-free halves)
constexpr int GT = 64, GKC = 16, GPITCH = 80;

// One operand's 64 rows x 16 k of chunk (k1, k2c): 4 elements per thread. Sixteen consecutive
// lanes take 16 consecutive elements along the operand's contiguous dimension (k when the
// operand is K-major, rows otherwise), so every load instruction reads 128-byte runs.
struct TileLoad {
    bool kinner;
    __device__ TileLoad(const GemmOp &op) : kinner(op.s2 == 1 || (op.sr != 1 && (op.s2 < op.sr))) {}
    __device__ void coords(int tid, int u, int &rr, int &kk) const {
        if (kinner) {
            rr = (tid >> 4) + 16 * u;
            kk = tid & 15;
        } else {
            kk = tid >> 4;
            rr = (tid & 15) + 16 * u;
        }
    }
};

__device__ inline void gemm_fetch(const GemmOp &op, const TileLoad &tl, int64_t base, int r0, int rext, int k1, int k2c,
                                  int K2, double *reg, int tid) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        int rr, kk;
        tl.coords(tid, u, rr, kk);
        const int gr = r0 + rr, gk = k2c * GKC + kk;
        reg[u] = gr < rext && gk < K2 ? op.base[base + (int64_t)gr * op.sr + (int64_t)k1 * op.s1 + (int64_t)gk * op.s2]
                                      : 0.0;
    }
}

__device__ inline void gemm_stash(const TileLoad &tl, const double *reg, double *lds, int tid) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        int rr, kk;
        tl.coords(tid, u, rr, kk);
        lds[kk * GPITCH + rr] = reg[u];
    }
}

// Paired loads for operands with a contiguous dimension and even strides (16-byte aligned pairs):
// MODE 1 rows contiguous (sr == 1), a pair is rows (2j, 2j + 1) at one k, 32 lanes per k row;
// MODE 2 K contiguous (s2 == 1), a pair is k (2j, 2j + 1) of one row, 8 lanes per row. Two pairs
// per thread; reg[2u], reg[2u + 1] hold pair u. Elements past rext / K2 are zero.
template <int MODE> __device__ inline void pair_coords(int tid, int u, int &rr, int &kk) {
    const int p = tid + 256 * u;
    if constexpr (MODE == 1) {
        kk = p >> 5;
        rr = (p & 31) * 2;
    } else {
        rr = p >> 3;
        kk = (p & 7) * 2;
    }
}

template <int MODE>
__device__ inline void gemm_fetch_pair(const GemmOp &op, int64_t base, int r0, int rext, int k1, int k2c, int K2,
                                       double *reg, int tid) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
        int rr, kk;
        pair_coords<MODE>(tid, u, rr, kk);
        const int gr = r0 + rr, gk = k2c * GKC + kk;
        const bool full = MODE == 1 ? (gk < K2 && gr + 1 < rext) : (gr < rext && gk + 1 < K2);
        const bool first = gr < rext && gk < K2;
        const double *src = op.base + base + (int64_t)gr * op.sr + (int64_t)k1 * op.s1 + (int64_t)gk * op.s2;
        if (full) {
            const d2 x = *(const d2 *)src;
            reg[2 * u] = x[0];
            reg[2 * u + 1] = x[1];
        } else {
            reg[2 * u] = first ? src[0] : 0.0;
            reg[2 * u + 1] = 0.0;
        }
    }
}

template <int MODE> __device__ inline void gemm_stash_pair(const double *reg, double *lds, int tid) {
#pragma unroll
    for (int u = 0; u < 2; u++) {
        int rr, kk;
        pair_coords<MODE>(tid, u, rr, kk);
        if constexpr (MODE == 1) {
            *(d2 *)&lds[kk * GPITCH + rr] = d2{reg[2 * u], reg[2 * u + 1]};
        } else {
            lds[kk * GPITCH + rr] = reg[2 * u];
            lds[(kk + 1) * GPITCH + rr] = reg[2 * u + 1];
        }
    }
}

// One operand of chunk (k1, k2c) in load mode MODE (0 element-wise, 1 / 2 paired)
template <int MODE>
__device__ inline void fetch_op(const GemmOp &op, int64_t base, int r0, int rext, int k1, int k2c, int K2, double *reg,
                                int tid) {
    if constexpr (MODE == 0)
        gemm_fetch(op, TileLoad(op), base, r0, rext, k1, k2c, K2, reg, tid);
    else
        gemm_fetch_pair<MODE>(op, base, r0, rext, k1, k2c, K2, reg, tid);
}
template <int MODE> __device__ inline void stash_op(const GemmOp &op, const double *reg, double *lds, int tid) {
    if constexpr (MODE == 0)
        gemm_stash(TileLoad(op), reg, lds, tid);
    else
        gemm_stash_pair<MODE>(reg, lds, tid);
}

// A0, A1, B0, B1: load modes of A and B in segments 0 and 1
template <int A0, int A1, int B0, int B1>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    // two LDS buffers; the global loads of chunk c + 1 are in flight while chunk c is multiplied
    __shared__ __attribute__((aligned(16))) double As[2][GKC * GPITCH];
    __shared__ __attribute__((aligned(16))) double Bs[2][GKC * GPITCH];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const int ntn = (g.N + GT - 1) / GT;
    // tiles of one (batch, split) on one XCD, so the tiles sharing an operand panel share its L2
    const int nx = gridDim.x, ny = gridDim.y;
    const int lin = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    const int bx = lin % nx, byz = lin / nx;
    const int tm = bx / ntn, tn = bx - tm * ntn;
    const int m0 = tm * GT, n0 = tn * GT;
    const int beta = byz % ny, split = byz / ny;

    // flat chunk index over the segments; this split takes a fixed contiguous range of it
    // (segment values as scalars: a runtime-indexed local array would live in scratch memory)
    const int nk20 = (g.seg[0].K2 + GKC - 1) / GKC, nk21 = g.nseg > 1 ? (g.seg[1].K2 + GKC - 1) / GKC : 1;
    const int nch0 = g.seg[0].K1 * nk20, nch1 = g.nseg > 1 ? g.seg[1].K1 * nk21 : 0;
    const int total = nch0 + nch1;
    const int c_begin = (int)((int64_t)total * split / g.splits), c_end = (int)((int64_t)total * (split + 1) / g.splits);

    d4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    // MFMA blocks of this wave wholly past M or N are skipped (wave-uniform)
    bool vm[2], vn[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
        vm[x] = m0 + wm * 32 + x * 16 < g.M;
        vn[x] = n0 + wn * 32 + x * 16 < g.N;
    }

    // two register sets: chunk c + 2 is loaded while chunk c is multiplied (c + 1 waits in the other)
    double ra0[4], rb0[4], ra1[4], rb1[4];
    auto fetch = [&](int ch, double *ra, double *rb) {
        if (ch < nch0) {
            const GemmSeg &sg = g.seg[0];
            const int k1 = ch / nk20, k2c = ch - k1 * nk20;
            fetch_op<A0>(sg.A, (int64_t)beta * sg.A.sb, m0, g.M, k1, k2c, sg.K2, ra, tid);
            fetch_op<B0>(sg.B, (int64_t)beta * sg.B.sb, n0, g.N, k1, k2c, sg.K2, rb, tid);
        } else {
            const GemmSeg &sg = g.seg[1];
            const int cl = ch - nch0, k1 = cl / nk21, k2c = cl - k1 * nk21;
            fetch_op<A1>(sg.A, (int64_t)beta * sg.A.sb, m0, g.M, k1, k2c, sg.K2, ra, tid);
            fetch_op<B1>(sg.B, (int64_t)beta * sg.B.sb, n0, g.N, k1, k2c, sg.K2, rb, tid);
        }
    };
    int buf = 0;
    auto step = [&](int ch, double *ra, double *rb) {
        if (ch < nch0) {
            stash_op<A0>(g.seg[0].A, ra, As[buf], tid);
            stash_op<B0>(g.seg[0].B, rb, Bs[buf], tid);
        } else {
            stash_op<A1>(g.seg[1].A, ra, As[buf], tid);
            stash_op<B1>(g.seg[1].B, rb, Bs[buf], tid);
        }
        __syncthreads();
        if (ch + 2 < c_end)
            fetch(ch + 2, ra, rb);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int kk = t * 4 + (lane >> 4);
            double af[2], bf[2];
#pragma unroll
            for (int x = 0; x < 2; x++)
                af[x] = As[buf][kk * GPITCH + wm * 32 + x * 16 + (lane & 15)];
#pragma unroll
            for (int y = 0; y < 2; y++)
                bf[y] = Bs[buf][kk * GPITCH + wn * 32 + y * 16 + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; x++)
                if (vm[x])
#pragma unroll
                    for (int y = 0; y < 2; y++)
                        if (vn[y])
                            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x], bf[y], acc[x][y], 0, 0, 0);
        }
        buf ^= 1;
    };
    if (c_begin < c_end)
        fetch(c_begin, ra0, rb0);
    if (c_begin + 1 < c_end)
        fetch(c_begin + 1, ra1, rb1);
    for (int ch = c_begin; ch < c_end; ch += 2) {
        step(ch, ra0, rb0);
        if (ch + 1 < c_end)
            step(ch + 1, ra1, rb1);
    }

#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            const int n = n0 + wn * 32 + y * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * r;
                if (m < g.M && n < g.N) {
                    if (g.splits == 1) {
                        double *cp = g.C + (int64_t)beta * g.sCb + (int64_t)m * g.sCm + (int64_t)n * g.sCn;
                        *cp += g.alpha * acc[x][y][r];
                    } else {
                        g.scratch[((int64_t)split * g.M + m) * g.N + n] = acc[x][y][r];
                    }
                }
            }
        }
}

// The same contraction for the layouts of the (T) back-contractions, with direct-to-LDS operand loads:
// B has unit stride along n (rows of T2 and of the integral blocks), A (a gradient array G) unit stride
// along m (AM = 0) or along k2 (AM = 1). A 16-deep chunk of a unit-stride-n (or m) operand is a [k][x]
// tile of 16 rows of 64 doubles (32 sixteen-byte vectors, vector p of row k at slot p ^ 8 (k & 1)); of a
// unit-stride-k operand a [m][k] tile of 64 rows of 16 doubles (8 vectors, vector p of row m at slot
// p ^ ((m >> 1) & 7)). Both put the 16 lanes of every MFMA operand read on distinct banks. Loads go
// global -> LDS (global_load_lds, 16 bytes per lane, 1 KiB per wave instruction) into a two-stage ring;
// vectors past M or N are clamped into the padded rows (cext), B rows with k2 >= K2 to row K2 - 1 (A
// reads zero there: kzero). Tiles, waves, MFMA blocks, splits and the store are those of gemm_kernel.
constexpr int KN_TILE = 16 * 64;  // doubles per operand chunk (8 KiB)

// source of the 16-byte vector that lane `lane` of wave load L (0..7) puts into an operand chunk;
// mfold > 0 (A of a folded batch): tile row x is row x % mfold of batch entry x / mfold
template <int MK, int W = 64>
__device__ inline const double *kn_src(const GemmOp &op, int64_t base, int x0, int xext, int k1, int k2c, int K2, int L,
                                       int lane, int mfold) {
    if constexpr (MK) {  // [m][k] tile: wave load L covers rows 8 L .. 8 L + 7
        const int mm = 8 * L + (lane >> 3), p = lane & 7;
        const int lv = p ^ ((mm >> 1) & 7);
        const int64_t k2 = min((int64_t)k2c * 16 + 2 * lv, op.cext - 2);
        int m = min(x0 + mm, xext - 1);
        if (mfold) {
            const int b = m / mfold;
            base += b * op.sb;
            m -= b * mfold;
        }
        return op.base + base + (int64_t)m * op.sr + (int64_t)k1 * op.s1 + k2;
    } else {  // [k][x] tile of W-double rows: wave load L covers rows 2 L, 2 L + 1 (W = 64) or row L (128)
        const int kk = W == 64 ? 2 * L + (lane >> 5) : L, p = W == 64 ? lane & 31 : lane;
        const int lv = p ^ ((kk & 1) << 3);
        int64_t x;
        if (mfold) {  // mfold is even, so the two rows of a vector lie in one batch entry
            const int xf = min(x0 + 2 * lv, xext - 2), b = xf / mfold;
            base += b * op.sb;
            x = xf - b * mfold;
        } else {
            x = min((int64_t)x0 + 2 * lv, op.cext - 2);
        }
        const int k2 = op.kzero ? k2c * 16 + kk : min(k2c * 16 + kk, K2 - 1);
        return op.base + base + (int64_t)k1 * op.s1 + (int64_t)k2 * op.s2 + x;
    }
}

// element (x, k) of an operand chunk in LDS (x = row of A or B within the tile, k = 0..15; [k][x] tiles
// with rows of W doubles)
template <int MK, int W = 64> __device__ inline double kn_frag(const double *t, int x, int k) {
    if constexpr (MK)
        return t[x * 16 + (((k >> 1) ^ ((x >> 1) & 7)) << 1) + (k & 1)];
    else
        return t[k * W + (((x >> 1) ^ ((k & 1) << 3)) << 1) + (x & 1)];
}

// WIDE = 2: tiles of 64 rows by 128 columns (each wave 32 x 64, four B wave loads per chunk), twice the
// MFMAs per barrier for 1.5x the staged bytes; two stages only.
template <int AM0, int AM1, int NST, int WIDE>
__global__ __launch_bounds__(256) void gemm_kn_kernel(GemmGroup gg) {
    static_assert(NST == 2 || NST == 3, "stages");
    static_assert(WIDE == 1 || (WIDE == 2 && NST == 2), "wide tiles: two stages");
    constexpr int BW = 64 * WIDE, NBY = 2 * WIDE;  // tile columns; MFMA blocks per wave along n
    __shared__ __attribute__((aligned(1024))) double lds[NST][KN_TILE + 16 * BW];  // [stage][A chunk, B chunk]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const int nx = gridDim.x, ny = gridDim.y;
    const int lin = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    const int bx = lin % nx, byz = lin / nx;
    const int beta = byz % ny, zz = byz / ny;
    const int prob = zz % gg.n, split = zz / gg.n;
    const GemmArgs &g = gg.g[prob];
    const int ntn = (g.N + BW - 1) / BW;
    const int tm = bx / ntn, tn = bx - tm * ntn;
    const int m0 = tm * GT, n0 = tn * BW;

    const int nk20 = (g.seg[0].K2 + 15) / 16, nk21 = g.nseg > 1 ? (g.seg[1].K2 + 15) / 16 : 1;
    const int nch0 = g.seg[0].K1 * nk20, nch1 = g.nseg > 1 ? g.seg[1].K1 * nk21 : 0;
    const int total = nch0 + nch1;
    const int c_begin = (int)((int64_t)total * split / g.splits), c_end = (int)((int64_t)total * (split + 1) / g.splits);

    d4 acc[2][NBY];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < NBY; y++)
            acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    bool vm[2], vn[NBY];
#pragma unroll
    for (int x = 0; x < 2; x++)
        vm[x] = m0 + wm * 32 + x * 16 < g.M;
#pragma unroll
    for (int y = 0; y < NBY; y++)
        vn[y] = n0 + wn * 32 * WIDE + y * 16 < g.N;

    // wave loads 2 wave and 2 wave + 1 of the A chunk and 2 WIDE wave .. 2 WIDE wave + 2 WIDE - 1 of the
    // B chunk are this wave's
    auto issue_seg = [&](const GemmSeg &sg, auto amk, int cl, int nk2, int st) {
        constexpr int MK = decltype(amk)::value;
        const int k1 = cl / nk2, k2c = cl - k1 * nk2;
        double *As = lds[st], *Bs = lds[st] + KN_TILE;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int L = 2 * wave + h;
            const double *sa = kn_src<MK>(sg.A, (int64_t)beta * sg.A.sb, m0, g.M, k1, k2c, sg.K2, L, lane, g.mfold);
            __builtin_amdgcn_global_load_lds((const void *)sa, (lds_void *)(As + L * 128), 16, 0, 0);
        }
#pragma unroll
        for (int h = 0; h < 2 * WIDE; h++) {
            const int L = 2 * WIDE * wave + h;
            const double *sb = kn_src<0, BW>(sg.B, (int64_t)beta * sg.B.sb, n0, g.N, k1, k2c, sg.K2, L, lane, 0);
            __builtin_amdgcn_global_load_lds((const void *)sb, (lds_void *)(Bs + L * 128), 16, 0, 0);
        }
    };
    auto issue = [&](int ch, int st) {
        if (ch < nch0)
            issue_seg(g.seg[0], std::integral_constant<int, AM0>{}, ch, nk20, st);
        else
            issue_seg(g.seg[1], std::integral_constant<int, AM1>{}, ch - nch0, nk21, st);
    };
    auto compute = [&](int st, auto amk) {
        constexpr int MK = decltype(amk)::value;
        const double *As = lds[st], *Bs = lds[st] + KN_TILE;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int kk = t * 4 + (lane >> 4);
            double af[2], bf[NBY];
#pragma unroll
            for (int x = 0; x < 2; x++)
                af[x] = kn_frag<MK>(As, wm * 32 + x * 16 + (lane & 15), kk);
#pragma unroll
            for (int y = 0; y < NBY; y++)
                bf[y] = kn_frag<0, BW>(Bs, wn * 32 * WIDE + y * 16 + (lane & 15), kk);
#pragma unroll
            for (int x = 0; x < 2; x++)
                if (vm[x])
#pragma unroll
                    for (int y = 0; y < NBY; y++)
                        if (vn[y])
                            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[x], bf[y], acc[x][y], 0, 0, 0);
        }
    };
#pragma unroll
    for (int p = 0; p < NST - 1; p++)
        if (c_begin + p < c_end)
            issue(c_begin + p, p);
    int st = 0;
    for (int ch = c_begin; ch < c_end; ch++) {
        // this wave's four loads per chunk: chunk ch has landed once only later chunks' are outstanding
        if (NST > 2 && ch + 1 < c_end)
            wait_vm<(NST > 2 ? 4 : 0)>();
        else
            wait_vm<0>();
        __syncthreads();  // chunk ch is in LDS; the stage of chunk ch - 1 is free again
        if (ch + NST - 1 < c_end)
            issue(ch + NST - 1, st == 0 ? NST - 1 : st - 1);
        if (ch < nch0)
            compute(st, std::integral_constant<int, AM0>{});
        else
            compute(st, std::integral_constant<int, AM1>{});
        st = st == NST - 1 ? 0 : st + 1;
    }

#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < NBY; y++) {
            const int n = n0 + wn * 32 * WIDE + y * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * r;
                if (m < g.M && n < g.N) {
                    if (g.splits == 1) {
                        int64_t row = (int64_t)beta * g.sCb + (int64_t)m * g.sCm;
                        if (g.mfold) {  // row m of the folded batch: row m % mfold of entry m / mfold
                            const int bq = m / g.mfold;
                            row = (int64_t)bq * g.sCb + (int64_t)(m - bq * g.mfold) * g.sCm;
                        }
                        double *cp = g.C + row + (int64_t)n * g.sCn;
                        *cp += g.alpha * acc[x][y][r];
                    } else {
                        g.scratch[((int64_t)split * g.M + m) * g.N + n] = acc[x][y][r];
                    }
                }
            }
        }
}

__global__ __launch_bounds__(256) void gemm_reduce_kernel(GemmArgs g) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)g.M * g.N)
        return;
    const int m = (int)(idx / g.N), n = (int)(idx - (int64_t)m * g.N);
    const int64_t mn = (int64_t)g.M * g.N;
    const double *src = g.scratch + idx;
    double s = 0.0;
    int p = 0;
    for (; p + 8 <= g.splits; p += 8) {  // eight loads in flight, added in split order
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            x[u] = src[(int64_t)(p + u) * mn];
#pragma unroll
        for (int u = 0; u < 8; u++)
            s += x[u];
    }
    for (; p < g.splits; p++)
        s += src[(int64_t)p * mn];
    g.C[(int64_t)m * g.sCm + (int64_t)n * g.sCn] += g.alpha * s;
}

// gemm_reduce_kernel for problem blockIdx.y of a group
__global__ __launch_bounds__(256) void gemm_reduce_group_kernel(GemmGroup gg) {
    const GemmArgs &g = gg.g[blockIdx.y];
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)g.M * g.N)
        return;
    const int m = (int)(idx / g.N), n = (int)(idx - (int64_t)m * g.N);
    const int64_t mn = (int64_t)g.M * g.N;
    const double *src = g.scratch + idx;
    double s = 0.0;
    int p = 0;
    for (; p + 8 <= g.splits; p += 8) {  // eight loads in flight, added in split order
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            x[u] = src[(int64_t)(p + u) * mn];
#pragma unroll
        for (int u = 0; u < 8; u++)
            s += x[u];
    }
    for (; p < g.splits; p++)
        s += src[(int64_t)p * mn];
    g.C[(int64_t)m * g.sCm + (int64_t)n * g.sCn] += g.alpha * s;
}

// ------------------------------------------- GEMM with K-contiguous operands
// The same contraction when every segment has A.s2 == B.s2 == 1 (the inner K index is the
// contiguous one, as the einsum engine lays its operands out) and even strides: rows are read as
// 16-byte vectors (two doubles of K), staged through registers into a double-buffered LDS ring
// with the rows' vectors XOR-permuted (conflict-free 16-lane fragment reads), and each of the 2x2
// waves holds BLK x BLK MFMA 16x16 blocks. K tails and rows past M / N are zero-filled.
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ inline int kswz(int r) { return (r ^ (r >> 3)) & 7; }

typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

__device__ inline short bf16_bits(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float bf16_value(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

// DMA: operand chunks go from global memory straight into LDS (global_load_lds_dwordx4, each lane fetching
// the logical vector its swizzled LDS slot holds), as in the W-build; vectors reaching past K2 load the
// row's first vector instead and are rewritten (value or zero) after the chunk has landed.
// WIDE = 2 (direct-to-LDS only): tiles of BM rows by 2 BM columns, each wave BLK x 2 BLK blocks, so a
// wave issues twice the MFMAs per chunk for 1.5x the staged bytes.
template <int BLK, bool VEC, int PREC, bool DMA, int WIDE>
__global__ __launch_bounds__(256) void gemm_kc_kernel(GemmArgs g) {
    constexpr int BM = 2 * BLK * 16;   // tile rows
    constexpr int BN = BM * WIDE;      // tile columns
    constexpr int BLKN = BLK * WIDE;   // MFMA blocks per wave along n
    constexpr int RB = 128;            // bytes of one 16-double K chunk of a row
    constexpr int NV = BM * 8 / 256;   // 16-byte vectors per thread per operand per chunk (square tiles)
    static_assert(!DMA || (VEC && BM * RB % 4096 == 0), "direct-to-LDS loads need 16-byte vectors, whole KiB");
    static_assert(WIDE == 1 || (WIDE == 2 && DMA), "wide tiles use the direct-to-LDS loads");
    __shared__ __attribute__((aligned(1024))) char lds[2][(BM + BN) * RB];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const int lr = lane & 15, lq = lane >> 4;
    const int ntn = (g.N + BN - 1) / BN;
    // tiles of one (batch, split) on one XCD, so the tiles sharing an operand panel share its L2
    const int nx = gridDim.x, ny = gridDim.y;
    const int lin = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
    const int bx = lin % nx, byz = lin / nx;
    const int tm = bx / ntn, tn = bx - tm * ntn;
    const int m0 = tm * BM, n0 = tn * BN;
    const int beta = byz % ny, split = byz / ny;

    // segment values as scalars: a runtime-indexed local array would live in scratch memory
    const int nk20 = (g.seg[0].K2 + 15) / 16, nk21 = g.nseg > 1 ? (g.seg[1].K2 + 15) / 16 : 1;
    const int nch0 = g.seg[0].K1 * nk20, nch1 = g.nseg > 1 ? g.seg[1].K1 * nk21 : 0;
    const int total = nch0 + nch1;
    const int c_begin = (int)((int64_t)total * split / g.splits), c_end = (int)((int64_t)total * (split + 1) / g.splits);

    d4 acc[BLK][BLKN];
#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLKN; y++)
            acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
    // MFMA blocks of this wave wholly past M or N are skipped (wave-uniform)
    bool vm[BLK], vn[BLKN];
#pragma unroll
    for (int x = 0; x < BLK; x++)
        vm[x] = m0 + wm * BLK * 16 + x * 16 < g.M;
#pragma unroll
    for (int y = 0; y < BLKN; y++)
        vn[y] = n0 + wn * BLKN * 16 + y * 16 < g.N;

    // thread t loads vectors q = t + 256 u (u < NV) of the BM x 8 vector tile: row q / 8, vector q % 8
    d2v ra[NV], rb[NV];
    auto fetch = [&](int ch) {
        const bool s0 = ch < nch0;
        const int cl = s0 ? ch : ch - nch0, nk = s0 ? nk20 : nk21;
        const GemmSeg &sg = s0 ? g.seg[0] : g.seg[1];
        const int k1 = cl / nk, k0 = (cl - k1 * nk) * 16;
        const double *pa = sg.A.base + (int64_t)beta * sg.A.sb + (int64_t)k1 * sg.A.s1;
        const double *pb = sg.B.base + (int64_t)beta * sg.B.sb + (int64_t)k1 * sg.B.s1;
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int q = tid + 256 * u, row = q >> 3, k = k0 + 2 * (q & 7);
            const int gm = m0 + row, gn = n0 + row;
            d2v a = {0.0, 0.0}, b = {0.0, 0.0};
            if (gm < g.M) {
                const double *p = pa + (int64_t)gm * sg.A.sr + k;
                if (VEC && k + 1 < sg.K2) {
                    a = *(const d2v *)p;
                } else {
                    if (k < sg.K2)
                        a.x = p[0];
                    if (k + 1 < sg.K2)
                        a.y = p[1];
                }
            }
            if (gn < g.N) {
                const double *p = pb + (int64_t)gn * sg.B.sr + k;
                if (VEC && k + 1 < sg.K2) {
                    b = *(const d2v *)p;
                } else {
                    if (k < sg.K2)
                        b.x = p[0];
                    if (k + 1 < sg.K2)
                        b.y = p[1];
                }
            }
            ra[u] = a;
            rb[u] = b;
        }
    };
    auto stash = [&](int buf) {
        char *As = lds[buf], *Bs = As + BM * RB;
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int q = tid + 256 * u, row = q >> 3, vec = q & 7;
            *(d2v *)(As + row * RB + ((vec ^ kswz(row)) << 4)) = ra[u];
            *(d2v *)(Bs + row * RB + ((vec ^ kswz(row)) << 4)) = rb[u];
        }
    };

    auto compute = [&](int buf) {
        const char *As = lds[buf], *Bs = As + BM * RB;
        // lane (lr, lq) holds k = 4 lq .. 4 lq + 3 of its row: vectors 2 lq and 2 lq + 1
        if constexpr (PREC == P64) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                d2v af[BLK], bf[BLKN];
#pragma unroll
                for (int b = 0; b < BLK; b++) {
                    const int rowa = wm * BLK * 16 + b * 16 + lr;
                    af[b] = *(const d2v *)(As + rowa * RB + (((2 * lq + h) ^ kswz(rowa)) << 4));
                }
#pragma unroll
                for (int b = 0; b < BLKN; b++) {
                    const int rowb = wn * BLKN * 16 + b * 16 + lr;
                    bf[b] = *(const d2v *)(Bs + rowb * RB + (((2 * lq + h) ^ kswz(rowb)) << 4));
                }
#pragma unroll
                for (int t = 0; t < 2; t++)
#pragma unroll
                    for (int x = 0; x < BLK; x++)
                        if (vm[x])
#pragma unroll
                            for (int y = 0; y < BLKN; y++)
                                if (vn[y])
                                    acc[x][y] =
                                        __builtin_amdgcn_mfma_f64_16x16x4f64(af[x][t], bf[y][t], acc[x][y], 0, 0, 0);
            }
        } else if constexpr (PREC == PB3S) {
            // a 16-slot chunk holds 32 k of pre-split rows: vector 2 lq = the hi bf16 of k = 8 lq .. 8 lq + 7,
            // vector 2 lq + 1 their lo bf16 (the 16x16x32 operand layout); hi*hi + hi*lo + lo*hi in FP32 per
            // chunk, added to the FP64 accumulators
            b8v ah[BLK], al[BLK], bh[BLKN], bl[BLKN];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                const int rowa = wm * BLK * 16 + b * 16 + lr;
                ah[b] = *(const b8v *)(As + rowa * RB + (((2 * lq) ^ kswz(rowa)) << 4));
                al[b] = *(const b8v *)(As + rowa * RB + (((2 * lq + 1) ^ kswz(rowa)) << 4));
            }
#pragma unroll
            for (int b = 0; b < BLKN; b++) {
                const int rowb = wn * BLKN * 16 + b * 16 + lr;
                bh[b] = *(const b8v *)(Bs + rowb * RB + (((2 * lq) ^ kswz(rowb)) << 4));
                bl[b] = *(const b8v *)(Bs + rowb * RB + (((2 * lq + 1) ^ kswz(rowb)) << 4));
            }
#pragma unroll
            for (int x = 0; x < BLK; x++)
#pragma unroll
                for (int y = 0; y < BLKN; y++) {
                    if (!vm[x] || !vn[y])
                        continue;
                    f4v c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], f4v{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], c, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        acc[x][y][r] += (double)c[r];
                }
        } else {
            // the lane's 4 k values of each block row, rounded (FP32) or split (hi + lo BF16)
            f4v a32[BLK], b32[BLKN];
#pragma unroll
            for (int b = 0; b < BLK; b++) {
                const int rowa = wm * BLK * 16 + b * 16 + lr;
                const d2v a0 = *(const d2v *)(As + rowa * RB + (((2 * lq) ^ kswz(rowa)) << 4));
                const d2v a1 = *(const d2v *)(As + rowa * RB + (((2 * lq + 1) ^ kswz(rowa)) << 4));
                a32[b] = f4v{(float)a0.x, (float)a0.y, (float)a1.x, (float)a1.y};
            }
#pragma unroll
            for (int b = 0; b < BLKN; b++) {
                const int rowb = wn * BLKN * 16 + b * 16 + lr;
                const d2v b0 = *(const d2v *)(Bs + rowb * RB + (((2 * lq) ^ kswz(rowb)) << 4));
                const d2v b1 = *(const d2v *)(Bs + rowb * RB + (((2 * lq + 1) ^ kswz(rowb)) << 4));
                b32[b] = f4v{(float)b0.x, (float)b0.y, (float)b1.x, (float)b1.y};
            }
            f4v c32[BLK][BLKN];
#pragma unroll
            for (int x = 0; x < BLK; x++)
#pragma unroll
                for (int y = 0; y < BLKN; y++)
                    c32[x][y] = f4v{0.f, 0.f, 0.f, 0.f};
            if constexpr (PREC == P32) {
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int x = 0; x < BLK; x++)
                        if (vm[x])
#pragma unroll
                            for (int y = 0; y < BLKN; y++)
                                if (vn[y])
                                    c32[x][y] =
                                        __builtin_amdgcn_mfma_f32_16x16x4f32(a32[x][t], b32[y][t], c32[x][y], 0, 0, 0);
            } else {
                s4v ah[BLK], al[BLK], bh[BLKN], bl[BLKN];
#pragma unroll
                for (int b = 0; b < BLK; b++)
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        ah[b][t] = bf16_bits(a32[b][t]);
                        al[b][t] = bf16_bits(a32[b][t] - bf16_value(ah[b][t]));
                    }
#pragma unroll
                for (int b = 0; b < BLKN; b++)
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        bh[b][t] = bf16_bits(b32[b][t]);
                        bl[b][t] = bf16_bits(b32[b][t] - bf16_value(bh[b][t]));
                    }
#pragma unroll
                for (int x = 0; x < BLK; x++)
#pragma unroll
                    for (int y = 0; y < BLKN; y++) {
                        if (!vm[x] || !vn[y])
                            continue;
                        c32[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[x], bh[y], c32[x][y], 0, 0, 0);
                        c32[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[x], bl[y], c32[x][y], 0, 0, 0);
                        c32[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[x], bh[y], c32[x][y], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int x = 0; x < BLK; x++)
#pragma unroll
                for (int y = 0; y < BLKN; y++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        acc[x][y][r] += (double)c32[x][y][r];
        }
    };

    int buf = 0;
    if constexpr (DMA) {
        constexpr int OPB = BM * RB, NIW = OPB / 4096, NIWB = BN * RB / 4096;  // A bytes; wave loads of A, B
        // this lane's DMA units: tile row and the first k (even) of the logical vector its LDS slot holds
        // (units 0 .. NIW-1 of A; the B units are the same for u < NIW and continue the rows past it)
        int lrow[NIWB], lk[NIWB];
#pragma unroll
        for (int u = 0; u < NIWB; u++) {
            const int qq = (wave * NIWB + u) * 64 + lane, r = qq >> 3;
            lrow[u] = r;
            lk[u] = 2 * ((qq & 7) ^ kswz(r));
        }
        int larow[NIW], lak[NIW];
#pragma unroll
        for (int u = 0; u < NIW; u++) {
            const int qq = (wave * NIW + u) * 64 + lane, r = qq >> 3;
            larow[u] = r;
            lak[u] = 2 * ((qq & 7) ^ kswz(r));
        }
        auto chunk_pos = [&](int ch, const GemmSeg *&sg, int &k1, int &k0) {
            const bool s0 = ch < nch0;
            const int cl = s0 ? ch : ch - nch0, nk = s0 ? nk20 : nk21;
            sg = s0 ? &g.seg[0] : &g.seg[1];
            k1 = cl / nk;
            k0 = (cl - k1 * nk) * 16;
        };
        auto issue = [&](int ch, int st) {
            const GemmSeg *sg;
            int k1, k0;
            chunk_pos(ch, sg, k1, k0);
            const double *pa = sg->A.base + (int64_t)beta * sg->A.sb + (int64_t)k1 * sg->A.s1;
            const double *pb = sg->B.base + (int64_t)beta * sg->B.sb + (int64_t)k1 * sg->B.s1;
            char *base = lds[st];
#pragma unroll
            for (int u = 0; u < NIW; u++) {
                const int k = k0 + lak[u], kk = k + 1 < sg->K2 ? k : 0;  // inside the row, else its first vector
                const int gm = min(m0 + larow[u], g.M - 1);
                __builtin_amdgcn_global_load_lds((const void *)(pa + (int64_t)gm * sg->A.sr + kk),
                                                 (lds_void *)(base + (wave * NIW + u) * 1024), 16, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < NIWB; u++) {
                const int k = k0 + lk[u], kk = k + 1 < sg->K2 ? k : 0;
                const int gn = min(n0 + lrow[u], g.N - 1);
                __builtin_amdgcn_global_load_lds((const void *)(pb + (int64_t)gn * sg->B.sr + kk),
                                                 (lds_void *)(base + OPB + (wave * NIWB + u) * 1024), 16, 0, 0);
            }
        };
        // after the chunk has landed: the vectors reaching past K2 get (x[k], 0) or (0, 0)
        auto fixup = [&](int ch, int st) {
            const GemmSeg *sg;
            int k1, k0;
            chunk_pos(ch, sg, k1, k0);
            if (k0 + 16 <= sg->K2)
                return false;
            const double *pa = sg->A.base + (int64_t)beta * sg->A.sb + (int64_t)k1 * sg->A.s1;
            const double *pb = sg->B.base + (int64_t)beta * sg->B.sb + (int64_t)k1 * sg->B.s1;
            char *base = lds[st];
#pragma unroll
            for (int u = 0; u < NIW; u++) {
                const int k = k0 + lak[u];
                if (k + 1 < sg->K2)
                    continue;
                const int gm = min(m0 + larow[u], g.M - 1);
                *(d2v *)(base + (wave * NIW + u) * 1024 + lane * 16) =
                    d2v{k < sg->K2 ? pa[(int64_t)gm * sg->A.sr + k] : 0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < NIWB; u++) {
                const int k = k0 + lk[u];
                if (k + 1 < sg->K2)
                    continue;
                const int gn = min(n0 + lrow[u], g.N - 1);
                *(d2v *)(base + OPB + (wave * NIWB + u) * 1024 + lane * 16) =
                    d2v{k < sg->K2 ? pb[(int64_t)gn * sg->B.sr + k] : 0.0, 0.0};
            }
            return true;  // wave-uniform: every lane of the block sees the same chunk
        };
        if (c_begin < c_end)
            issue(c_begin, 0);
        for (int ch = c_begin; ch < c_end; ch++) {
            wait_vm<0>();
            __syncthreads();  // chunk ch is in LDS; the other stage is free again
            if (fixup(ch, buf))
                __syncthreads();
            if (ch + 1 < c_end)
                issue(ch + 1, buf ^ 1);
            compute(buf);
            buf ^= 1;
        }
    } else {
        if (c_begin < c_end)
            fetch(c_begin);
        for (int ch = c_begin; ch < c_end; ch++) {
            stash(buf);
            __syncthreads();
            if (ch + 1 < c_end)
                fetch(ch + 1);
            compute(buf);
            buf ^= 1;
        }
    }

#pragma unroll
    for (int x = 0; x < BLK; x++)
#pragma unroll
        for (int y = 0; y < BLKN; y++) {
            const int n = n0 + wn * BLKN * 16 + y * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // output layout of the 16x16 MFMAs: row lq + 4 r (FP64), 4 lq + r (FP32 accumulators)
                const int m = m0 + wm * BLK * 16 + x * 16 + (PREC == P64 ? lq + 4 * r : 4 * lq + r);
                if (m < g.M && n < g.N) {
                    if (g.splits == 1) {
                        double *cp = g.C + (int64_t)beta * g.sCb + (int64_t)m * g.sCm + (int64_t)n * g.sCn;
                        *cp += g.alpha * acc[x][y][r];
                    } else {
                        g.scratch[((int64_t)split * g.M + m) * g.N + n] = acc[x][y][r];
                    }
                }
            }
        }
}

// 0: not K-contiguous; 1: K-contiguous; 2: K-contiguous with 16-byte aligned rows (vector loads)
static int k_contiguous(const GemmArgs &g) {
    int r = 2;
    for (int s = 0; s < g.nseg; s++) {
        const GemmSeg &sg = g.seg[s];
        const GemmOp *ops[2] = {&sg.A, &sg.B};
        for (const GemmOp *op : ops) {
            if (op->s2 != 1)
                return 0;
            if ((op->sr & 1) || (op->s1 & 1) || (op->sb & 1) || (((uintptr_t)op->base) & 15))
                r = 1;
        }
    }
    return r;
}

// Paired-load mode of one operand: 1 rows contiguous, 2 K contiguous (both with even strides and a
// 16-byte aligned base), 0 otherwise
static int pair_mode(const GemmOp &op) {
    const bool even = !(op.s1 & 1) && !(op.sb & 1) && !(((uintptr_t)op.base) & 15);
    if (even && op.sr == 1 && !(op.s2 & 1))
        return 1;
    if (even && op.s2 == 1 && !(op.sr & 1))
        return 2;
    return 0;
}

// Whether gemm_kn_kernel applies (FP64 products, B of unit stride in n, A of unit stride in m or in
// k2, every stride and base even so that operand vectors are 16-byte aligned, padded extents known, A
// zero past K2): 0 if not, else 1 + (segment 0's A unit-stride in k2) + 2 (segment 1's).
// SCESS_GEMM_KN=0 keeps the register-staged kernel.
static int kn_modes(const GemmArgs &g) {
    if (g.prec != P64)
        return 0;
    if (const char *e = std::getenv("SCESS_GEMM_KN"))
        if (std::atoi(e) == 0)
            return 0;
    auto even = [](const GemmOp &op) {
        return !(op.sr & 1) || op.sr == 1 ? !(op.s1 & 1) && !(op.sb & 1) && !(((uintptr_t)op.base) & 15) : false;
    };
    int am = 1;
    for (int q = 0; q < g.nseg; q++) {
        const GemmOp &A = g.seg[q].A, &B = g.seg[q].B;
        if (B.sr != 1 || (B.s2 & 1) || B.cext < 2 || (B.cext & 1) || !even(B) || g.seg[q].K2 < 1 || g.seg[q].K1 < 1)
            return 0;
        if (!A.kzero || A.cext < 2 || (A.cext & 1) || !even(A))
            return 0;
        if (A.sr == 1 && !(A.s2 & 1))
            continue;  // [k][m] tile
        if (A.s2 == 1 && !(A.sr & 1) && A.cext >= (g.seg[q].K2 + 15) / 16 * 16) {
            am += q == 0 ? 1 : 2;  // [m][k] tile
            continue;
        }
        return 0;
    }
    if (g.nseg == 1 && am == 2)
        am = 4;  // segment 1 is not used: give it segment 0's mode
    return am;
}

hipError_t launch_gemm(const GemmArgs &g, int batch, hipStream_t s) {
    {  // the kernels index K chunks (16 deep) with 32-bit integers
        int64_t chunks = 0;
        for (int q = 0; q < g.nseg; q++)
            chunks += (int64_t)g.seg[q].K1 * ((g.seg[q].K2 + 15) / 16);
        if (chunks >= (int64_t)1 << 31 || g.nseg < 1 || g.nseg > 2)
            return hipErrorInvalidValue;
    }
    if (const int kc = k_contiguous(g)) {
        // 64x64 tiles: for FP64 MFMA more resident work-groups beat larger register tiles
        int blk = 2;
        if (const char *e = std::getenv("SCESS_GEMM_BLK"))  // tuning override: 2 or 4 (FP64 only)
            blk = std::atoi(e) == 4 && g.prec == P64 ? 4 : 2;
        const int bm = 32 * blk;
        const int nt = ((g.M + bm - 1) / bm) * ((g.N + bm - 1) / bm);
        const dim3 grid((unsigned)nt, (unsigned)batch, (unsigned)g.splits);
        // direct-to-LDS loads when the rows are 16-byte vectors and at least two elements long (the first
        // vector of every row is then inside it); SCESS_GEMM_DMA=0 keeps the register-staged loads
        bool dma = kc == 2 && blk == 2;
        for (int q = 0; q < g.nseg; q++)
            dma = dma && g.seg[q].K2 >= 2 && g.M >= 1 && g.N >= 1;
        if (const char *e = std::getenv("SCESS_GEMM_DMA"))
            dma = dma && std::atoi(e) != 0;
        // 64 x 128 tiles for the direct-to-LDS FP64 and pre-split BF16x3 products: twice the MFMAs per
        // barrier for 1.5x the staged bytes (CCSD update at o=50, v=500: FP64 17.9 s against 19.5 s,
        // BF16x3 6.66 s against 7.47 s with 64 x 64). SCESS_GEMM_WIDE=0 keeps 64 x 64 (tuning).
        const char *ew = std::getenv("SCESS_GEMM_WIDE");
        const bool wide = dma && blk == 2 && (g.prec == P64 || g.prec == PB3S) && !(ew && std::atoi(ew) == 0);
#define SCESS_KC(B, V, P, D)                                                                      \
    if (blk == B && (kc == 2) == V && g.prec == P && dma == D) {                                  \
        hipLaunchKernelGGL((gemm_kc_kernel<B, V, P, D, 1>), grid, dim3(256), 0, s, g);            \
    } else
        if (wide) {  // 64 x 128 tiles
            const dim3 gw((unsigned)(((g.M + 63) / 64) * ((g.N + 127) / 128)), (unsigned)batch, (unsigned)g.splits);
            if (g.prec == PB3S)
                hipLaunchKernelGGL((gemm_kc_kernel<2, true, PB3S, true, 2>), gw, dim3(256), 0, s, g);
            else
                hipLaunchKernelGGL((gemm_kc_kernel<2, true, P64, true, 2>), gw, dim3(256), 0, s, g);
        } else
        SCESS_KC(2, true, P64, true) SCESS_KC(2, true, P32, true) SCESS_KC(2, true, PB3, true)
        SCESS_KC(2, true, PB3S, true) SCESS_KC(2, true, PB3S, false)
        SCESS_KC(2, true, P64, false) SCESS_KC(2, false, P64, false) SCESS_KC(2, true, P32, false)
        SCESS_KC(2, false, P32, false) SCESS_KC(2, true, PB3, false) SCESS_KC(2, false, PB3, false)
        SCESS_KC(4, true, P64, false) SCESS_KC(4, false, P64, false)
        return hipErrorInvalidValue;
#undef SCESS_KC
    } else if (kn_modes(g)) {
        GemmGroup gg;
        gg.g[0] = g;
        gg.n = 1;
        return launch_gemm_group(gg, batch, s);
    } else {
        const int ntiles = ((g.M + GT - 1) / GT) * ((g.N + GT - 1) / GT);
        const dim3 grid((unsigned)ntiles, (unsigned)batch, (unsigned)g.splits);
        int a0 = pair_mode(g.seg[0].A), b0 = pair_mode(g.seg[0].B);
        int a1 = g.nseg > 1 ? pair_mode(g.seg[1].A) : a0, b1 = g.nseg > 1 ? pair_mode(g.seg[1].B) : b0;
        if (!a0 || !a1 || !b0 || !b1 || b0 != b1)
            a0 = a1 = b0 = b1 = 0;  // element-wise loads for every operand
#define SCESS_GP(A0, A1, B)                                                                       \
    if (a0 == A0 && a1 == A1 && b0 == B) {                                                        \
        hipLaunchKernelGGL((gemm_kernel<A0, A1, B, B>), grid, dim3(256), 0, s, g);                \
    } else
        SCESS_GP(1, 1, 1) SCESS_GP(1, 2, 1) SCESS_GP(2, 1, 1) SCESS_GP(2, 2, 1)
        SCESS_GP(1, 1, 2) SCESS_GP(1, 2, 2) SCESS_GP(2, 1, 2) SCESS_GP(2, 2, 2)
        hipLaunchKernelGGL((gemm_kernel<0, 0, 0, 0>), grid, dim3(256), 0, s, g);
#undef SCESS_GP
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || g.splits == 1)
        return e;
    const int64_t n = (int64_t)g.M * g.N;
    hipLaunchKernelGGL(gemm_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g);
    return hipGetLastError();
}

static hipError_t launch_kn_group(const GemmGroup &gg0, int am, int batch, hipStream_t s) {
    // A batch whose B operands do not depend on the batch index (the dX and dTM back-contractions) is
    // folded into the rows: M' = M x batch rows of 64-row tiles leave one partly filled tile instead of
    // one per batch entry (at v = 200 the fourth tile of every entry holds 8 rows). Needs M even (the
    // two rows of a 16-byte A vector then share a batch entry) and one split.
    GemmGroup gg = gg0;
    const char *ef = std::getenv("SCESS_KN_FOLD");  // tuning override: 0 keeps one tile row set per entry
    bool fold = batch > 1 && !(ef && std::atoi(ef) == 0);
    for (int p = 0; p < gg.n && fold; p++) {
        const GemmArgs &q = gg.g[p];
        fold = q.splits == 1 && q.M % 2 == 0 && (int64_t)q.M * batch < ((int64_t)1 << 30);
        for (int h = 0; h < q.nseg && fold; h++)
            fold = q.seg[h].B.sb == 0;
    }
    if (fold) {
        for (int p = 0; p < gg.n; p++) {
            gg.g[p].mfold = gg.g[p].M;
            gg.g[p].M *= batch;
        }
        batch = 1;
    }
    const GemmArgs &g = gg.g[0];
    // am: 1 + (A of segment 0 unit-stride in k) + 2 (A of segment 1 unit-stride in k)
    const char *e3 = std::getenv("SCESS_KN_STAGES");  // tuning override: 2 or 3 LDS stages
    const char *ew = std::getenv("SCESS_KN_WIDE");    // tuning override: 1 for 64 x 128 tiles
    const bool three = e3 && std::atoi(e3) == 3;
    // 64 x 128 tiles where N spans more than one 64-wide tile: slower here (4.07 s per backward at o=20,
    // v=200 against 3.70 s), as fewer resident work-groups cost these short-K GEMMs more than the MFMAs
    // per barrier gain
    const bool wide = !three && g.N > 64 && ew && std::atoi(ew) == 1;
    const int ntiles = ((g.M + GT - 1) / GT) * ((g.N + (wide ? 2 * GT : GT) - 1) / (wide ? 2 * GT : GT));
    const dim3 grid((unsigned)ntiles, (unsigned)batch, (unsigned)(g.splits * gg.n));
    if (wide) {
        switch (am) {
        case 1: hipLaunchKernelGGL((gemm_kn_kernel<0, 0, 2, 2>), grid, dim3(256), 0, s, gg); break;
        case 2: hipLaunchKernelGGL((gemm_kn_kernel<1, 0, 2, 2>), grid, dim3(256), 0, s, gg); break;
        case 3: hipLaunchKernelGGL((gemm_kn_kernel<0, 1, 2, 2>), grid, dim3(256), 0, s, gg); break;
        default: hipLaunchKernelGGL((gemm_kn_kernel<1, 1, 2, 2>), grid, dim3(256), 0, s, gg); break;
        }
    } else if (three) {
        switch (am) {
        case 1: hipLaunchKernelGGL((gemm_kn_kernel<0, 0, 3, 1>), grid, dim3(256), 0, s, gg); break;
        case 2: hipLaunchKernelGGL((gemm_kn_kernel<1, 0, 3, 1>), grid, dim3(256), 0, s, gg); break;
        case 3: hipLaunchKernelGGL((gemm_kn_kernel<0, 1, 3, 1>), grid, dim3(256), 0, s, gg); break;
        default: hipLaunchKernelGGL((gemm_kn_kernel<1, 1, 3, 1>), grid, dim3(256), 0, s, gg); break;
        }
    } else {
        switch (am) {
        case 1: hipLaunchKernelGGL((gemm_kn_kernel<0, 0, 2, 1>), grid, dim3(256), 0, s, gg); break;
        case 2: hipLaunchKernelGGL((gemm_kn_kernel<1, 0, 2, 1>), grid, dim3(256), 0, s, gg); break;
        case 3: hipLaunchKernelGGL((gemm_kn_kernel<0, 1, 2, 1>), grid, dim3(256), 0, s, gg); break;
        default: hipLaunchKernelGGL((gemm_kn_kernel<1, 1, 2, 1>), grid, dim3(256), 0, s, gg); break;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || g.splits == 1)
        return e;
    const int64_t n = (int64_t)g.M * g.N;
    hipLaunchKernelGGL(gemm_reduce_group_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)gg.n), dim3(256), 0, s,
                       gg);
    return hipGetLastError();
}

hipError_t launch_gemm_group(const GemmGroup &gg, int batch, hipStream_t s) {
    if (gg.n < 1 || gg.n > kMaxGemmGroup)
        return hipErrorInvalidValue;
    const GemmArgs &g0 = gg.g[0];
    for (int p = 1; p < gg.n; p++)  // one shape per group
        if (gg.g[p].M != g0.M || gg.g[p].N != g0.N || gg.g[p].splits != g0.splits)
            return hipErrorInvalidValue;
    for (int p = 0; p < gg.n; p++) {
        int64_t chunks = 0;  // the kernels index K chunks (16 deep) with 32-bit integers
        for (int q = 0; q < gg.g[p].nseg; q++)
            chunks += (int64_t)gg.g[p].seg[q].K1 * ((gg.g[p].seg[q].K2 + 15) / 16);
        if (chunks >= (int64_t)1 << 31 || gg.g[p].nseg < 1 || gg.g[p].nseg > 2)
            return hipErrorInvalidValue;
    }
    // one launch per operand layout class (problems keep their relative order; outputs are disjoint)
    bool done[kMaxGemmGroup] = {};
    for (int p = 0; p < gg.n; p++) {
        if (done[p])
            continue;
        const int am = kn_modes(gg.g[p]);
        if (!am) {
            hipError_t e = launch_gemm(gg.g[p], batch, s);  // not a direct-to-LDS layout
            if (e != hipSuccess)
                return e;
            done[p] = true;
            continue;
        }
        GemmGroup sub;
        sub.n = 0;
        for (int q = p; q < gg.n; q++)
            if (!done[q] && kn_modes(gg.g[q]) == am) {
                sub.g[sub.n++] = gg.g[q];
                done[q] = true;
            }
        hipError_t e = launch_kn_group(sub, am, batch, s);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// ------------------------------------------------------------- V backward
// V(a,b,c) = U0[a][b] t_k^c + U1[a][c] t_j^b + U2[b][c] t_i^a (U rows pitched vR, t1 rows vR):
//   dU0[a][b] = sum_c G_V t_k^c,  dU1[a][c] = sum_b G_V t_j^b,  dU2[b][c] = sum_a G_V t_i^a
//   dt_k^c = sum_ab G_V U0[a][b], dt_j^b = sum_ac G_V U1[a][c], dt_i^a = sum_bc G_V U2[b][c]
// sum_z f(z), z = 0..n-1, in order, with eight loads in flight
template <typename F> __device__ inline void nsum8(int n, F f, double &s, double &c) {
    int z = 0;
    for (; z + 8 <= n; z += 8) {
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; q++)
            x[q] = f(z + q);
#pragma unroll
        for (int q = 0; q < 8; q++)
            nadd_a(s, c, x[q]);
    }
    for (; z < n; z++)
        nadd_a(s, c, f(z));
}

// dU (w = 0, 1, 2) and H[a][c] = sum_b G_V[a][b][c] U0[a][b] (w = 3, into H), one thread per output
// element; the threads of a wave take consecutive y, so every G_V read but dU0's is coalesced
__global__ __launch_bounds__(256) void vback_du_kernel(const double *GV, const double *t1, const double *U, int i,
                                                        int j, int k, int v, int vR, double *dU, double *H) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t vv = (int64_t)v * v;
    if (idx >= 4 * vv)
        return;
    const int w = (int)(idx / vv);
    const int64_t xy = idx - w * vv;
    const int x = (int)(xy / v), y = (int)(xy - (int64_t)x * v);
    const int64_t R2 = (int64_t)vR * vR;
    double s = 0.0, c = 0.0;
    if (w == 0) {  // dU0[x=a][y=b]
        const double *g = GV + x * R2 + (int64_t)y * vR;
        const double *t = t1 + (int64_t)k * vR;
        nsum8(v, [&](int z) { return g[z] * t[z]; }, s, c);
    } else if (w == 1) {  // dU1[x=a][y=c]
        const double *t = t1 + (int64_t)j * vR;
        nsum8(v, [&](int z) { return GV[x * R2 + (int64_t)z * vR + y] * t[z]; }, s, c);
    } else if (w == 2) {  // dU2[x=b][y=c]
        const double *t = t1 + (int64_t)i * vR;
        nsum8(v, [&](int z) { return GV[z * R2 + (int64_t)x * vR + y] * t[z]; }, s, c);
    } else {  // H[x=a][y=c]
        const double *u0 = U + (int64_t)x * vR;
        nsum8(v, [&](int z) { return GV[x * R2 + (int64_t)z * vR + y] * u0[z]; }, s, c);
        H[xy] = s + c;
        return;
    }
    dU[idx] = s + c;
}

__global__ __launch_bounds__(256) void vback_dt_kernel(const double *GV, const double *U, const double *H, int v, int vR,
                                                        double *dt) {
    __shared__ double rs[256], rc[256];
    const int w = blockIdx.x / v, x = blockIdx.x - w * v;
    const int64_t R2 = (int64_t)vR * vR, vr = (int64_t)v * vR;
    const int64_t n = w == 0 ? v : (int64_t)v * v;
    double s = 0.0, c = 0.0;
    // terms t = p v + q: dt_j[x=b] = sum_ac G[a][b][c] U1[a][c] (w = 1), dt_i[x=a] = sum_bc G[a][b][c]
    // U2[b][c] (w = 2); dt_k[x=c] = sum_a H[a][c] (w = 0, terms t = a, H from vback_du_kernel). This thread
    // takes t = tid, tid + 256, ... in order, eight loads in flight, (p, q) advanced without divisions
    const double *Uw = U + (int64_t)w * vr;
    int p = (int)(threadIdx.x / v), q = (int)(threadIdx.x % v);
    for (int64_t t0 = threadIdx.x; t0 < n; t0 += 256 * 8) {
        double xs[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            double term = 0.0;  // past the end: adds nothing
            if (t0 + 256 * u < n) {
                if (w == 0)
                    term = H[(t0 + 256 * u) * v + x];
                else
                    term = (w == 1 ? GV[p * R2 + (int64_t)x * vR + q] : GV[x * R2 + (int64_t)p * vR + q]) *
                           Uw[(int64_t)p * vR + q];
            }
            xs[u] = term;
            q += 256;
            while (q >= v) {
                q -= v;
                p++;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            nadd_a(s, c, xs[u]);
    }
    rs[threadIdx.x] = s;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
            const double a1 = rs[threadIdx.x], b1 = rs[threadIdx.x + m];
            const double ss = a1 + b1, bp = ss - a1;
            rc[threadIdx.x] = rc[threadIdx.x] + rc[threadIdx.x + m] + ((a1 - (ss - bp)) + (b1 - bp));
            rs[threadIdx.x] = ss;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        dt[blockIdx.x] = rs[0] + rc[0];
}

// dS[p][q][x][y] = d/d(px|qy): the six updates dS[pq_u] += 1/2 dU[u / 2] (transposed for odd u), in u
// order, and the three dt1 row updates (k from dt[0..v), j from dt[v..2v), i from dt[2v..3v)), in order.
// Targets may coincide (two equal occupied indices); each thread owns one (x, y) of every target, so the
// updates of one element stay in that order and no two threads touch the same element.
__global__ __launch_bounds__(256) void vback_acc_kernel(double *dS, double *dt1, const double *dU, const double *dt,
                                                         int i, int j, int k, int o, int v, int vR) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t vv = (int64_t)v * v;
    if (idx < vv) {
        const int x = (int)(idx / v), y = (int)(idx - (int64_t)x * v);
        const int pq[6][2] = {{i, j}, {j, i}, {i, k}, {k, i}, {j, k}, {k, j}};
#pragma unroll
        for (int u = 0; u < 6; u++) {
            const double *src = dU + (u / 2) * vv;
            dS[((int64_t)pq[u][0] * o + pq[u][1]) * vv + idx] += 0.5 * ((u & 1) ? src[(int64_t)y * v + x] : src[idx]);
        }
    }
    if (idx < v) {
        const int rows[3] = {k, j, i};
#pragma unroll
        for (int u = 0; u < 3; u++)
            dt1[(int64_t)rows[u] * vR + idx] += dt[(int64_t)u * v + idx];
    }
}

hipError_t launch_vback(const double *GV, const double *U, const double *t1, int i, int j, int k, int o, int v, int vR,
                        double *dU, double *dt, double *H, double *dS, double *dt1, hipStream_t s) {
    const int64_t vv = (int64_t)v * v;
    hipLaunchKernelGGL(vback_du_kernel, dim3((unsigned)((4 * vv + 255) / 256)), dim3(256), 0, s, GV, t1, U, i, j, k, v,
                       vR, dU, H);
    hipLaunchKernelGGL(vback_dt_kernel, dim3((unsigned)(3 * v)), dim3(256), 0, s, GV, U, H, v, vR, dt);
    hipLaunchKernelGGL(vback_acc_kernel, dim3((unsigned)((std::max<int64_t>(vv, v) + 255) / 256)), dim3(256), 0, s, dS,
                       dt1, dU, dt, i, j, k, o, v, vR);
    return hipGetLastError();
}

// --------------------------------------------------------------- unpacking
// ColMajor host-layout gradients from the packed FP64 accumulators.
__global__ __launch_bounds__(256) void unpack_kernel(int mode, const double *A, const double *B, double *out, int o_,
                                                     int v_, int64_t vK, int64_t oK, int64_t vR) {
    const int64_t o = o_, v = v_;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t n = 0;
    switch (mode) {
    case 0: n = o * v * v * v; break;  // ovvv(p,x,y,f) = dX[p][x][y][f]
    case 1: n = o * v * o * o; break;  // ovoo(p,x,m,q) = -dRn[p][q][x][m]
    case 2: n = o * v * o * v; break;  // ovov(p,x,q,y) = dS[p][q][x][y]
    case 3: n = o * o * v * v; break;  // t2(q,r,z,f) = dT[q][r][z][f] + dTM[r][z][f][q]
    case 4: n = o * v; break;          // t1(k,c) = dt1[k][c]
    }
    if (gid >= n)
        return;
    double val = 0.0;
    switch (mode) {
    case 0: {
        const int64_t p = gid % o, t = gid / o, x = t % v, t2 = t / v, y = t2 % v, f = t2 / v;
        val = A[((p * v + x) * v + y) * vK + f];
    } break;
    case 1: {
        const int64_t p = gid % o, t = gid / o, x = t % v, t2 = t / v, m = t2 % o, q = t2 / o;
        val = -A[((p * o + q) * v + x) * oK + m];
    } break;
    case 2: {
        const int64_t p = gid % o, t = gid / o, x = t % v, t2 = t / v, q = t2 % o, y = t2 / o;
        val = A[((p * o + q) * v + x) * v + y];
    } break;
    case 3: {
        const int64_t q = gid % o, t = gid / o, r = t % o, t2 = t / o, z = t2 % v, f = t2 / v;
        val = A[((q * o + r) * v + z) * vK + f] + B[((r * v + z) * v + f) * oK + q];
    } break;
    case 4: {
        const int64_t kk = gid % o, c = gid / o;
        val = A[kk * vR + c];
    } break;
    }
    out[gid] = val;
}

hipError_t launch_unpack(int mode, const double *A, const double *B, double *out, int o, int v, int64_t vK, int64_t oK,
                         int64_t vR, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mode, A, B, out, o, v, vK, oK,
                       vR);
    return hipGetLastError();
}

}  // namespace scess
