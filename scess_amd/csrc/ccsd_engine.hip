// Closed-shell CCSD residual engine on gfx950 (FP64).
//
// A device einsum engine with the semantics of the reference's compile-time
// Einsum (src/Tensor/TensorEinsum.h:9-17,137-205): labels shared by both
// operands and absent from the output are summed; labels of the output found
// in both operands are batch indices. Each contraction is evaluated as
// permute(A) -> [batch][M][K], permute(B) -> [batch][N][K], one FP64 MFMA GEMM
// (launch_gemm of triples_adjoint.hip), and a permuting accumulate into the
// output. On top of it, UpdateAmplitudes (src/Quantum/CC/RCCSD.cpp:101-237) is
// transcribed term by term; the particle-particle ladder sum_cd Wvvvv tau never
// forms a v^4 tensor when the Cholesky factor of (ac|bd) is given:
//   sum_cd (ac|bd) tau_ij^cd   from (ac|bd) tiles regenerated from L per (a,b) tile,
//   sum_cd (kd|ac) t_k^b tau_ij^cd and sum_cd (kc|bd) t_k^a tau_ij^cd
//                              through Z[k][a][i][j] = sum_cd (kd|ac) tau_ij^cd.
#include "ccsd_internal.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace scess {

// ------------------------------------------------------------------ kernels
constexpr int MAXD = 8;
struct PermArgs {
    int nd;
    int64_t odims[MAXD];   // output dims (row-major)
    int64_t istride[MAXD]; // input stride of each output dim
    const double *in;
    double *out;
    double alpha, beta;    // out = beta * out + alpha * in[perm]
    int64_t n;
};

// Input offset of an output position: the output multi-index decomposed with 64-bit divisions. The
// loop runs over all MAXD dimensions, unrolled, with the ones past a.nd skipped, so the index lives in
// registers.
struct PermWalk {
    int64_t idx[MAXD];
    int64_t off;
    __device__ void init(const PermArgs &a, int64_t lin) {
        int64_t r = lin;
        off = 0;
#pragma unroll
        for (int d = MAXD - 1; d >= 0; d--) {
            idx[d] = 0;
            if (d >= a.nd)
                continue;
            idx[d] = r % a.odims[d];
            r /= a.odims[d];
            off += idx[d] * a.istride[d];
        }
    }
};

// One output per thread, so that the lanes of a wave store (and, with beta != 0, load) consecutive
// outputs: walking several consecutive outputs per thread saves index divisions but spreads every
// store instruction over 32-byte strides (FP64 update at o=50, v=500: 1.33 s of permutes against 0.70 s).
__global__ __launch_bounds__(256) void permute_kernel(PermArgs a) {
    const int64_t lin = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (lin >= a.n)
        return;
    PermWalk w;
    w.init(a, lin);
    const double x = a.alpha * a.in[w.off];
    a.out[lin] = a.beta == 0.0 ? x : a.beta * a.out[lin] + x;
}

// Permute into pre-split BF16x3 rows: element (row, k) of the permuted operand (a.odims, flattened
// row * K + k) goes to row `row` of `out`, k padded to Kp (a multiple of 32) with zeros, as hi = bf16(x)
// and lo = bf16(x - hi) (round to nearest even) in groups of 8 k: 8 hi then 8 lo (32 bytes per group).
__device__ inline uint16_t bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// one thread per group of 8 k of one row (two 16-byte stores)
__global__ __launch_bounds__(256) void permute_split_kernel(PermArgs a, int64_t K, int64_t Kp, uint16_t *out,
                                                            int64_t nrows) {
    // one element per thread: the lanes of a wave read consecutive k of a row (walking eight k per
    // thread and storing 16-byte vectors spread each load over 64-byte strides: 0.74 s against 0.63 s
    // of split permutes per BF16x3 update at o=50, v=500)
    const int64_t lin = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (lin >= nrows * Kp)
        return;
    const int64_t row = lin / Kp, k = lin - row * Kp;
    double x = 0.0;
    if (k < K) {
        PermWalk w;
        w.init(a, row * K + k);
        x = a.alpha * a.in[w.off];
    }
    const uint16_t hi = bf16_rn((float)x);
    const uint16_t lo = bf16_rn((float)(x - (double)__uint_as_float((uint32_t)hi << 16)));
    uint16_t *o = out + row * 2 * Kp + (k >> 3) * 16 + (k & 7);
    o[0] = hi;
    o[8] = lo;
}

// out = alpha * x + beta * y (elementwise)
__global__ __launch_bounds__(256) void axpby_kernel(double *out, const double *x, const double *y, double alpha,
                                                    double beta, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        out[i] = alpha * x[i] + (y ? beta * y[i] : 0.0);
}

// tau[i][j][a][b] = t2 + t1[i][a] t1[j][b]
__global__ __launch_bounds__(256) void tau_kernel(double *tau, const double *t2, const double *t1, int o, int v) {
    const int64_t n = (int64_t)o * o * v * v;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n)
        return;
    const int b = (int)(idx % v);
    int64_t r = idx / v;
    const int a = (int)(r % v);
    r /= v;
    const int j = (int)(r % o), i = (int)(r / o);
    tau[idx] = t2[idx] + t1[(int64_t)i * v + a] * t1[(int64_t)j * v + b];
}

// t1 /= (e_i - e_a), t2 /= (e_i + e_j - e_a - e_b)   (RCCSD.cpp:235-236, :321-331)
__global__ __launch_bounds__(256) void jacobi_kernel(double *t1, double *t2, const double *e, int o, int v) {
    const int64_t n2 = (int64_t)o * o * v * v, n1 = (int64_t)o * v;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx < n2) {
        const int b = (int)(idx % v);
        int64_t r = idx / v;
        const int a = (int)(r % v);
        r /= v;
        const int j = (int)(r % o), i = (int)(r / o);
        t2[idx] /= (e[i] - e[o + a]) + (e[j] - e[o + b]);
    }
    if (idx < n1) {
        const int a = (int)(idx % v), i = (int)(idx / v);
        t1[idx] /= e[i] - e[o + a];
    }
}

// partial dot products, fixed order: part[block] = sum over the block's range
__global__ __launch_bounds__(256) void dot_kernel(const double *x, const double *y, int64_t n, double *part) {
    __shared__ double rs[256], rc[256];
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(n, b0 + per);
    double s = 0.0, c = 0.0;
    for (int64_t t = b0 + threadIdx.x; t < b1; t += 256) {
        const double p = x[t] * y[t];
        const double u = s + p;
        if (fabs(s) >= fabs(p))
            c += (s - u) + p;
        else
            c += (p - u) + s;
        s = u;
    }
    rs[threadIdx.x] = s;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) {
            const double a1 = rs[threadIdx.x], b1v = rs[threadIdx.x + m];
            const double ss = a1 + b1v, bp = ss - a1;
            rc[threadIdx.x] = rc[threadIdx.x] + rc[threadIdx.x + m] + ((a1 - (ss - bp)) + (b1v - bp));
            rs[threadIdx.x] = ss;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = rs[0] + rc[0];
}

hipError_t ccsd_jacobi(double *t1, double *t2, const double *e, int o, int v, hipStream_t s) {
    const int64_t n = (int64_t)o * o * v * v;
    hipLaunchKernelGGL(jacobi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t1, t2, e, o, v);
    return hipGetLastError();
}

// ------------------------------------------------------------------- engine
DTensor::~DTensor() { release(); }

void DTensor::release() {
    if (p)
        (void)hipFreeAsync(p, s);
    p = nullptr;
}

DTensor::DTensor(DTensor &&o) noexcept : p(o.p), dims(std::move(o.dims)), s(o.s) { o.p = nullptr; }

DTensor &DTensor::operator=(DTensor &&o) noexcept {
    if (this != &o) {
        release();
        p = o.p;
        dims = std::move(o.dims);
        s = o.s;
        o.p = nullptr;
    }
    return *this;
}

int64_t DTensor::size() const {
    int64_t n = 1;
    for (auto d : dims)
        n *= d;
    return n;
}

static void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DTensor Engine::alloc(std::vector<int64_t> dims, bool zero) {
    DTensor t;
    t.dims = std::move(dims);
    t.s = stream;
    const size_t bytes = sizeof(double) * (size_t)std::max<int64_t>(t.size(), 1);
    check(hipMallocAsync((void **)&t.p, bytes, stream), "hipMallocAsync (CCSD tensor)");
    if (zero)
        check(hipMemsetAsync(t.p, 0, bytes, stream), "hipMemsetAsync");
    return t;
}

void Engine::upload(DTensor &t, const double *host) {
    check(hipMemcpyAsync(t.p, host, sizeof(double) * (size_t)t.size(), hipMemcpyHostToDevice, stream), "upload");
}

void Engine::download(const DTensor &t, double *host) {
    check(hipMemcpyAsync(host, t.p, sizeof(double) * (size_t)t.size(), hipMemcpyDeviceToHost, stream), "download");
    check(hipStreamSynchronize(stream), "download sync");
}

void Engine::copy(DTensor &dst, const DTensor &src) {
    check(hipMemcpyAsync(dst.p, src.p, sizeof(double) * (size_t)src.size(), hipMemcpyDeviceToDevice, stream), "copy");
}

void Engine::axpby(double *out, const double *x, const double *y, double alpha, double beta, int64_t n) {
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, out, x, y, alpha, beta, n);
    check(hipGetLastError(), "axpby");
}

// out (labels lo, row-major dims) = beta*out + alpha*in[permuted], in has labels li
static PermArgs perm_args(const std::string &li, const double *in, const std::vector<int64_t> &idims,
                          const std::string &lo, double *out, const std::vector<int64_t> &odims, double alpha,
                          double beta) {
    PermArgs a{};
    a.nd = (int)lo.size();
    if (a.nd > MAXD || (int)li.size() != a.nd)
        throw std::runtime_error("permute: rank");
    std::vector<int64_t> is(li.size());
    int64_t s = 1;
    for (int d = (int)li.size() - 1; d >= 0; d--) {
        is[d] = s;
        s *= idims[d];
    }
    a.n = 1;
    for (int d = 0; d < a.nd; d++) {
        const auto pos = li.find(lo[d]);
        if (pos == std::string::npos)
            throw std::runtime_error("permute: label " + std::string(1, lo[d]) + " missing");
        a.odims[d] = odims[d];
        a.istride[d] = is[pos];
        a.n *= odims[d];
    }
    a.in = in;
    a.out = out;
    a.alpha = alpha;
    a.beta = beta;
    return a;
}

void Engine::permute(const std::string &li, const double *in, const std::vector<int64_t> &idims,
                     const std::string &lo, double *out, const std::vector<int64_t> &odims, double alpha,
                     double beta) {
    const PermArgs a = perm_args(li, in, idims, lo, out, odims, alpha, beta);
    if (a.n == 0)
        return;
    hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, stream, a);
    check(hipGetLastError(), "permute");
}

DTensor Engine::permute_split(const std::string &li, const double *in, const std::vector<int64_t> &idims,
                              const std::string &lo, const std::vector<int64_t> &odims, int64_t K) {
    const PermArgs a = perm_args(li, in, idims, lo, nullptr, odims, 1.0, 0.0);
    if (K < 1 || a.n % K)
        throw std::runtime_error("permute_split: K does not divide the operand");
    const int64_t Kp = (K + 31) / 32 * 32, nrows = a.n / K;
    DTensor out = alloc({nrows, Kp / 2});  // Kp bf16 pairs per row = Kp / 2 eight-byte slots
    if (nrows > 0) {
        hipLaunchKernelGGL(permute_split_kernel, dim3((unsigned)((nrows * Kp + 255) / 256)), dim3(256), 0, stream, a,
                           K, Kp, (uint16_t *)out.p, nrows);
        check(hipGetLastError(), "permute_split");
    }
    return out;
}

static void parse_spec(const std::string &spec, std::string &la, std::string &lb, std::string &lc) {
    const auto arrow = spec.find("->");
    const auto comma = spec.find(',');
    if (arrow == std::string::npos)
        throw std::runtime_error("einsum: no ->");
    lc = spec.substr(arrow + 2);
    if (comma == std::string::npos || comma > arrow) {
        la = spec.substr(0, arrow);
        lb.clear();
    } else {
        la = spec.substr(0, comma);
        lb = spec.substr(comma + 1, arrow - comma - 1);
    }
}

// C = beta*C + alpha * einsum(spec, A, B)
void Engine::einsum(const std::string &spec, const DTensor &A, const DTensor *B, DTensor &C, double alpha,
                    double beta) {
    std::string la, lb, lc;
    parse_spec(spec, la, lb, lc);
    if (la.size() != A.dims.size() || lc.size() != C.dims.size() || (B && lb.size() != B->dims.size()))
        throw std::runtime_error("einsum " + spec + ": rank mismatch");
    auto dim_of = [&](char x) -> int64_t {
        auto p = la.find(x);
        if (p != std::string::npos)
            return A.dims[p];
        p = lb.find(x);
        if (B && p != std::string::npos)
            return B->dims[p];
        p = lc.find(x);
        return C.dims[p];
    };
    if (!B) {  // pure permutation
        permute(la, A.p, A.dims, lc, C.p, C.dims, alpha, beta);
        return;
    }
    std::string Lb, Lm, Ln, Lk;
    for (char x : lc) {
        const bool ina = la.find(x) != std::string::npos, inb = lb.find(x) != std::string::npos;
        if (ina && inb)
            Lb += x;
        else if (ina)
            Lm += x;
        else if (inb)
            Ln += x;
        else
            throw std::runtime_error("einsum " + spec + ": output label not in an operand");
    }
    for (char x : la) {
        const bool inb = lb.find(x) != std::string::npos, inc = lc.find(x) != std::string::npos;
        if (inb && !inc)
            Lk += x;
        else if (!inb && !inc)
            throw std::runtime_error("einsum " + spec + ": trace over a single operand is not supported");
    }
    for (char x : lb)
        if (la.find(x) == std::string::npos && lc.find(x) == std::string::npos)
            throw std::runtime_error("einsum " + spec + ": trace over a single operand is not supported");
    auto prod = [&](const std::string &ls) {
        int64_t n = 1;
        for (char x : ls)
            n *= dim_of(x);
        return n;
    };
    auto dims_of = [&](const std::string &ls) {
        std::vector<int64_t> d;
        for (char x : ls)
            d.push_back(dim_of(x));
        return d;
    };
    const int64_t nb = prod(Lb), M = prod(Lm), N = prod(Ln), K = prod(Lk);
    // operands in [batch][M][K] and [batch][N][K]: FP64, or (BF16x3 products) pre-split rows of
    // Ks = Kp / 2 eight-byte slots, written by the permute itself
    const bool split = prec == PB3 && M > 0 && N > 0 && K > 0;
    const int64_t Ks = split ? (K + 31) / 32 * 16 : K;
    DTensor Ap, Bp, Cp = alloc({nb, M, N}, true);
    if (split) {
        Ap = permute_split(la, A.p, A.dims, Lb + Lm + Lk, dims_of(Lb + Lm + Lk), K);
        Bp = permute_split(lb, B->p, B->dims, Lb + Ln + Lk, dims_of(Lb + Ln + Lk), K);
    } else {
        Ap = alloc({nb, M, K});
        Bp = alloc({nb, N, K});
        permute(la, A.p, A.dims, Lb + Lm + Lk, Ap.p, dims_of(Lb + Lm + Lk), 1.0, 0.0);
        permute(lb, B->p, B->dims, Lb + Ln + Lk, Bp.p, dims_of(Lb + Ln + Lk), 1.0, 0.0);
    }
    if (M > 0 && N > 0 && K > 0) {
        GemmArgs g{};
        g.nseg = 1;
        g.M = (int)M;
        g.N = (int)N;
        g.alpha = 1.0;
        g.splits = 1;
        g.seg[0].K1 = 1;
        g.seg[0].K2 = (int)Ks;
        g.seg[0].A = {Ap.p, Ks, 0, 1, M * Ks};
        g.seg[0].B = {Bp.p, Ks, 0, 1, N * Ks};
        g.C = Cp.p;
        g.sCm = N;
        g.sCn = 1;
        g.sCb = M * N;
        g.prec = split ? PB3S : prec;
        if (M > 0x7fffffff || N > 0x7fffffff || K > 0x7fffffff || nb > 65535)
            throw std::runtime_error("einsum " + spec + ": extent too large");
        // few output tiles over a long K (e.g. Foo, the T1 terms): split K so that ~512 work-groups
        // run; the partial tiles are added in split order (deterministic)
        const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
        DTensor scr;
        if (nb == 1 && tiles < 256 && K >= 8192) {
            g.splits = (int)std::min<int64_t>({512 / tiles, K / 2048, 256});
            if (g.splits > 1) {
                scr = alloc({g.splits, M, N});
                g.scratch = scr.p;
            } else {
                g.splits = 1;
            }
        }
        check(launch_gemm(g, (int)nb, stream), "einsum gemm");
    }
    permute(Lb + Lm + Ln, Cp.p, dims_of(Lb + Lm + Ln), lc, C.p, C.dims, alpha, beta);
}

double Engine::dot(const double *x, const double *y, int64_t n) {
    const int nblk = 256;
    if (!dotbuf)
        check(hipMalloc((void **)&dotbuf, sizeof(double) * nblk), "hipMalloc dot");
    hipLaunchKernelGGL(dot_kernel, dim3(nblk), dim3(256), 0, stream, x, y, n, dotbuf);
    check(hipGetLastError(), "dot");
    double h[nblk];
    check(hipMemcpyAsync(h, dotbuf, sizeof(h), hipMemcpyDeviceToHost, stream), "dot copy");
    check(hipStreamSynchronize(stream), "dot sync");
    double s = 0.0, c = 0.0;
    for (int i = 0; i < nblk; i++) {
        const double u = s + h[i];
        if (std::fabs(s) >= std::fabs(h[i]))
            c += (s - u) + h[i];
        else
            c += (h[i] - u) + s;
        s = u;
    }
    return s + c;
}

Engine::~Engine() {
    if (dotbuf)
        (void)hipFree(dotbuf);
}

// ------------------------------------------------------------- CCSD update
void CcsdState::make_tau(Engine &E, const DTensor &t1, const DTensor &t2, DTensor &tau) {
    const int64_t n = t2.size();
    hipLaunchKernelGGL(tau_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E.stream, tau.p, t2.p, t1.p, o, v);
    check(hipGetLastError(), "tau");
}

// R += X_ij^ab + X_ji^ba (scaled)
static void plus_transpose_add(Engine &E, DTensor &R, const DTensor &X, double alpha) {
    E.permute("ijab", X.p, X.dims, "ijab", R.p, R.dims, alpha, 1.0);
    E.permute("jiba", X.p, X.dims, "ijab", R.p, R.dims, alpha, 1.0);
}

// Particle-particle ladder sum_cd (ac|bd) tau_ij^cd from the Cholesky factor, (a,b) tiles of TAB x TAB
void CcsdState::ladder_from_factor(Engine &E, const DTensor &tau, DTensor &R) {
    // (a, b) tile edge: 64 fills the GEMM's 64-row tiles; smaller when one tile of integrals
    // (TAB^2 v^2 doubles) would pass 8 GiB
    int64_t TAB = 64;
    while (TAB > 16 && TAB * TAB * v * v * 8 > ((int64_t)8 << 30))
        TAB /= 2;
    const int64_t nP = naux;
    DTensor I = E.alloc({TAB, v, TAB, v});  // I[a][c][b][d] of one tile
    // BF16x3 products: the factor (rows (a, c), K = P), tau (rows (i, j, c), K = d) and each tile of I
    // (rows (a, c, b), K = d) in pre-split rows of Ps / Vs eight-byte slots
    const bool split = E.prec == PB3;
    const int64_t Ps = (nP + 31) / 32 * 16, Vs = (v + 31) / 32 * 16;
    DTensor Ls, taus;
    if (split) {
        Ls = E.permute_split("acP", Lvv.p, {v, v, nP}, "acP", {v, v, nP}, nP);
        taus = E.permute_split("ijcd", tau.p, {o, o, v, v}, "ijcd", {o, o, v, v}, v);
    }
    for (int64_t a0 = 0; a0 < v; a0 += TAB) {
        const int64_t ta = std::min(TAB, v - a0);
        for (int64_t b0 = 0; b0 < v; b0 += TAB) {
            const int64_t tb = std::min(TAB, v - b0);
            // I[(a,c)][(b,d)] = sum_P L[a][c][P] L[b][d][P]
            GemmArgs g{};
            g.nseg = 1;
            g.M = (int)(ta * v);
            g.N = (int)(tb * v);
            g.alpha = 1.0;
            g.splits = 1;
            g.seg[0].K1 = 1;
            if (split) {
                g.seg[0].K2 = (int)Ps;
                g.seg[0].A = {Ls.p + a0 * v * Ps, Ps, 0, 1, 0};
                g.seg[0].B = {Ls.p + b0 * v * Ps, Ps, 0, 1, 0};
            } else {
                g.seg[0].K2 = (int)nP;
                g.seg[0].A = {Lvv.p + a0 * v * nP, nP, 0, 1, 0};  // rows (a, c), K = P contiguous
                g.seg[0].B = {Lvv.p + b0 * v * nP, nP, 0, 1, 0};
            }
            check(hipMemsetAsync(I.p, 0, sizeof(double) * (size_t)I.size(), E.stream), "memset I");
            g.C = I.p;
            g.prec = split ? PB3S : E.prec;
            g.sCm = tb * v;
            g.sCn = 1;
            g.sCb = 0;
            check(launch_gemm(g, 1, E.stream), "ladder integrals");
            // R[i][j][a][b] += sum_{c,d} I[a][c][b][d] tau[i][j][c][d]; batch a, rows b, K = (c, d)
            GemmArgs h{};
            h.nseg = 1;
            h.M = (int)tb;
            h.N = (int)(o * o);
            h.alpha = 1.0;
            h.splits = 1;
            h.seg[0].K1 = (int)v;
            DTensor Is;
            if (split) {  // I of this tile in pre-split rows (a, c, b), K = d
                Is = E.permute_split("acbd", I.p, {ta, v, tb, v}, "acbd", {ta, v, tb, v}, v);
                h.seg[0].K2 = (int)Vs;
                h.seg[0].A = {Is.p, Vs, tb * Vs, 1, v * tb * Vs};
                h.seg[0].B = {taus.p, v * Vs, Vs, 1, 0};
            } else {
                h.seg[0].K2 = (int)v;
                h.seg[0].A = {I.p, v, tb * v, 1, v * tb * v};
                h.seg[0].B = {tau.p, v * v, v, 1, 0};
            }
            h.prec = split ? PB3S : E.prec;
            h.C = R.p + a0 * v + b0;
            h.sCb = v;
            h.sCm = 1;
            h.sCn = v * v;
            check(launch_gemm(h, (int)ta, E.stream), "ladder contraction");
        }
    }
}

void CcsdState::update(Engine &E, DTensor &t1, DTensor &t2) {
    const int64_t O = o, Vv = v;
    DTensor tau = E.alloc({O, O, Vv, Vv});
    make_tau(E, t1, t2, tau);
    DTensor Foo = E.alloc({O, O}, true), Fvv = E.alloc({Vv, Vv}, true), Fov = E.alloc({O, Vv}, true);
    E.einsum("kcld,ilcd->ki", L_ovov, &tau, Foo, 1.0, 0.0);
    E.einsum("kcld,klad->ac", L_ovov, &tau, Fvv, -1.0, 0.0);
    E.einsum("kcld,ld->kc", L_ovov, &t1, Fov, 1.0, 0.0);

    // -- T1 equation (RCCSD.cpp:117-138)
    DTensor t1n = E.alloc({O, Vv}, true);
    E.einsum("ac,ic->ia", Fvv, &t1, t1n, 1.0, 1.0);
    E.einsum("ki,ka->ia", Foo, &t1, t1n, -1.0, 1.0);
    E.einsum("kc,kica->ia", Fov, &t2, t1n, 2.0, 1.0);
    E.einsum("kc,ikca->ia", Fov, &t2, t1n, -1.0, 1.0);
    {
        DTensor Fov_t1 = E.alloc({O, O}, true);
        E.einsum("kc,ic->ki", Fov, &t1, Fov_t1, 1.0, 0.0);
        E.einsum("ki,ka->ia", Fov_t1, &t1, t1n, 1.0, 1.0);
    }
    E.einsum("kcai,kc->ia", ovvo, &t1, t1n, 2.0, 1.0);
    E.einsum("kiac,kc->ia", oovv, &t1, t1n, -1.0, 1.0);
    E.einsum("kdac,ikcd->ia", ovvv, &tau, t1n, 2.0, 1.0);
    E.einsum("kcad,ikcd->ia", ovvv, &tau, t1n, -1.0, 1.0);
    E.einsum("lcki,klac->ia", ovoo, &tau, t1n, -2.0, 1.0);
    E.einsum("kcli,klac->ia", ovoo, &tau, t1n, 1.0, 1.0);

    // -- T2 equation
    DTensor t2n = E.alloc({O, O, Vv, Vv}, true);
    {
        // tmp2(a,b,i,c) = -kibc,ka + (ia|cb); abic,jc->ijab, without forming the o v^3 tmp2:
        //   sum_c (ia|cb) t_j^c  and  -sum_k t_k^a sum_c (ki|bc) t_j^c
        DTensor X = E.alloc({O, O, Vv, Vv}, true);
        E.einsum("iacb,jc->ijab", ovvv, &t1, X, 1.0, 0.0);
        DTensor M = E.alloc({O, O, Vv, O}, true);  // M[k][i][b][j] = sum_c (ki|bc) t_j^c
        E.einsum("kibc,jc->kibj", oovv, &t1, M, 1.0, 0.0);
        E.einsum("kibj,ka->ijab", M, &t1, X, -1.0, 1.0);
        plus_transpose_add(E, t2n, X, 1.0);
    }
    {
        // tmp2(a,k,i,j) = kcai,jc + (ia|jk); akij,kb->ijab
        DTensor tmp = E.alloc({Vv, O, O, O}, true);
        E.einsum("kcai,jc->akij", ovvo, &t1, tmp, 1.0, 0.0);
        E.einsum("iajk->akij", ovoo, nullptr, tmp, 1.0, 1.0);
        DTensor X = E.alloc({O, O, Vv, Vv}, true);
        E.einsum("akij,kb->ijab", tmp, &t1, X, 1.0, 0.0);
        plus_transpose_add(E, t2n, X, -1.0);
    }
    E.axpby(t2n.p, t2n.p, oovv_phys.p, 1.0, 1.0, t2n.size());
    if (t2driver.p)
        E.axpby(t2n.p, t2n.p, t2driver.p, 1.0, 1.0, t2n.size());

    DTensor Loo = E.alloc({O, O}, true), Lvv = E.alloc({Vv, Vv}, true);
    E.copy(Loo, Foo);
    E.einsum("lcki,lc->ki", ovoo, &t1, Loo, 2.0, 1.0);
    E.einsum("kcli,lc->ki", ovoo, &t1, Loo, -1.0, 1.0);
    E.copy(Lvv, Fvv);
    E.einsum("kdac,kd->ac", ovvv, &t1, Lvv, 2.0, 1.0);
    E.einsum("kcad,kd->ac", ovvv, &t1, Lvv, -1.0, 1.0);

    {
        DTensor Woooo = E.alloc({O, O, O, O});
        E.copy(Woooo, oooo_phys);
        E.einsum("lcki,jc->klij", ovoo, &t1, Woooo, 1.0, 1.0);
        E.einsum("kclj,ic->klij", ovoo, &t1, Woooo, 1.0, 1.0);
        E.einsum("kcld,ijcd->klij", ovov, &tau, Woooo, 1.0, 1.0);
        E.einsum("klij,klab->ijab", Woooo, &tau, t2n, 1.0, 1.0);
    }
    // particle-particle ladder (RCCSD.cpp:177-180, :215-216)
    if (vvvv_phys.p) {
        DTensor Wvvvv = E.alloc({Vv, Vv, Vv, Vv});
        E.copy(Wvvvv, vvvv_phys);
        E.einsum("kdac,kb->abcd", ovvv, &t1, Wvvvv, -1.0, 1.0);
        E.einsum("kcbd,ka->abcd", ovvv, &t1, Wvvvv, -1.0, 1.0);
        E.einsum("abcd,ijcd->ijab", Wvvvv, &tau, t2n, 1.0, 1.0);
    } else {
        ladder_from_factor(E, tau, t2n);
        DTensor Z = E.alloc({O, Vv, O, O}, true);  // Z[k][a][i][j] = sum_cd (kd|ac) tau_ij^cd
        E.einsum("kdac,ijcd->kaij", ovvv, &tau, Z, 1.0, 0.0);
        E.einsum("kaij,kb->ijab", Z, &t1, t2n, -1.0, 1.0);
        // sum_cd (kc|bd) tau_ij^cd = Z[k][b][j][i] (tau_ij^cd = tau_ji^dc)
        E.einsum("kbji,ka->ijab", Z, &t1, t2n, -1.0, 1.0);
    }
    {
        DTensor X = E.alloc({O, O, Vv, Vv}, true);
        E.einsum("ac,ijcb->ijab", Lvv, &t2, X, 1.0, 0.0);
        plus_transpose_add(E, t2n, X, 1.0);
        E.einsum("ki,kjab->ijab", Loo, &t2, X, 1.0, 0.0);
        plus_transpose_add(E, t2n, X, -1.0);
    }

    // ring and exchange terms (RCCSD.cpp:182-211, :221-233)
    DTensor Wvoov = E.alloc({Vv, O, O, Vv}), Wvovo = E.alloc({Vv, O, Vv, O});
    E.einsum("kcai->akic", ovvo, nullptr, Wvoov, 1.0, 0.0);
    E.einsum("kcad,id->akic", ovvv, &t1, Wvoov, 1.0, 1.0);
    E.einsum("kcli,la->akic", ovoo, &t1, Wvoov, -1.0, 1.0);
    E.einsum("ldkc,ilda->akic", ovov, &t2, Wvoov, -0.5, 1.0);
    E.einsum("lckd,ilad->akic", ovov, &t2, Wvoov, -0.5, 1.0);
    E.einsum("ldkc,ilad->akic", ovov, &t2, Wvoov, 1.0, 1.0);
    {
        DTensor tmp = E.alloc({O, O, Vv, O}, true);
        E.einsum("ldkc,id->lkci", ovov, &t1, tmp, 1.0, 0.0);
        E.einsum("lkci,la->akic", tmp, &t1, Wvoov, -1.0, 1.0);
    }
    E.einsum("kiac->akci", oovv, nullptr, Wvovo, 1.0, 0.0);
    E.einsum("kdac,id->akci", ovvv, &t1, Wvovo, 1.0, 1.0);
    E.einsum("lcki,la->akci", ovoo, &t1, Wvovo, -1.0, 1.0);
    E.einsum("lckd,ilda->akci", ovov, &t2, Wvovo, -0.5, 1.0);
    {
        DTensor tmp = E.alloc({O, Vv, O, O}, true);
        E.einsum("lckd,id->lcki", ovov, &t1, tmp, 1.0, 0.0);
        E.einsum("lcki,la->akci", tmp, &t1, Wvovo, -1.0, 1.0);
    }
    {
        DTensor X = E.alloc({O, O, Vv, Vv}, true);
        E.einsum("akic,kjcb->ijab", Wvoov, &t2, X, 2.0, 0.0);
        E.einsum("akci,kjcb->ijab", Wvovo, &t2, X, -1.0, 1.0);
        plus_transpose_add(E, t2n, X, 1.0);
        E.einsum("akic,kjbc->ijab", Wvoov, &t2, X, 1.0, 0.0);
        plus_transpose_add(E, t2n, X, -1.0);
        E.einsum("bkci,kjac->ijab", Wvovo, &t2, X, 1.0, 0.0);
        plus_transpose_add(E, t2n, X, -1.0);
    }

    // Jacobi update with FP64 denominators (RCCSD.cpp:235-236)
    const int64_t n = t2n.size();
    hipLaunchKernelGGL(jacobi_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E.stream, t1n.p, t2n.p, e.p, o,
                       v);
    check(hipGetLastError(), "jacobi");
    t1 = std::move(t1n);
    t2 = std::move(t2n);
}

double CcsdState::energy(Engine &E, const DTensor &t1, const DTensor &t2) {
    // sum_ijab (2 <ij|ab> - <ij|ba>) tau_ij^ab (RCCSD.cpp:85-89) (+ F12 coupling :91-98)
    DTensor tau = E.alloc({o, o, v, v});
    make_tau(E, t1, t2, tau);
    double Ecorr = E.dot(L_oovv_phys.p, tau.p, tau.size());
    if (f12energy.p) {
        DTensor Le = E.alloc({o, o, v, v});
        E.permute("ijab", f12energy.p, f12energy.dims, "ijab", Le.p, Le.dims, 2.0, 0.0);
        E.permute("ijba", f12energy.p, f12energy.dims, "ijab", Le.p, Le.dims, -1.0, 1.0);
        Ecorr += E.dot(Le.p, t2.p, t2.size());
    }
    return Ecorr;
}

}  // namespace scess
