// C ABI of libscess_gpu.so (include/scess_gpu.h): context, resident inputs and
// the (T) driver that schedules the gfx950 kernels of triples_kernels.hip.
#include "../../include/scess_gpu.h"
#include "ccsd_internal.h"
#include "triples_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace scess;

struct TriplesState {
    bool have = false;
    int o = 0, v = 0, vR = 0;
    bool have_prec[kNumPrec] = {false, false, false, false};
    Operands ops[kNumPrec];
    double *S = nullptr, *t1 = nullptr, *e = nullptr, *ev = nullptr;  // FP64 epilogue inputs
    std::vector<int> ijk_all;                          // 3 per unique occupied triple
    std::vector<int> work;                             // indices (into ijk_all) with i != k
    std::vector<int> work_ijk;                         // 3 per work item, as evaluated (j == k -> (j, j, i))
    int *d_ijk_all = nullptr;
    std::vector<int> orbits;
    int *d_orbits = nullptr;
    uint64_t *d_orbtab = nullptr;  // orbit_table entries
    int nb = 0, norb = 0, ngrp = 0;  // blocks per dim, orbits, epilogue partials per ijk
    // work buffers
    void *W1 = nullptr, *W2 = nullptr;
    size_t wbytes = 0;
    int *d_work_ijk = nullptr, *d_work_index = nullptr;  // per work item (i != k)
    double *partials = nullptr;
    size_t partials_cap = 0;
    double *U = nullptr;  // V pieces of one batch
    size_t u_cap = 0;
    double *F = nullptr, *E = nullptr;
};

struct scess_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // (T) epilogue stream, overlapping the next batch's W-build
    std::string err;
    int (*should_stop)(void *) = nullptr;
    void *user = nullptr;
    int64_t nbatch_req = 0;
    int64_t shard = 0, nshards = 1;  // (T) evaluates work items [n*shard/nshards, n*(shard+1)/nshards)
    int ccsd_prec = 0;               // product precision of the CCSD residual contractions
    void *scratch = nullptr;         // this device's address of the static scratch array
    int int8_digits = kI8DefaultDigits;  // digits per element of SCESS_PREC_OZAKI (next scess_t_set_inputs)
    TriplesState t;
    scess::Engine eng;
    scess::CcsdState cc;
};

namespace {

// Direct-to-LDS W-build kernel unless SCESS_WBUILD_DMA=0 (the register-staged kernel, for A/B timing)
bool wbuild_dma() {
    const char *e = std::getenv("SCESS_WBUILD_DMA");
    return !(e && e[0] == '0');
}


int fail(scess_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    return code;
}

int hip_fail(scess_ctx *c, hipError_t e, const char *what) {
    const int code = (e == hipErrorOutOfMemory) ? SCESS_ERR_OOM : SCESS_ERR_HIP;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(c, expr)                                  \
    do {                                                  \
        hipError_t e__ = (expr);                          \
        if (e__ != hipSuccess)                            \
            return hip_fail((c), e__, #expr);             \
    } while (0)

template <class T>
void dfree(T *&p) {
    if (p) {
        (void)hipFree((void *)p);
        p = nullptr;
    }
}

void free_triples(TriplesState &t) {
    for (auto &op : t.ops) {
        dfree(op.X);
        dfree(op.T);
        dfree(op.Rn);
        dfree(op.TM);
        op = Operands{};
    }
    dfree(t.S);
    dfree(t.t1);
    dfree(t.e);
    dfree(t.ev);
    dfree(t.d_ijk_all);
    dfree(t.d_orbits);
    dfree(t.d_orbtab);
    dfree(t.W1);
    dfree(t.W2);
    dfree(t.d_work_ijk);
    dfree(t.d_work_index);
    dfree(t.partials);
    dfree(t.U);
    dfree(t.F);
    dfree(t.E);
    t = TriplesState{};
}

int upload(scess_ctx *c, const double *h, size_t n, double **d) {
    HIP_TRY(c, hipMalloc((void **)d, sizeof(double) * std::max<size_t>(n, 1)));
    HIP_TRY(c, hipMemcpyAsync(*d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return SCESS_OK;
}

// Device scratch of scess_t_forward_noalloc: W of both families, V pieces and epilogue partials of one
// occupied triple, F_ijk and E(T). Every GPU the library runs on holds its own copy until the process
// ends. Host code never uses the name `scratch` as an address: each device's address comes from
// hipGetSymbolAddress, taken once per device in scess_create with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved once per device.
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

int scess_create(scess_ctx **out, int device) {
    if (!out)
        return SCESS_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return SCESS_ERR_HIP;
    if (device < 0 || device >= n)
        return SCESS_ERR_ARG;
    auto *c = new scess_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return SCESS_ERR_HIP;
    }
    if (hipStreamCreate(&c->stream2) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return SCESS_ERR_HIP;
    }
    c->eng.stream = c->stream;
    if (scratch_address(device, &c->scratch) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->stream2);
        delete c;
        return SCESS_ERR_HIP;
    }
    *out = c;
    return SCESS_OK;
}

void scess_destroy(scess_ctx *c) {
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_triples(c->t);
    c->cc = scess::CcsdState{};
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->stream2);
    delete c;
}

const char *scess_last_error(const scess_ctx *c) { return c ? c->err.c_str() : "null context"; }

int scess_set_should_stop(scess_ctx *c, int (*fn)(void *), void *user) {
    if (!c)
        return SCESS_ERR_ARG;
    c->should_stop = fn;
    c->user = user;
    return SCESS_OK;
}

int scess_set_batch(scess_ctx *c, int64_t nbatch) {
    if (!c || nbatch < 0)
        return SCESS_ERR_ARG;
    c->nbatch_req = nbatch;
    return SCESS_OK;
}

int scess_t_set_shard(scess_ctx *c, int64_t shard, int64_t nshards) {
    if (!c)
        return SCESS_ERR_ARG;
    if (nshards < 1 || shard < 0 || shard >= nshards)
        return fail(c, SCESS_ERR_ARG, "scess_t_set_shard: need 0 <= shard < nshards");
    c->shard = shard;
    c->nshards = nshards;
    return SCESS_OK;
}

int scess_t_set_int8_digits(scess_ctx *c, int digits) {
    if (!c)
        return SCESS_ERR_ARG;
    if (digits < kI8MinDigits || digits > kI8MaxDigits)
        return fail(c, SCESS_ERR_ARG, "scess_t_set_int8_digits: digits must be 5, 6 or 7");
    c->int8_digits = digits;
    return SCESS_OK;
}

int64_t scess_t_num_ijk(const scess_ctx *c) { return c && c->t.have ? (int64_t)c->t.ijk_all.size() / 3 : 0; }

int64_t scess_t_num_partials(const scess_ctx *c) { return c && c->t.have ? c->t.ngrp : 0; }

int scess_t_set_inputs(scess_ctx *c, const scess_t_in *in, int prec_mask) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!in || in->o < 1 || in->v < 1 || !in->e || !in->t1 || !in->t2 || !in->ovvv || !in->ovoo || !in->ovov)
        return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: null array or non-positive size");
    if (in->o > 4096 || in->v > 4096)
        return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: o and v must be <= 4096");
    if ((prec_mask & 15) == 0 || (prec_mask & ~15))
        return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: prec_mask must be a non-empty subset of bits 0..3");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_triples(c->t);
    TriplesState &t = c->t;
    const int64_t o = in->o, v = in->v;
    t.o = (int)o;
    t.v = (int)v;
    t.vR = (int)round_up(v, 16);

    // raw ColMajor inputs (staging)
    double *r_ovvv = nullptr, *r_ovoo = nullptr, *r_ovov = nullptr, *r_t2 = nullptr;
    int rc = SCESS_OK;
    auto cleanup = [&]() {
        dfree(r_ovvv);
        dfree(r_ovoo);
        dfree(r_ovov);
        dfree(r_t2);
    };
    if ((rc = upload(c, in->ovvv, (size_t)(o * v * v * v), &r_ovvv)) ||
        (rc = upload(c, in->ovoo, (size_t)(o * v * o * o), &r_ovoo)) ||
        (rc = upload(c, in->ovov, (size_t)(o * v * o * v), &r_ovov)) ||
        (rc = upload(c, in->t2, (size_t)(o * o * v * v), &r_t2))) {
        cleanup();
        free_triples(t);
        return rc;
    }
    // t1 as [o][vR] (ColMajor (i,a) -> row-major i, zero-padded rows) and the virtual energies padded to vR
    const int64_t vRp = t.vR;
    std::vector<double> t1r((size_t)(o * vRp), 0.0), evr((size_t)vRp, 0.0);
    for (int64_t i = 0; i < o; i++)
        for (int64_t a = 0; a < v; a++)
            t1r[(size_t)(i * vRp + a)] = in->t1[i + o * a];
    for (int64_t a = 0; a < v; a++)
        evr[(size_t)a] = in->e[o + a];
    if ((rc = upload(c, t1r.data(), t1r.size(), &t.t1)) || (rc = upload(c, in->e, (size_t)(o + v), &t.e)) ||
        (rc = upload(c, evr.data(), evr.size(), &t.ev))) {
        cleanup();
        free_triples(t);
        return rc;
    }
    auto pack = [&](int mode, int prec, const double *src, void **dst, int64_t rows, int64_t kpad) -> int {
        const size_t bytes = (size_t)rows * (size_t)kpad * (prec == P64 ? 8 : 4);
        hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
        if (e != hipSuccess)
            return hip_fail(c, e, "hipMalloc (packed operand)");
        e = launch_pack(mode, prec, src, *dst, (int)o, (int)v, kpad, c->stream);
        if (e != hipSuccess)
            return hip_fail(c, e, "pack kernel");
        return SCESS_OK;
    };
    // PI8: one power-of-two scale per tensor (|x| 2^-e <= 0.498), then the digit slices
    auto pack_i8 = [&](Operands &op, const double *ovvv, const double *t2, const double *ovoo) -> int {
        unsigned long long *dmax = nullptr;
        hipError_t e = hipMalloc((void **)&dmax, 3 * sizeof(unsigned long long));
        if (e != hipSuccess)
            return hip_fail(c, e, "hipMalloc");
        unsigned long long hmax[3] = {0, 0, 0};
        e = hipMemsetAsync(dmax, 0, sizeof(hmax), c->stream);
        if (e == hipSuccess)
            e = launch_absmax(ovvv, o * v * v * v, dmax, c->stream);
        if (e == hipSuccess)
            e = launch_absmax(t2, o * o * v * v, dmax + 1, c->stream);
        if (e == hipSuccess)
            e = launch_absmax(ovoo, o * v * o * o, dmax + 2, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hmax, dmax, sizeof(hmax), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        (void)hipFree(dmax);
        if (e != hipSuccess)
            return hip_fail(c, e, "PI8 scales");
        int ex[3];
        for (int n = 0; n < 3; n++) {
            double mx;
            std::memcpy(&mx, &hmax[n], sizeof(mx));
            if (!std::isfinite(mx))
                return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: non-finite input");
            // the smallest e with mx 2^-e <= 0.498 (mx = f 2^E, f in [1/2, 1): e = E + 1 unless f > 0.996)
            int E = 0;
            const double f = std::frexp(mx, &E);
            ex[n] = mx > 0.0 ? (0.5 * f <= 0.498 ? E + 1 : E + 2) : 0;
        }
        // one scale for both integral blocks, so that the f- and m-term digit products of a W element
        // share their int32 groups (the smaller block loses its log2 of the max ratio in digits)
        op.eX = op.eR = std::max(ex[0], ex[2]);
        op.eT = ex[1];
        // int32 groups: |digit product| <= 2^14, at most `slices` pairs per group, 4 vK + 4 oK deep
        if (16384.0 * op.slices * (4.0 * (double)op.vK + 4.0 * (double)op.oK) >= 2147483647.0)
            return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: too large for the int8 mode");
        struct Job {
            int mode;
            const double *src;
            void **dst;
            int64_t rows, kpad;
            int gran, e;
        } jobs[4] = {{PACK_X, ovvv, &op.X, o * v * v, op.vK, 32, op.eX},
                     {PACK_T, t2, &op.T, o * o * v, op.vK, 32, op.eT},
                     {PACK_RN, ovoo, &op.Rn, o * o * v, op.oK, 32, op.eR},
                     {PACK_TM, t2, &op.TM, o * v * v, op.oK, 32, op.eT}};
        for (const Job &jb : jobs) {
            const size_t bytes = (size_t)jb.rows * (size_t)jb.kpad * (size_t)op.slices;
            e = hipMalloc(jb.dst, std::max<size_t>(bytes, 16));
            if (e != hipSuccess)
                return hip_fail(c, e, "hipMalloc (packed operand)");
            e = launch_pack_i8(jb.mode, jb.src, *jb.dst, (int)o, (int)v, jb.kpad, jb.gran, std::ldexp(1.0, -jb.e),
                               op.slices, c->stream);
            if (e != hipSuccess)
                return hip_fail(c, e, "pack_i8 kernel");
        }
        return SCESS_OK;
    };
    if ((rc = pack(PACK_S, P64, r_ovov, (void **)&t.S, o * o * v, v))) {
        cleanup();
        free_triples(t);
        return rc;
    }
    for (int p = 0; p < kNumPrec; p++) {
        if (!(prec_mask & (1 << p)))
            continue;
        Operands &op = t.ops[p];
        op.slices = p == PI8 ? c->int8_digits : 0;
        op.vK = round_up(v, kChunk[p]);
        op.oK = round_up(o, kChunkO[p]);
        op.rbv = op.vK * kElemBytes[p] * (p == PI8 ? op.slices : 1);
        op.rbo = op.oK * kElemBytes[p] * (p == PI8 ? op.slices : 1);
        if (p == PI8) {
            if ((rc = pack_i8(op, r_ovvv, r_t2, r_ovoo))) {
                cleanup();
                free_triples(t);
                return rc;
            }
            t.have_prec[p] = true;
            continue;
        }
        if ((rc = pack(PACK_X, p, r_ovvv, &op.X, o * v * v, op.vK)) ||
            (rc = pack(PACK_T, p, r_t2, &op.T, o * o * v, op.vK)) ||
            (rc = pack(PACK_RN, p, r_ovoo, &op.Rn, o * o * v, op.oK)) ||
            (rc = pack(PACK_TM, p, r_t2, &op.TM, o * v * v, op.oK))) {
            cleanup();
            free_triples(t);
            return rc;
        }
        t.have_prec[p] = true;
    }
    hipError_t se = hipStreamSynchronize(c->stream);
    cleanup();
    if (se != hipSuccess) {
        free_triples(t);
        return hip_fail(c, se, "scess_t_set_inputs");
    }

    // unique occupied triples i >= j >= k (RCCSD_T.cpp enumerates a >= b >= c the same way)
    for (int i = 0; i < o; i++)
        for (int j = 0; j <= i; j++)
            for (int k = 0; k <= j; k++) {
                if (i != k)
                    t.work.push_back((int)(t.ijk_all.size() / 3));
                t.ijk_all.push_back(i);
                t.ijk_all.push_back(j);
                t.ijk_all.push_back(k);
            }
    // virtual block orbits A >= B >= C over 8-wide blocks
    t.nb = t.vR / 8;
    for (int A = 0; A < t.nb; A++)
        for (int B = 0; B <= A; B++)
            for (int C = 0; C <= B; C++) {
                t.orbits.push_back(A);
                t.orbits.push_back(B);
                t.orbits.push_back(C);
            }
    t.norb = (int)(t.orbits.size() / 3);
    t.ngrp = epilogue_groups(t.norb);
    {
        std::vector<uint64_t> tab;
        if (!orbit_table(t.orbits, tab)) {  // unreachable for v <= 4096 (nb <= 512)
            free_triples(t);
            return fail(c, SCESS_ERR_ARG, "scess_t_set_inputs: too many virtual blocks for the orbit table");
        }
        HIP_TRY(c, hipMalloc((void **)&t.d_orbtab, sizeof(uint64_t) * tab.size()));
        HIP_TRY(c, hipMemcpy(t.d_orbtab, tab.data(), sizeof(uint64_t) * tab.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMalloc((void **)&t.d_ijk_all, sizeof(int) * t.ijk_all.size()));
    HIP_TRY(c, hipMemcpy(t.d_ijk_all, t.ijk_all.data(), sizeof(int) * t.ijk_all.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMalloc((void **)&t.d_orbits, sizeof(int) * t.orbits.size()));
    HIP_TRY(c, hipMemcpy(t.d_orbits, t.orbits.data(), sizeof(int) * t.orbits.size(), hipMemcpyHostToDevice));
    {
        // F_ijk is invariant under permutations of (i, j, k), so a triple with j == k is evaluated
        // as (j, j, i): every degenerate triple then has i == j, the form the halved W-build takes
        std::vector<int> wijk;
        for (int w : t.work) {
            const int i = t.ijk_all[3 * w], j = t.ijk_all[3 * w + 1], k = t.ijk_all[3 * w + 2];
            if (j == k && i != j) {
                wijk.push_back(j);
                wijk.push_back(j);
                wijk.push_back(i);
            } else {
                wijk.push_back(i);
                wijk.push_back(j);
                wijk.push_back(k);
            }
        }
        t.work_ijk = wijk;
        const size_t nw = std::max<size_t>(t.work.size(), 1);
        HIP_TRY(c, hipMalloc((void **)&t.d_work_ijk, sizeof(int) * 3 * nw));
        HIP_TRY(c, hipMalloc((void **)&t.d_work_index, sizeof(int) * nw));
        if (!t.work.empty()) {
            HIP_TRY(c, hipMemcpy(t.d_work_ijk, wijk.data(), sizeof(int) * wijk.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(t.d_work_index, t.work.data(), sizeof(int) * t.work.size(), hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(c, hipMalloc((void **)&t.F, sizeof(double) * (t.ijk_all.size() / 3)));
    HIP_TRY(c, hipMalloc((void **)&t.E, sizeof(double)));
    t.have = true;
    return SCESS_OK;
}

// W-build arguments of one precision (the batch's ijk and W pointers are set per launch)
static WBuildArgs wbuild_args(const TriplesState &t, int prec, int blk, int64_t wstride) {
    const Operands &op = t.ops[prec];
    const int64_t v = t.v;
    const int BM = 2 * blk * 16;
    WBuildArgs wa{};
    wa.X = (const char *)op.X;
    wa.T = (const char *)op.T;
    wa.Rn = (const char *)op.Rn;
    wa.TM = (const char *)op.TM;
    wa.o = t.o;
    wa.v = t.v;
    wa.vR = t.vR;
    wa.rbv = op.rbv;
    wa.rbo = op.rbo;
    wa.nchv = (int)(op.vK / kChunk[prec]);
    wa.ncho = (int)(op.oK / kChunkO[prec]);
    wa.nt = (int)((v + BM - 1) / BM);
    wa.ntile1 = (int)(v * wa.nt * wa.nt);
    wa.wstride = wstride;
    wa.halve = 1;
    wa.scale = std::ldexp(1.0, op.eX + op.eT);
    wa.slices = op.slices;
    return wa;
}

int scess_t_forward(scess_ctx *c, int prec, double *e_t, scess_t_stats *stats) {
    return scess_t_forward_ex(c, prec, e_t, nullptr, nullptr, stats);
}

int scess_t_forward_ex(scess_ctx *c, int prec, double *e_t, double *F_ijk, double *partials_out,
                       scess_t_stats *stats) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!e_t)
        return fail(c, SCESS_ERR_ARG, "scess_t_forward: e_t is null");
    *e_t = 0.0;
    if (prec < 0 || prec >= kNumPrec)
        return fail(c, SCESS_ERR_ARG, "scess_t_forward: unknown precision");
    TriplesState &t = c->t;
    if (!t.have)
        return fail(c, SCESS_ERR_STATE, "scess_t_forward: no inputs (call scess_t_set_inputs first)");
    if (!t.have_prec[prec])
        return fail(c, SCESS_ERR_STATE, "scess_t_forward: precision not packed by scess_t_set_inputs");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t o = t.o, v = t.v, vR = t.vR;
    const int64_t nall = (int64_t)t.ijk_all.size() / 3;
    const int64_t nwork = (int64_t)t.work.size();
    const int64_t w_begin = nwork * c->shard / c->nshards, w_end = nwork * (c->shard + 1) / c->nshards;
    const int64_t nshare = w_end - w_begin;
    const size_t accb = prec == P64 || prec == PI8 ? 8 : 4;
    const int64_t wstride = vR * vR * vR;  // W in 8x8x8 blocks over vR^3

    // batch size: a multiple of 8 occupied triples per launch (8 for 8-byte W, 16 for 4-byte W), so that
    // the XCD-ordered work-groups give each of the 8 XCDs whole triples (measured: the FP64 W-build at
    // o=50, v=500 is 27% faster than with one triple per launch, at the same HBM bytes per triple);
    // fewer when the W sets (2 arrays x 2 sets x nbatch x vR^3) exceed a quarter of the free device memory
    int64_t nbatch = c->nbatch_req;
    if (nbatch <= 0) {
        const double per = 4.0 * (double)wstride * (double)accb;
        size_t fr = 0, tot = 0;
        const double budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? 0.25 * (double)fr : 4.0e9;
        nbatch = accb == 4 ? 16 : 8;
        while (nbatch > 1 && nbatch * per > budget)
            nbatch /= 2;
    }
    nbatch = std::max<int64_t>(1, std::min<int64_t>(nbatch, std::max<int64_t>(nshare, 1)));
    if (nbatch > 65535)
        nbatch = 65535;
    // two sets of W, U and partials: the epilogue of batch b reads set b % 2 while the W-build of
    // batch b + 1 writes the other one
    const size_t wneed = 2 * (size_t)nbatch * (size_t)wstride * accb;
    if (t.wbytes < wneed) {
        dfree(t.W1);
        dfree(t.W2);
        t.wbytes = 0;
        HIP_TRY(c, hipMalloc(&t.W1, wneed));
        HIP_TRY(c, hipMalloc(&t.W2, wneed));
        t.wbytes = wneed;
    }
    // per-orbit partials: whole job if the caller wants them, else one batch
    const size_t pneed = (size_t)(partials_out ? std::max<int64_t>(nwork, 1) : 2 * nbatch) * (size_t)t.ngrp;
    if (t.partials_cap < pneed) {
        dfree(t.partials);
        HIP_TRY(c, hipMalloc((void **)&t.partials, sizeof(double) * pneed));
        t.partials_cap = pneed;
    }
    const size_t uneed = 2 * (size_t)nbatch * 3 * (size_t)(v * vR);
    if (t.u_cap < uneed) {
        dfree(t.U);
        HIP_TRY(c, hipMalloc((void **)&t.U, sizeof(double) * uneed));
        t.u_cap = uneed;
    }
    HIP_TRY(c, hipMemsetAsync(t.F, 0, sizeof(double) * nall, c->stream));
    const size_t wset = (size_t)nbatch * (size_t)wstride * accb;  // bytes of one set of W1 (or W2)

    const Operands &op = t.ops[prec];
    int blk = wbuild_tile(prec, (int)v);
    if (const char *ov = std::getenv("SCESS_WBUILD_BLK")) {  // tuning override: 1, 2, 4 or 5
        const int b = std::atoi(ov);
        if (prec != PI8 && (b == 1 || b == 2 || b == 4 || b == 5))
            blk = b;
    }
    WBuildArgs wa = wbuild_args(t, prec, blk, wstride);
    wa.W1 = t.W1;
    wa.W2 = t.W2;

    EpiArgs ea;
    ea.W1 = t.W1;
    ea.W2 = t.W2;
    ea.wstride = wstride;
    ea.o = (int)o;
    ea.v = (int)v;
    ea.vR = (int)vR;
    ea.nb = t.nb;
    ea.orbits = t.d_orbits;
    ea.orbtab = t.d_orbtab;
    ea.norb = t.norb;
    ea.ngrp = t.ngrp;
    ea.t1 = t.t1;
    ea.e = t.e;
    ea.ev = t.ev;
    ea.U = t.U;
    ea.halve = 1;

    std::vector<hipEvent_t> ev;
    auto new_event = [&](hipEvent_t *e) -> hipError_t {
        hipError_t r = hipEventCreate(e);
        if (r == hipSuccess)
            ev.push_back(*e);
        return r;
    };
    auto destroy_events = [&]() {
        for (auto e : ev)
            (void)hipEventDestroy(e);
        ev.clear();
    };
    hipEvent_t e_begin, e_end;
    if (new_event(&e_begin) != hipSuccess || new_event(&e_end) != hipSuccess) {
        destroy_events();
        return fail(c, SCESS_ERR_HIP, "hipEventCreate");
    }
    // a, b: W-build of the batch on stream (its HIP-event duration); e, c: its V pieces, epilogue and
    // partial reduction on stream2, which waits for b; the W-build two batches later waits for c
    struct BatchEvents {
        hipEvent_t a, b, e, c;
    };
    std::vector<BatchEvents> bev;
    hipStream_t s0 = c->stream, s1 = c->stream2;
    hipEvent_t e_ready, e_join;
    if (new_event(&e_ready) != hipSuccess || new_event(&e_join) != hipSuccess) {
        destroy_events();
        return fail(c, SCESS_ERR_HIP, "hipEventCreate");
    }
    (void)hipEventRecord(e_begin, s0);
    (void)hipEventRecord(e_ready, s0);  // F cleared
    (void)hipStreamWaitEvent(s1, e_ready, 0);

    int rc = SCESS_OK;
    bool cancelled = false;
    for (int64_t b0 = w_begin; b0 < w_end; b0 += nbatch) {
        const size_t k = bev.size(), set = k & 1;
        // keep at most two batches in flight so cancellation is polled while the device works
        if (k >= 2 && hipEventSynchronize(bev[k - 2].c) != hipSuccess) {
            rc = fail(c, SCESS_ERR_HIP, "hipEventSynchronize");
            break;
        }
        if (c->should_stop && c->should_stop(c->user)) {
            cancelled = true;
            break;
        }
        const int nb = (int)std::min<int64_t>(nbatch, w_end - b0);
        wa.ijk = t.d_work_ijk + 3 * b0;
        ea.ijk = t.d_work_ijk + 3 * b0;
        wa.W1 = (char *)t.W1 + set * wset;
        wa.W2 = (char *)t.W2 + set * wset;
        ea.W1 = wa.W1;
        ea.W2 = wa.W2;
        double *U = t.U + set * (size_t)nbatch * 3 * (size_t)(v * vR);
        ea.U = U;
        hipError_t he = hipSuccess;
        BatchEvents be{};
        if (new_event(&be.a) != hipSuccess || new_event(&be.b) != hipSuccess || new_event(&be.e) != hipSuccess ||
            new_event(&be.c) != hipSuccess) {
            rc = fail(c, SCESS_ERR_HIP, "hipEventCreate");
            break;
        }
        double *pbase = t.partials + (partials_out ? (size_t)b0 * t.ngrp : set * (size_t)nbatch * t.ngrp);
        ea.partials = pbase;
        if (k >= 2)
            he = hipStreamWaitEvent(s0, bev[k - 2].c, 0);  // this W set is no longer read
        if (he == hipSuccess)
            he = hipEventRecord(be.a, s0);
        if (he == hipSuccess)
            he = launch_wbuild(prec, blk, wbuild_dma(), wa, nb, s0);
        if (he == hipSuccess)
            he = hipEventRecord(be.b, s0);
        if (he == hipSuccess)
            he = hipStreamWaitEvent(s1, be.b, 0);
        if (he == hipSuccess)
            he = hipEventRecord(be.e, s1);
        if (he == hipSuccess)
            he = launch_vpieces(t.S, ea.ijk, (int)o, (int)v, (int)vR, U, nb, s1);
        if (he == hipSuccess)
            he = launch_epilogue(prec, ea, nb, s1, wa.slices);
        if (he == hipSuccess)
            he = launch_reduce_partials(pbase, t.ngrp, t.d_work_index + b0, t.F, nb, s1);
        if (he == hipSuccess)
            he = hipEventRecord(be.c, s1);
        if (he != hipSuccess) {
            rc = hip_fail(c, he, "kernel launch");
            break;
        }
        bev.push_back(be);
    }
    if (rc == SCESS_OK && !cancelled) {
        hipError_t he = launch_reduce_energy(t.F, t.d_ijk_all, (int)nall, t.E, s1);
        if (he != hipSuccess)
            rc = hip_fail(c, he, "reduce_energy launch");
    }
    (void)hipEventRecord(e_join, s1);
    (void)hipStreamWaitEvent(s0, e_join, 0);
    (void)hipEventRecord(e_end, s0);
    hipError_t se = hipStreamSynchronize(s0);
    if (se == hipSuccess)
        se = hipStreamSynchronize(s1);
    if (rc == SCESS_OK && se != hipSuccess)
        rc = hip_fail(c, se, "scess_t_forward");
    if (rc == SCESS_OK && cancelled)
        rc = fail(c, SCESS_CANCELLED, "scess_t_forward: cancelled");
    if (rc == SCESS_OK) {
        double E = 0.0;
        hipError_t he = hipMemcpy(&E, t.E, sizeof(double), hipMemcpyDeviceToHost);
        if (he == hipSuccess && F_ijk)
            he = hipMemcpy(F_ijk, t.F, sizeof(double) * nall, hipMemcpyDeviceToHost);
        if (he == hipSuccess && partials_out) {
            // scatter per-work-item partials to the full unique-ijk enumeration (zeros for i == j == k)
            std::memset(partials_out, 0, sizeof(double) * (size_t)nall * t.ngrp);
            std::vector<double> hp((size_t)nwork * t.ngrp);
            he = hipMemcpy(hp.data(), t.partials, sizeof(double) * hp.size(), hipMemcpyDeviceToHost);
            for (int64_t w = w_begin; he == hipSuccess && w < w_end; w++)
                std::memcpy(partials_out + (size_t)t.work[(size_t)w] * t.ngrp, hp.data() + (size_t)w * t.ngrp,
                            sizeof(double) * t.ngrp);
        }
        if (he != hipSuccess)
            rc = hip_fail(c, he, "result copy");
        else
            *e_t = E;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e_begin, e_end) == hipSuccess)
            stats->seconds = ms * 1e-3;
        for (auto &be : bev) {
            float m1 = 0.0f, m2 = 0.0f;
            if (hipEventElapsedTime(&m1, be.a, be.b) == hipSuccess)
                stats->w_build_ms += m1;
            if (hipEventElapsedTime(&m2, be.e, be.c) == hipSuccess)
                stats->epilogue_ms += m2;
        }
        stats->w_build_launches = (int64_t)bev.size();
        stats->epilogue_launches = (int64_t)bev.size();
        stats->n_ijk = nshare;
        // the shard's share of the o^3 ordered occupied triples: 6 per distinct, 3 per two-equal
        double ordered = 0.0, ordered_all = 0.0;
        int64_t ndeg = 0;  // triples with two equal indices build half of W
        for (int64_t w = 0; w < nwork; w++) {
            const int wi = t.work[(size_t)w];
            const bool deg = t.ijk_all[3 * wi] == t.ijk_all[3 * wi + 1] || t.ijk_all[3 * wi + 1] == t.ijk_all[3 * wi + 2];
            ordered_all += deg ? 3.0 : 6.0;
            if (w >= w_begin && w < w_end) {
                ordered += deg ? 3.0 : 6.0;
                ndeg += deg;
            }
        }
        stats->flops_algorithmic = 2.0 * (double)o * o * o * (double)v * v * v * (double)(o + v) *
                                   (ordered_all > 0.0 ? ordered / ordered_all : 1.0);
        // rows and columns as the kernels issue them: MFMA blocks of 16 (blocks wholly past v skipped);
        // the int8-digit kernel computes every 32-row block of its 64 x 64 tiles
        const double rows = (double)round_up(v, prec == PI8 ? 64 : 16);
        stats->flops_executed = ((double)nshare - 0.5 * (double)ndeg) * 2.0 * (double)v * rows * rows *
                                (6.0 * (double)op.vK + 6.0 * (double)op.oK);
    }
    destroy_events();
    return rc;
}


int scess_t_forward_noalloc(scess_ctx *c, int prec, double *e_t) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!e_t)
        return fail(c, SCESS_ERR_ARG, "scess_t_forward_noalloc: e_t is null");
    *e_t = 0.0;
    if (prec < 0 || prec >= kNumPrec)
        return fail(c, SCESS_ERR_ARG, "scess_t_forward_noalloc: unknown precision");
    TriplesState &t = c->t;
    if (!t.have)
        return fail(c, SCESS_ERR_STATE, "scess_t_forward_noalloc: no inputs (call scess_t_set_inputs first)");
    if (!t.have_prec[prec])
        return fail(c, SCESS_ERR_STATE, "scess_t_forward_noalloc: precision not packed by scess_t_set_inputs");
    if (!c->scratch)
        return fail(c, SCESS_ERR_STATE, "scess_t_forward_noalloc: no scratch address for this device");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t o = t.o, v = t.v, vR = t.vR;
    const int64_t nall = (int64_t)t.ijk_all.size() / 3;
    const int64_t nwork = (int64_t)t.work.size();
    const int64_t w_begin = nwork * c->shard / c->nshards, w_end = nwork * (c->shard + 1) / c->nshards;
    const size_t accb = prec == P64 || prec == PI8 ? 8 : 4;
    const int64_t wstride = vR * vR * vR;
    // scratch layout (256-byte aligned pieces): W1, W2 [vR^3 accb], U [3 v vR], partials [ngrp], F [nall], E
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t wb = al((size_t)wstride * accb), ub = al(sizeof(double) * 3 * (size_t)(v * vR)),
                 pb = al(sizeof(double) * (size_t)t.ngrp), fb = al(sizeof(double) * (size_t)nall);
    const size_t need = 2 * wb + ub + pb + fb + 256;
    if (need > kNoAllocScratchBytes)
        return fail(c, SCESS_ERR_ARG,
                    "scess_t_forward_noalloc: o=" + std::to_string(o) + ", v=" + std::to_string(v) + " needs " +
                        std::to_string(need) + " bytes of scratch, more than its " +
                        std::to_string(kNoAllocScratchBytes) + " (use scess_t_forward)");
    char *base = (char *)c->scratch;
    void *W1 = base, *W2 = base + wb;
    double *U = (double *)(base + 2 * wb), *partials = (double *)(base + 2 * wb + ub),
           *F = (double *)(base + 2 * wb + ub + pb), *E = (double *)(base + 2 * wb + ub + pb + fb);

    const int blk = wbuild_tile(prec, (int)v);
    WBuildArgs wa = wbuild_args(t, prec, blk, wstride);
    wa.W1 = W1;
    wa.W2 = W2;
    EpiArgs ea{};
    ea.W1 = W1;
    ea.W2 = W2;
    ea.wstride = wstride;
    ea.o = (int)o;
    ea.v = (int)v;
    ea.vR = (int)vR;
    ea.nb = t.nb;
    ea.orbits = t.d_orbits;
    ea.orbtab = t.d_orbtab;
    ea.norb = t.norb;
    ea.ngrp = t.ngrp;
    ea.t1 = t.t1;
    ea.e = t.e;
    ea.ev = t.ev;
    ea.U = U;
    ea.partials = partials;
    ea.halve = 1;

    hipStream_t s = c->stream;
    hipError_t he = launch_fill_zero(F, nall, s);
    bool cancelled = false;
    for (int64_t w = w_begin; w < w_end && he == hipSuccess; w++) {
        if (c->should_stop && ((w - w_begin) & 63) == 0) {
            if ((he = hipStreamSynchronize(s)) != hipSuccess)
                break;
            if (c->should_stop(c->user)) {
                cancelled = true;
                break;
            }
        }
        wa.ijk = t.d_work_ijk + 3 * w;
        ea.ijk = wa.ijk;
        he = launch_wbuild(prec, blk, wbuild_dma(), wa, 1, s);
        if (he == hipSuccess)
            he = launch_vpieces(t.S, ea.ijk, (int)o, (int)v, (int)vR, U, 1, s);
        if (he == hipSuccess)
            he = launch_epilogue(prec, ea, 1, s, wa.slices);
        if (he == hipSuccess)
            he = launch_reduce_partials(partials, t.ngrp, t.d_work_index + w, F, 1, s);
    }
    if (he == hipSuccess && !cancelled)
        he = launch_reduce_energy(F, t.d_ijk_all, (int)nall, E, s);
    if (he == hipSuccess)
        he = hipStreamSynchronize(s);
    if (he != hipSuccess)
        return hip_fail(c, he, "scess_t_forward_noalloc");
    if (cancelled)
        return fail(c, SCESS_CANCELLED, "scess_t_forward_noalloc: cancelled");
    double Eh = 0.0;
    HIP_TRY(c, hipMemcpy(&Eh, E, sizeof(double), hipMemcpyDeviceToHost));
    *e_t = Eh;
    return SCESS_OK;
}

int scess_t_backward(scess_ctx *c, int prec, double dE, scess_t_grad *grad) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!grad || !grad->ovvv || !grad->ovoo || !grad->ovov || !grad->t1 || !grad->t2)
        return fail(c, SCESS_ERR_ARG, "scess_t_backward: null gradient array");
    if (prec < 0 || prec >= kNumPrec)
        return fail(c, SCESS_ERR_ARG, "scess_t_backward: unknown precision");
    TriplesState &t = c->t;
    if (!t.have)
        return fail(c, SCESS_ERR_STATE, "scess_t_backward: no inputs (call scess_t_set_inputs first)");
    if (!t.have_prec[prec] || !t.have_prec[P64])
        return fail(c, SCESS_ERR_STATE, "scess_t_backward: needs SCESS_PREC_FP64 (and prec) packed");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t o = t.o, v = t.v, vR = t.vR;
    const Operands &op = t.ops[prec];
    const Operands &o64 = t.ops[P64];
    const int64_t vK = o64.vK, oK = o64.oK;
    const size_t accb = prec == P64 || prec == PI8 ? 8 : 4;
    const int64_t wstride = vR * vR * vR;

    // device buffers of this call, freed on every exit
    std::vector<void *> bufs;
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (e == hipSuccess) {
            bufs.push_back(*p);
            e = hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 16), c->stream);
        }
        return e;
    };
    // the W of the next batch of triples is built on a second stream into the other of two W sets,
    // while the G kernels and back-contractions of the current batch run on c->stream
    hipStream_t ws = nullptr;
    hipEvent_t evW[2] = {nullptr, nullptr}, evDone[2] = {nullptr, nullptr};
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        if (ws)
            (void)hipStreamSynchronize(ws);
        for (void *p : bufs)
            (void)hipFree(p);
        bufs.clear();
        for (int q = 0; q < 2; q++) {
            if (evW[q])
                (void)hipEventDestroy(evW[q]);
            if (evDone[q])
                (void)hipEventDestroy(evDone[q]);
        }
        if (ws)
            (void)hipStreamDestroy(ws);
    };
    // K splits of dT (v x v outputs, K = v^2) and dRn (v x o outputs, K = v^2): about 1024 work-groups per
    // GEMM so that the long K loops overlap their memory latency across the chip (the terms of a triple
    // then run together, up to six GEMMs per launch). SCESS_ADJ_SPLITS = PT,PR overrides (tuning).
    const int64_t tilesT = ((v + 63) / 64) * ((v + 63) / 64), tilesR = ((v + 63) / 64) * ((o + 63) / 64);
    int PT = (int)std::min<int64_t>(64, std::max<int64_t>(16, 1024 / tilesT));
    int PR = (int)std::min<int64_t>(128, std::max<int64_t>(32, 1024 / tilesR));
    if (const char *e = std::getenv("SCESS_ADJ_SPLITS")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%d,%d", &a, &b) == 2 && a >= 1 && a <= 256 && b >= 1 && b <= 256) {
            PT = a;
            PR = b;
        }
    }
    double *dX, *dT, *dRn, *dTM, *dS, *dt1, *GW, *GV, *U, *dU, *dtv, *scr, *out;
    void *W1[2], *W2[2];
    // W of up to WB triples is built by one launch (2 buffers of WB x vR^3 within 4 GB per set)
    const int64_t WB = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t)(4.0e9 / (2.0 * (double)wstride * accb))));
    hipError_t he = hipSuccess;
    const size_t nX = (size_t)(o * v * v * vK), nT = (size_t)(o * o * v * vK), nR = (size_t)(o * o * v * oK),
                 nTM = (size_t)(o * v * v * oK), nS = (size_t)(o * o * v * v), nt1 = (size_t)(o * vR),
                 nG = (size_t)wstride, nscr = (size_t)kMaxGemmGroup * std::max<int64_t>(PT * v * v, PR * v * o);
    // staging of each unpacked gradient before its download: the largest of ovvv (o v^3), ovoo (o^3 v),
    // ovov and t2 (o^2 v^2) and t1 (o v)
    const size_t nout = std::max<size_t>({(size_t)(o * v * v * v), (size_t)(o * v * o * o), (size_t)(o * o * v * v),
                                          (size_t)(o * v)});
    if ((he = alloc(8 * nX, (void **)&dX)) != hipSuccess || (he = alloc(8 * nT, (void **)&dT)) != hipSuccess ||
        (he = alloc(8 * nR, (void **)&dRn)) != hipSuccess || (he = alloc(8 * nTM, (void **)&dTM)) != hipSuccess ||
        (he = alloc(8 * nS, (void **)&dS)) != hipSuccess || (he = alloc(8 * nt1, (void **)&dt1)) != hipSuccess ||
        (he = alloc(8 * nG, (void **)&GW)) != hipSuccess || (he = alloc(8 * nG, (void **)&GV)) != hipSuccess ||
        (he = alloc(8 * 3 * (size_t)(v * vR), (void **)&U)) != hipSuccess ||
        (he = alloc(8 * 3 * (size_t)(v * v), (void **)&dU)) != hipSuccess ||
        (he = alloc(8 * 3 * (size_t)v, (void **)&dtv)) != hipSuccess ||
        (he = alloc(8 * nscr, (void **)&scr)) != hipSuccess || (he = alloc(8 * nout, (void **)&out)) != hipSuccess ||
        (he = alloc(accb * (size_t)(WB * wstride), &W1[0])) != hipSuccess ||
        (he = alloc(accb * (size_t)(WB * wstride), &W2[0])) != hipSuccess ||
        (he = alloc(accb * (size_t)(WB * wstride), &W1[1])) != hipSuccess ||
        (he = alloc(accb * (size_t)(WB * wstride), &W2[1])) != hipSuccess) {
        release();
        return hip_fail(c, he, "scess_t_backward: device allocation");
    }
    if ((he = hipStreamCreateWithFlags(&ws, hipStreamNonBlocking)) != hipSuccess) {
        ws = nullptr;
        release();
        return hip_fail(c, he, "scess_t_backward: stream");
    }
    for (int q = 0; q < 2 && he == hipSuccess; q++) {
        he = hipEventCreateWithFlags(&evW[q], hipEventDisableTiming);
        if (he == hipSuccess)
            he = hipEventCreateWithFlags(&evDone[q], hipEventDisableTiming);
    }
    // the W set zeroed by alloc on c->stream must be clear before the second stream writes it
    if (he == hipSuccess)
        he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
        release();
        return hip_fail(c, he, "scess_t_backward: events");
    }

    const int blk = wbuild_tile(prec, (int)v);
    // triples with i == j build only W1; their W2 is read from W1 (as the forward epilogue)
    WBuildArgs wa = wbuild_args(t, prec, blk, wstride);
    EpiArgs ea{};
    ea.wstride = wstride;
    ea.o = (int)o;
    ea.v = (int)v;
    ea.vR = (int)vR;
    ea.nb = t.nb;
    ea.orbits = t.d_orbits;
    ea.orbtab = t.d_orbtab;
    ea.norb = t.norb;
    ea.ngrp = t.ngrp;
    ea.t1 = t.t1;
    ea.e = t.e;
    ea.ev = t.ev;
    ea.U = U;

    static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int64_t pos_stride[3] = {vR * vR, vR, 1};
    const double *X64 = (const double *)o64.X, *T64 = (const double *)o64.T, *Rn64 = (const double *)o64.Rn,
                 *TM64 = (const double *)o64.TM;
    int rc = SCESS_OK;
    bool cancelled = false;
    const int64_t nwork = (int64_t)t.work.size();
    const int64_t w_begin = nwork * c->shard / c->nshards, w_end = nwork * (c->shard + 1) / c->nshards;
    for (int64_t w = w_begin; w < w_end && rc == SCESS_OK; w++) {
        if ((w & 7) == 0) {
            if ((he = hipStreamSynchronize(c->stream)) != hipSuccess) {
                rc = hip_fail(c, he, "scess_t_backward");
                break;
            }
            if (c->should_stop && c->should_stop(c->user)) {
                cancelled = true;
                break;
            }
        }
        const int occ[3] = {t.work_ijk[3 * w], t.work_ijk[3 * w + 1], t.work_ijk[3 * w + 2]};
        const double deg = (occ[0] == occ[1] || occ[1] == occ[2]) ? 2.0 : 1.0;
        const int *dijk = t.d_work_ijk + 3 * w;
        const int64_t slot = (w - w_begin) % WB, batch = (w - w_begin) / WB, set = batch & 1;
        if (slot == 0) {
            // W of batch b + 1 on the second stream (W of batch 0 first), once batch b - 1 has released
            // that set; batch b waits for its own W
            auto build = [&](int64_t b) {
                const int64_t w0 = w_begin + b * WB;
                hipError_t e = hipSuccess;
                if (b >= 2)
                    e = hipStreamWaitEvent(ws, evDone[b & 1], 0);
                wa.ijk = t.d_work_ijk + 3 * w0;
                wa.W1 = W1[b & 1];
                wa.W2 = W2[b & 1];
                if (e == hipSuccess)
                    e = launch_wbuild(prec, blk, wbuild_dma(), wa, (int)std::min<int64_t>(WB, w_end - w0), ws);
                if (e == hipSuccess)
                    e = hipEventRecord(evW[b & 1], ws);
                return e;
            };
            if (batch == 0)
                he = build(0);
            if (he == hipSuccess && w + WB < w_end)
                he = build(batch + 1);
            if (he == hipSuccess)
                he = hipStreamWaitEvent(c->stream, evW[set], 0);
        }
        ea.ijk = dijk;
        ea.W1 = (const char *)W1[set] + (size_t)slot * wstride * accb;
        ea.W2 = occ[0] == occ[1] ? ea.W1 : (const char *)W2[set] + (size_t)slot * wstride * accb;
        if (he == hipSuccess)
            he = launch_vpieces(t.S, dijk, (int)o, (int)v, (int)vR, U, 1, c->stream);
        // G_W entries with an index in [v, vR) are zero: the buffer starts zeroed and the G kernel writes
        // exact zeros past v (the back-contraction GEMMs read them as K padding)
        if (he == hipSuccess)
            he = launch_adjoint_g(prec, ea, dE * 2.0 / deg, GW, GV, c->stream);
        // With occ[0] == occ[1] (every triple with two equal indices, as evaluated) G is symmetric in its
        // first two indices, and the permutation terms pair up: the term with the values 0 and 1 of its
        // permutation exchanged (PERM rows 0-2, 1-3, 4-5) reads the same operands and writes the same
        // output. Such a triple runs one term of each pair with alpha = 2 (half the GEMM work).
        const bool twin = occ[0] == occ[1];
        // dX[p][x][y][f] += sum_z G(x,y,z) T[r][q][z][f]: the two permutation terms with the same
        // position for x (PERM rows 2m, 2m+1) write the same dX[p], so they run as one GEMM with
        // two K segments (twin: pair 1 equals pair 0, and the two segments of pair 2 are equal)
        // The GEMMs of each kind write disjoint slices (dX[p] for the three values of p, dT[r][q] and
        // dRn[p][q] for six distinct pairs, dTM[r] for three values), so each kind runs as one group, and
        // every output element still receives this triple's contribution once, in triple order.
        GemmGroup grp{};
        grp.n = 0;
        for (int pr = 0; pr < 3 && he == hipSuccess; pr++) {
            if (twin && pr == 1)
                continue;
            const int64_t p = occ[PERM[2 * pr][0]], sx = pos_stride[PERM[2 * pr][0]];
            GemmArgs g{};
            g.nseg = twin && pr == 2 ? 1 : 2;
            g.M = (int)v;
            g.N = (int)v;
            g.alpha = twin ? 2.0 : 1.0;
            g.splits = 1;
            for (int h = 0; h < g.nseg; h++) {
                const int sgm = 2 * pr + h;
                const int64_t q = occ[PERM[sgm][1]], r = occ[PERM[sgm][2]];
                const int64_t sy = pos_stride[PERM[sgm][1]], sz = pos_stride[PERM[sgm][2]];
                GemmSeg &sg = g.seg[h];
                sg.K1 = 1;
                sg.K2 = (int)v;
                // rows x batch y, or rows y batch x
                sg.A = {GW, sx == 1 ? sx : sy, 0, sz, sx == 1 ? sy : sx, vR, true};
                sg.B = {T64 + (r * o + q) * v * vK, 1, 0, vK, 0, vK, false};
            }
            g.C = dX + p * v * v * vK;
            g.sCn = 1;
            g.sCb = sx == 1 ? vK : v * vK;
            g.sCm = sx == 1 ? v * vK : vK;
            grp.g[grp.n++] = g;
        }
        if (he == hipSuccess)
            he = launch_gemm_group(grp, (int)v, c->stream);
        GemmGroup grpT{}, grpR{};
        grpT.n = grpR.n = 0;
        for (int sgm = 0; sgm < 6 && he == hipSuccess; sgm++) {
            if (twin && (sgm == 2 || sgm == 3 || sgm == 5))
                continue;  // the twins of terms 0, 1, 4
            const double alpha = twin ? 2.0 : 1.0;
            const int64_t p = occ[PERM[sgm][0]], q = occ[PERM[sgm][1]], r = occ[PERM[sgm][2]];
            const int64_t sx = pos_stride[PERM[sgm][0]], sy = pos_stride[PERM[sgm][1]], sz = pos_stride[PERM[sgm][2]];
            // dT[r][q][z][f] += sum_{x,y} G(x,y,z) X[p][x][y][f]
            {
                GemmArgs g{};
                g.nseg = 1;
                g.M = (int)v;
                g.N = (int)v;
                g.alpha = alpha;
                g.splits = PT;
                g.scratch = scr + (size_t)grpT.n * PT * v * v;
                GemmSeg &sg = g.seg[0];
                sg.K1 = (int)v;
                sg.K2 = (int)v;
                if (sx < sy) {  // inner K over x
                    sg.A = {GW, sz, sy, sx, 0, vR, true};
                    sg.B = {X64 + p * v * v * vK, 1, vK, v * vK, 0, vK, false};
                } else {  // inner K over y
                    sg.A = {GW, sz, sx, sy, 0, vR, true};
                    sg.B = {X64 + p * v * v * vK, 1, v * vK, vK, 0, vK, false};
                }
                g.C = dT + (r * o + q) * v * vK;
                g.sCm = vK;
                g.sCn = 1;
                grpT.g[grpT.n++] = g;
            }
            // dRn[p][q][x][m] += sum_{y,z} G(x,y,z) TM[r][y][z][m] (its own group; run after the dT group,
            // so the two reuse one scratch area)
            {
                GemmArgs g{};
                g.nseg = 1;
                g.M = (int)v;
                g.N = (int)o;
                g.alpha = alpha;
                g.splits = PR;
                g.scratch = scr + (size_t)grpR.n * PR * v * o;
                GemmSeg &sg = g.seg[0];
                sg.K1 = (int)v;
                sg.K2 = (int)v;
                if (sy < sz) {  // inner K over y
                    sg.A = {GW, sx, sz, sy, 0, vR, true};
                    sg.B = {TM64 + r * v * v * oK, 1, oK, v * oK, 0, oK, false};
                } else {  // inner K over z
                    sg.A = {GW, sx, sy, sz, 0, vR, true};
                    sg.B = {TM64 + r * v * v * oK, 1, v * oK, oK, 0, oK, false};
                }
                g.C = dRn + (p * o + q) * v * oK;
                g.sCm = oK;
                g.sCn = 1;
                grpR.g[grpR.n++] = g;
            }
        }
        if (he == hipSuccess)
            he = launch_gemm_group(grpT, 1, c->stream);
        if (he == hipSuccess)
            he = launch_gemm_group(grpR, 1, c->stream);
        // dTM[r][y][z][m] += sum_x Rn[p][q][x][m] G(x,y,z): the two terms with the same position for z
        // write the same dTM[r]; one GEMM with two K segments each (rows z and batch y when z is the
        // contiguous position of G, else rows y and batch z)
        // (twin: the two segments of pair 0 are equal, and pair 2 equals pair 1)
        static const int ZPAIR[3][2] = {{0, 2}, {1, 4}, {3, 5}};
        grp.n = 0;
        for (int pr = 0; pr < 3 && he == hipSuccess; pr++) {
            if (twin && pr == 2)
                continue;
            const int64_t r = occ[PERM[ZPAIR[pr][0]][2]], sz = pos_stride[PERM[ZPAIR[pr][0]][2]];
            GemmArgs g{};
            g.nseg = twin && pr == 0 ? 1 : 2;
            g.M = (int)v;
            g.N = (int)o;
            g.alpha = twin ? 2.0 : 1.0;
            g.splits = 1;
            for (int h = 0; h < g.nseg; h++) {
                const int sgm = ZPAIR[pr][h];
                const int64_t p = occ[PERM[sgm][0]], q = occ[PERM[sgm][1]];
                const int64_t sx = pos_stride[PERM[sgm][0]], sy = pos_stride[PERM[sgm][1]];
                GemmSeg &sg = g.seg[h];
                sg.K1 = 1;
                sg.K2 = (int)v;
                sg.B = {Rn64 + (p * o + q) * v * oK, 1, 0, oK, 0, oK, false};
                sg.A = sz == 1 ? GemmOp{GW, sz, 0, sx, sy, vR, true} : GemmOp{GW, sy, 0, sx, sz, vR, true};
            }
            g.C = dTM + r * v * v * oK;
            g.sCn = 1;
            g.sCb = sz == 1 ? v * oK : oK;
            g.sCm = sz == 1 ? oK : v * oK;
            grp.g[grp.n++] = g;
        }
        if (he == hipSuccess)
            he = launch_gemm_group(grp, (int)v, c->stream);
        if (he == hipSuccess)
            he = launch_vback(GV, U, t.t1, occ[0], occ[1], occ[2], (int)o, (int)v, (int)vR, dU, dtv, scr, dS, dt1,
                              c->stream);  // scr (the split-K scratch, >= v^2, free again here) holds H
        // the last triple of a batch releases its W set
        if (he == hipSuccess && (slot == WB - 1 || w == w_end - 1))
            he = hipEventRecord(evDone[set], c->stream);
        if (he != hipSuccess)
            rc = hip_fail(c, he, "scess_t_backward: kernel launch");
    }
    if (rc == SCESS_OK && !cancelled) {
        struct {
            int mode;
            const double *A, *B;
            double *host;
            size_t n;
        } outs[5] = {{0, dX, nullptr, grad->ovvv, (size_t)(o * v * v * v)},
                     {1, dRn, nullptr, grad->ovoo, (size_t)(o * v * o * o)},
                     {2, dS, nullptr, grad->ovov, (size_t)(o * v * o * v)},
                     {3, dT, dTM, grad->t2, (size_t)(o * o * v * v)},
                     {4, dt1, nullptr, grad->t1, (size_t)(o * v)}};
        for (auto &u : outs) {
            he = launch_unpack(u.mode, u.A, u.B, out, (int)o, (int)v, vK, oK, vR, (int64_t)u.n, c->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync(u.host, out, 8 * u.n, hipMemcpyDeviceToHost, c->stream);
            if (he == hipSuccess)
                he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) {
                rc = hip_fail(c, he, "scess_t_backward: unpack");
                break;
            }
        }
    }
    release();
    if (rc == SCESS_OK && cancelled)
        rc = fail(c, SCESS_CANCELLED, "scess_t_backward: cancelled");
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------- CCSD ABI
namespace {

int engine_fail(scess_ctx *c, const std::exception &ex) {
    const std::string m = ex.what();
    const int code = m.find("out of memory") != std::string::npos ? SCESS_ERR_OOM : SCESS_ERR_HIP;
    return fail(c, code, m);
}

// Contraction precision of the engine for one scope (FP64 again afterwards, also on exceptions)
struct PrecScope {
    scess::Engine &E;
    PrecScope(scess::Engine &e, int prec) : E(e) { E.prec = prec; }
    ~PrecScope() { E.prec = 0; }
};

// ColMajor host tensor (labels in index order) -> row-major device tensor with the same label order
scess::DTensor upload_colmajor(scess::Engine &E, const double *h, std::vector<int64_t> dims) {
    std::vector<int64_t> rev(dims.rbegin(), dims.rend());
    scess::DTensor raw = E.alloc(rev);
    E.upload(raw, h);
    scess::DTensor out = E.alloc(dims);
    std::string lab = std::string("abcdefgh").substr(0, dims.size());
    std::string rl(lab.rbegin(), lab.rend());
    E.permute(rl, raw.p, raw.dims, lab, out.p, out.dims, 1.0, 0.0);
    return out;
}

void download_colmajor(scess::Engine &E, const scess::DTensor &t, double *h) {
    std::vector<int64_t> rev(t.dims.rbegin(), t.dims.rend());
    scess::DTensor tmp = E.alloc(rev);
    std::string lab = std::string("abcdefgh").substr(0, t.dims.size());
    std::string rl(lab.rbegin(), lab.rend());
    E.permute(lab, t.p, t.dims, rl, tmp.p, tmp.dims, 1.0, 0.0);
    E.download(tmp, h);
}

// Solve the bordered DIIS system with full pivoting (Eigen's fullPivLu in the reference)
std::vector<double> solve_full_pivot(std::vector<double> A, std::vector<double> b, int n) {
    std::vector<int> cp(n);
    for (int i = 0; i < n; i++)
        cp[i] = i;
    for (int k = 0; k < n; k++) {
        int pr = k, pc = k;
        double best = -1.0;
        for (int i = k; i < n; i++)
            for (int j = k; j < n; j++)
                if (std::fabs(A[i * n + j]) > best) {
                    best = std::fabs(A[i * n + j]);
                    pr = i;
                    pc = j;
                }
        if (best == 0.0)
            break;
        if (pr != k) {
            for (int j = 0; j < n; j++)
                std::swap(A[k * n + j], A[pr * n + j]);
            std::swap(b[k], b[pr]);
        }
        if (pc != k) {
            for (int i = 0; i < n; i++)
                std::swap(A[i * n + k], A[i * n + pc]);
            std::swap(cp[k], cp[pc]);
        }
        for (int i = k + 1; i < n; i++) {
            const double f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; j++)
                A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    std::vector<double> y(n, 0.0), x(n, 0.0);
    for (int k = n - 1; k >= 0; k--) {
        if (A[k * n + k] == 0.0) {
            y[k] = 0.0;
            continue;
        }
        double s2 = b[k];
        for (int j = k + 1; j < n; j++)
            s2 -= A[k * n + j] * y[j];
        y[k] = s2 / A[k * n + k];
    }
    for (int k = 0; k < n; k++)
        x[cp[k]] = y[k];
    return x;
}

}  // namespace

extern "C" {

int scess_ccsd_set_inputs(scess_ctx *c, const scess_ccsd_in *in) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!in || in->o < 1 || in->v < 1 || !in->e || !in->ovov || !in->oovv || !in->ovvo || !in->ovvv || !in->ovoo ||
        !in->oooo || (!in->vvvv && !in->chol_vv) || (in->chol_vv && in->naux < 1))
        return fail(c, SCESS_ERR_ARG, "scess_ccsd_set_inputs: missing block or non-positive size");
    HIP_TRY(c, hipSetDevice(c->device));
    try {
        scess::Engine &E = c->eng;
        scess::CcsdState &S = c->cc;
        S = scess::CcsdState{};
        const int64_t o = in->o, v = in->v;
        S.o = o;
        S.v = v;
        S.ovov = upload_colmajor(E, in->ovov, {o, v, o, v});
        S.oovv = upload_colmajor(E, in->oovv, {o, o, v, v});
        S.ovvo = upload_colmajor(E, in->ovvo, {o, v, v, o});
        S.ovvv = upload_colmajor(E, in->ovvv, {o, v, v, v});
        S.ovoo = upload_colmajor(E, in->ovoo, {o, v, o, o});
        {
            scess::DTensor oooo = upload_colmajor(E, in->oooo, {o, o, o, o});  // (k,i,l,j)
            S.oooo_phys = E.alloc({o, o, o, o});
            E.einsum("kilj->klij", oooo, nullptr, S.oooo_phys, 1.0, 0.0);
        }
        S.oovv_phys = E.alloc({o, o, v, v});
        E.einsum("iajb->ijab", S.ovov, nullptr, S.oovv_phys, 1.0, 0.0);
        S.L_oovv_phys = E.alloc({o, o, v, v});
        E.einsum("ijab->ijab", S.oovv_phys, nullptr, S.L_oovv_phys, 2.0, 0.0);
        E.einsum("ijba->ijab", S.oovv_phys, nullptr, S.L_oovv_phys, -1.0, 1.0);
        S.L_ovov = E.alloc({o, v, o, v});
        E.einsum("kcld->kcld", S.ovov, nullptr, S.L_ovov, 2.0, 0.0);
        E.einsum("kdlc->kcld", S.ovov, nullptr, S.L_ovov, -1.0, 1.0);
        if (in->vvvv) {
            scess::DTensor vvvv = upload_colmajor(E, in->vvvv, {v, v, v, v});  // (a,c,b,d)
            S.vvvv_phys = E.alloc({v, v, v, v});
            E.einsum("acbd->abcd", vvvv, nullptr, S.vvvv_phys, 1.0, 0.0);
        } else {
            S.naux = in->naux;
            // kept as [a][c][P]: the ladder's tile regeneration contracts over the contiguous P
            S.Lvv = upload_colmajor(E, in->chol_vv, {v, v, in->naux});
        }
        S.e = E.alloc({o + v});
        E.upload(S.e, in->e);
        if (in->t2_driver)
            S.t2driver = upload_colmajor(E, in->t2_driver, {o, o, v, v});
        if (in->f12_energy)
            S.f12energy = upload_colmajor(E, in->f12_energy, {o, o, v, v});
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        S.have = true;
    } catch (const std::exception &ex) {
        c->cc = scess::CcsdState{};
        return engine_fail(c, ex);
    }
    return SCESS_OK;
}

int scess_ccsd_update(scess_ctx *c, const double *t1, const double *t2, double *t1_new, double *t2_new) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!t1 || !t2 || !t1_new || !t2_new)
        return fail(c, SCESS_ERR_ARG, "scess_ccsd_update: null array");
    if (!c->cc.have)
        return fail(c, SCESS_ERR_STATE, "scess_ccsd_update: no inputs (call scess_ccsd_set_inputs first)");
    HIP_TRY(c, hipSetDevice(c->device));
    try {
        scess::Engine &E = c->eng;
        scess::CcsdState &S = c->cc;
        scess::DTensor d1 = upload_colmajor(E, t1, {S.o, S.v});
        scess::DTensor d2 = upload_colmajor(E, t2, {S.o, S.o, S.v, S.v});
        const PrecScope ps(E, c->ccsd_prec);
        S.update(E, d1, d2);
        download_colmajor(E, d1, t1_new);
        download_colmajor(E, d2, t2_new);
    } catch (const std::exception &ex) {
        return engine_fail(c, ex);
    }
    return SCESS_OK;
}

int scess_ccsd_energy(scess_ctx *c, const double *t1, const double *t2, double *e_corr) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!t1 || !t2 || !e_corr)
        return fail(c, SCESS_ERR_ARG, "scess_ccsd_energy: null argument");
    if (!c->cc.have)
        return fail(c, SCESS_ERR_STATE, "scess_ccsd_energy: no inputs");
    HIP_TRY(c, hipSetDevice(c->device));
    try {
        scess::Engine &E = c->eng;
        scess::CcsdState &S = c->cc;
        scess::DTensor d1 = upload_colmajor(E, t1, {S.o, S.v});
        scess::DTensor d2 = upload_colmajor(E, t2, {S.o, S.o, S.v, S.v});
        *e_corr = S.energy(E, d1, d2);
    } catch (const std::exception &ex) {
        return engine_fail(c, ex);
    }
    return SCESS_OK;
}

int scess_ccsd_set_precision(scess_ctx *c, int prec) {
    if (!c)
        return SCESS_ERR_ARG;
    if (prec < SCESS_PREC_FP64 || prec > SCESS_PREC_BF16X3)
        return fail(c, SCESS_ERR_ARG, "scess_ccsd_set_precision: unknown precision");
    c->ccsd_prec = prec;
    return SCESS_OK;
}

int scess_transform_tei(scess_ctx *c, int64_t nao, int64_t nmo, const double *tei, const double *C,
                        int chem_to_phys, double *out) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!tei || !C || !out || nao <= 0 || nmo <= 0)
        return fail(c, SCESS_ERR_ARG, "scess_transform_tei: bad argument");
    HIP_TRY(c, hipSetDevice(c->device));
    try {
        scess::Engine &E = c->eng;
        // one index at a time, as TransformTEI (IntegralTransforms.cpp:33-58), last index first
        scess::DTensor cm = upload_colmajor(E, C, {nao, nmo});
        scess::DTensor g = upload_colmajor(E, tei, {nao, nao, nao, nao});
        scess::DTensor h = E.alloc({nao, nao, nao, nmo});
        E.einsum("abcd,ds->abcs", g, &cm, h, 1.0, 0.0);
        g = E.alloc({nao, nao, nmo, nmo});
        E.einsum("abcs,cr->abrs", h, &cm, g, 1.0, 0.0);
        h = E.alloc({nao, nmo, nmo, nmo});
        E.einsum("abrs,bq->aqrs", g, &cm, h, 1.0, 0.0);
        g = E.alloc({nmo, nmo, nmo, nmo});
        E.einsum("aqrs,ap->pqrs", h, &cm, g, 1.0, 0.0);
        if (chem_to_phys) {
            h = E.alloc({nmo, nmo, nmo, nmo});
            E.permute("pqrs", g.p, g.dims, "prqs", h.p, h.dims, 1.0, 0.0);
            download_colmajor(E, h, out);
        } else {
            download_colmajor(E, g, out);
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } catch (const std::exception &ex) {
        return engine_fail(c, ex);
    }
    return SCESS_OK;
}

int scess_ccsd_solve(scess_ctx *c, double *t1_out, double *t2_out, scess_ccsd_result *res, scess_ccsd_log log,
                     void *user) {
    if (!c)
        return SCESS_ERR_ARG;
    if (!res)
        return fail(c, SCESS_ERR_ARG, "scess_ccsd_solve: null result");
    std::memset(res, 0, sizeof(*res));
    if (!c->cc.have)
        return fail(c, SCESS_ERR_STATE, "scess_ccsd_solve: no inputs");
    HIP_TRY(c, hipSetDevice(c->device));
    // thresholds of SolveRCCSD_DIIS (RCCSD.cpp:302-307)
    constexpr double ENERGY_CONVERGENCE = 1.0e-10, AMPLITUDE_CONVERGENCE = 1.0e-8;
    constexpr int MAX_ITERATIONS = 100, DIIS_START = 2, MAX_DIIS_VECS = 8;
    try {
        scess::Engine &E = c->eng;
        scess::CcsdState &S = c->cc;
        const int64_t o = S.o, v = S.v, n1 = o * v, n2 = o * o * v * v, nt = n1 + n2;
        // flattened amplitudes [t1 | t2] (device row-major order); t1/t2 view the current vector
        scess::DTensor t1 = E.alloc({o, v}, true), t2 = E.alloc({o, o, v, v});
        // MP2 guess (RCCSD.cpp:333-339)
        E.copy(t2, S.oovv_phys);
        if (S.t2driver.p)
            E.axpby(t2.p, t2.p, S.t2driver.p, 1.0, 1.0, n2);
        {
            // t2 = (<ij|ab> + driver) / D with the update's Jacobi kernel (t1 stays 0)
            scess::DTensor z1 = E.alloc({o, v}, true);
            HIP_TRY(c, scess::ccsd_jacobi(z1.p, t2.p, S.e.p, (int)o, (int)v, c->stream));
        }
        std::vector<scess::DTensor> amps, errs;
        double E_last = 1.7976931348623157e308;
        bool converged = false;
        int it = 0;
        scess::DTensor prev = E.alloc({nt}), upd = E.alloc({nt}), change = E.alloc({nt});
        for (it = 0; it < MAX_ITERATIONS; it++) {
            if (c->should_stop && c->should_stop(c->user)) {
                res->iterations = it;
                return fail(c, SCESS_CANCELLED, "scess_ccsd_solve: cancelled");
            }
            HIP_TRY(c, hipMemcpyAsync(prev.p, t1.p, 8 * n1, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(prev.p + n1, t2.p, 8 * n2, hipMemcpyDeviceToDevice, c->stream));
            const PrecScope ps(E, c->ccsd_prec);
            S.update(E, t1, t2);
            HIP_TRY(c, hipMemcpyAsync(upd.p, t1.p, 8 * n1, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(upd.p + n1, t2.p, 8 * n2, hipMemcpyDeviceToDevice, c->stream));
            E.axpby(change.p, upd.p, prev.p, 1.0, -1.0, nt);
            // AmplitudeDIIS::Extrapolate (RCCSD.cpp:246-274)
            {
                scess::DTensor a = E.alloc({nt}), er = E.alloc({nt});
                E.copy(a, upd);
                E.copy(er, change);
                amps.push_back(std::move(a));
                errs.push_back(std::move(er));
                if ((int)errs.size() > MAX_DIIS_VECS) {
                    amps.erase(amps.begin());
                    errs.erase(errs.begin());
                }
            }
            if (it >= DIIS_START) {
                const int n = (int)errs.size();
                std::vector<double> B((n + 1) * (n + 1), -1.0), rhs(n + 1, 0.0);
                B[n * (n + 1) + n] = 0.0;
                for (int i = 0; i < n; i++)
                    for (int j = 0; j <= i; j++) {
                        const double d = E.dot(errs[i].p, errs[j].p, nt);
                        B[i * (n + 1) + j] = B[j * (n + 1) + i] = d;
                    }
                double mx = 0.0;
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++)
                        mx = std::max(mx, std::fabs(B[i * (n + 1) + j]));
                if (mx > 0.0)
                    for (int i = 0; i < n; i++)
                        for (int j = 0; j < n; j++)
                            B[i * (n + 1) + j] /= mx;
                rhs[n] = -1.0;
                const std::vector<double> coef = solve_full_pivot(B, rhs, n + 1);
                HIP_TRY(c, hipMemsetAsync(upd.p, 0, 8 * nt, c->stream));
                for (int i = 0; i < n; i++)
                    E.axpby(upd.p, upd.p, amps[i].p, 1.0, coef[i], nt);
                HIP_TRY(c, hipMemcpyAsync(t1.p, upd.p, 8 * n1, hipMemcpyDeviceToDevice, c->stream));
                HIP_TRY(c, hipMemcpyAsync(t2.p, upd.p + n1, 8 * n2, hipMemcpyDeviceToDevice, c->stream));
            }
            const double Ecur = S.energy(E, t1, t2);
            const double dE = std::fabs(Ecur - E_last);
            const double rms = std::sqrt(std::max(0.0, E.dot(change.p, change.p, nt))) / std::sqrt((double)nt);
            E_last = Ecur;
            res->last_dE = dE;
            res->last_rms = rms;
            if (log)
                log(user, it, Ecur, dE, rms);
            if (dE < ENERGY_CONVERGENCE && rms < AMPLITUDE_CONVERGENCE) {
                converged = true;
                break;
            }
        }
        res->converged = converged ? 1 : 0;
        res->iterations = converged ? it + 1 : it;
        res->E_CCSD = E_last;
        if (S.f12energy.p) {
            scess::DTensor Le = E.alloc({o, o, v, v});
            E.permute("ijab", S.f12energy.p, S.f12energy.dims, "ijab", Le.p, Le.dims, 2.0, 0.0);
            E.permute("ijba", S.f12energy.p, S.f12energy.dims, "ijab", Le.p, Le.dims, -1.0, 1.0);
            res->E_F12_Coupling = E.dot(Le.p, t2.p, n2);
        }
        if (t1_out)
            download_colmajor(E, t1, t1_out);
        if (t2_out)
            download_colmajor(E, t2, t2_out);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } catch (const std::exception &ex) {
        return engine_fail(c, ex);
    }
    return SCESS_OK;
}

}  // extern "C"
