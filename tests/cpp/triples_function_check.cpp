// Check of SCESS::CC::TriplesFunction (include/scess_cc.h) on the GPU, run by tests/test_cpp_api_gpu.py.
// Inputs: o = 4, v = 9, <pq|rs> = (pr|qs) from a symmetric factor (8-fold symmetric), t2 with
// t_ij^ab = t_ji^ba. Checks: forward equals RCCSD_TriplesCorrection bitwise; backward is linear in dE;
// dE/d<pq|rs>, dE/dt1 and dE/dt2 agree with central differences of RCCSD_TriplesCorrection (E is
// quadratic in each entry, so the difference is exact up to rounding); dE/d<pq|rs> is zero outside the
// three index patterns the functional reads; cancellation returns 0.0 and leaves no state to
// differentiate; the FP32, BF16X3 and OZAKI forwards stay within 1e-5, 1e-4 and 1e-12 (relative) of
// FP64, and the backward with the W recomputation in OZAKI stays within 1e-11 of the FP64 one. Prints one line per check; exit status 0 when all pass.
#include "scess_cc.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

using namespace SCESS;
using namespace SCESS::CC;

namespace {

struct Rng {  // 64-bit LCG and Box-Muller: deterministic normal deviates
    uint64_t s;
    double uniform() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal() {
        const double u = uniform(), w = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * w);
    }
};

struct Stop : Thread {
    bool ShouldStop() const override { return true; }
};

int failures = 0;
void check(bool ok, const char *what, double a, double b) {
    std::printf("%s %s: %.15e %.15e\n", ok ? "ok  " : "FAIL", what, a, b);
    if (!ok)
        failures++;
}

}  // namespace

int main() {
    const int64_t o = 4, v = 9, n = o + v, naux = 12;
    Rng rng{12345};
    std::vector<double> L((size_t)(naux * n * n));
    for (int64_t P = 0; P < naux; P++)
        for (int64_t p = 0; p < n; p++)
            for (int64_t q = 0; q <= p; q++)
                L[(P * n + p) * n + q] = L[(P * n + q) * n + p] = rng.normal() / std::sqrt((double)naux);
    TEITensor eri(n, n, n, n);  // <pq|rs> = (pr|qs)
    for (int64_t p = 0; p < n; p++)
        for (int64_t q = 0; q < n; q++)
            for (int64_t r = 0; r < n; r++)
                for (int64_t s = 0; s < n; s++) {
                    double x = 0.0;
                    for (int64_t P = 0; P < naux; P++)
                        x += L[(P * n + p) * n + r] * L[(P * n + q) * n + s];
                    eri(p, q, r, s) = x;
                }
    std::vector<double> e(n);
    for (int64_t i = 0; i < o; i++)
        e[i] = -1.5 + 0.2 * (double)i;
    for (int64_t a = 0; a < v; a++)
        e[o + a] = 0.4 + 0.15 * (double)a;
    T1Tensor t1(o, v);
    for (int64_t a = 0; a < v; a++)
        for (int64_t i = 0; i < o; i++)
            t1(i, a) = 0.05 * rng.normal();
    T2Tensor t2(o, o, v, v);
    for (int64_t b = 0; b < v; b++)
        for (int64_t a = 0; a < v; a++)
            for (int64_t j = 0; j < o; j++)
                for (int64_t i = 0; i < o; i++)
                    if (i < j || (i == j && a <= b)) {
                        const double x = 0.1 * rng.normal();
                        t2(i, j, a, b) = x;
                        t2(j, i, b, a) = x;
                    }

    try {
        const double Eref = RCCSD_TriplesCorrection(eri, e, o, t1, t2);
        TriplesFunction tf;
        const double E = tf.forward(eri, e, o, t1, t2);
        check(E == Eref && E != 0.0, "forward equals RCCSD_TriplesCorrection", E, Eref);

        const TriplesGradient g1 = tf.backward(1.0);
        const TriplesGradient g2 = tf.backward(2.0);
        double gmax = 0.0, lin = 0.0;
        for (int64_t k = 0; k < g1.eri.size(); k++) {
            gmax = std::max(gmax, std::fabs(g1.eri.data()[k]));
            lin = std::max(lin, std::fabs(g2.eri.data()[k] - 2.0 * g1.eri.data()[k]));
        }
        for (int64_t k = 0; k < g1.t2.size(); k++)
            lin = std::max(lin, std::fabs(g2.t2.data()[k] - 2.0 * g1.t2.data()[k]));
        check(gmax > 0.0 && lin <= 1e-12 * gmax, "backward linear in dE (max |g(2) - 2 g(1)|, max |g|)", lin, gmax);

        // entries no block reads: first index virtual, or third index occupied
        double outside = 0.0;
        for (int64_t p = o; p < n; p++)
            for (int64_t q = 0; q < n; q++)
                for (int64_t r = 0; r < n; r++)
                    for (int64_t s = 0; s < n; s++)
                        outside += std::fabs(g1.eri(p, q, r, s));
        for (int64_t i = 0; i < o; i++)
            for (int64_t q = 0; q < n; q++)
                for (int64_t r = 0; r < o; r++)
                    for (int64_t s = 0; s < n; s++)
                        outside += std::fabs(g1.eri(i, q, r, s));  // third index occupied: never read
        check(outside == 0.0, "dE/d<pq|rs> zero outside the three read patterns", outside, 0.0);

        // central differences of RCCSD_TriplesCorrection
        const double h = 1e-5;
        auto fd_eri = [&](int64_t p, int64_t q, int64_t r, int64_t s) {
            TEITensor x = eri;
            x(p, q, r, s) += h;
            const double ep = RCCSD_TriplesCorrection(x, e, o, t1, t2);
            x(p, q, r, s) -= 2.0 * h;
            const double em = RCCSD_TriplesCorrection(x, e, o, t1, t2);
            return (ep - em) / (2.0 * h);
        };
        auto fd_t1 = [&](int64_t i, int64_t a) {
            T1Tensor x = t1;
            x(i, a) += h;
            const double ep = RCCSD_TriplesCorrection(eri, e, o, x, t2);
            x(i, a) -= 2.0 * h;
            const double em = RCCSD_TriplesCorrection(eri, e, o, x, t2);
            return (ep - em) / (2.0 * h);
        };
        auto fd_t2 = [&](int64_t i, int64_t j, int64_t a, int64_t b) {
            T2Tensor x = t2;
            x(i, j, a, b) += h;
            const double ep = RCCSD_TriplesCorrection(eri, e, o, t1, x);
            x(i, j, a, b) -= 2.0 * h;
            const double em = RCCSD_TriplesCorrection(eri, e, o, t1, x);
            return (ep - em) / (2.0 * h);
        };
        auto near = [&](double an, double fd) {
            return std::fabs(an - fd) <= 1e-6 * std::fabs(fd) + 1e-9 * gmax;
        };
        const int64_t eidx[][4] = {{0, o + 1, o + 2, o + 3}, {2, o + 6, o + 0, o + 4},   // (ia|bf) pattern
                                   {1, 2, o + 3, 0},         {2, 0, o + 5, 1},           // (ia|mj) pattern
                                   {0, 1, o + 2, o + 4},     {2, 2, o + 6, o + 1}};      // (ia|jb) pattern
        for (const auto &ix : eidx) {
            const double an = g1.eri(ix[0], ix[1], ix[2], ix[3]), fd = fd_eri(ix[0], ix[1], ix[2], ix[3]);
            check(near(an, fd), "dE/d<pq|rs> vs central difference", an, fd);
        }
        for (const auto &ix : {std::array<int64_t, 2>{0, 3}, std::array<int64_t, 2>{2, 6}}) {
            const double an = g1.t1(ix[0], ix[1]), fd = fd_t1(ix[0], ix[1]);
            check(near(an, fd), "dE/dt1 vs central difference", an, fd);
        }
        for (const auto &ix : {std::array<int64_t, 4>{0, 1, 2, 5}, std::array<int64_t, 4>{2, 2, 0, 6},
                               std::array<int64_t, 4>{1, 0, 4, 4}}) {
            const double an = g1.t2(ix[0], ix[1], ix[2], ix[3]), fd = fd_t2(ix[0], ix[1], ix[2], ix[3]);
            check(near(an, fd), "dE/dt2 vs central difference", an, fd);
        }

        // the forward in the other product precisions, against the FP64 E(T)
        {
            TriplesFunction tm;
            using P = TriplesFunction::Precision;
            const double Ef = tm.forward(eri, e, o, t1, t2, nullptr, P::FP32);
            check(std::fabs(Ef - E) <= 1e-5 * std::fabs(E), "FP32 forward within 1e-5 (relative) of FP64", Ef, E);
            const double Eb = tm.forward(eri, e, o, t1, t2, nullptr, P::BF16X3);
            check(std::fabs(Eb - E) <= 1e-4 * std::fabs(E), "BF16X3 forward within 1e-4 (relative) of FP64", Eb, E);
            const double Ez = tm.forward(eri, e, o, t1, t2, nullptr, P::OZAKI);
            check(std::fabs(Ez - E) <= 1e-12 * std::fabs(E), "OZAKI forward within 1e-12 (relative) of FP64", Ez, E);
            // W recomputed from int8 digit products in the backward: the gradients of the FP64 pair
            TriplesFunction tz(P::OZAKI);
            (void)tz.forward(eri, e, o, t1, t2);
            const TriplesGradient gz = tz.backward(1.0);
            double dmax = 0.0, gmax = 0.0;
            for (int64_t n = 0; n < g1.t2.size(); n++) {
                dmax = std::max(dmax, std::fabs(gz.t2.data()[n] - g1.t2.data()[n]));
                gmax = std::max(gmax, std::fabs(g1.t2.data()[n]));
            }
            for (int64_t n = 0; n < g1.eri.size(); n++) {
                dmax = std::max(dmax, std::fabs(gz.eri.data()[n] - g1.eri.data()[n]));
                gmax = std::max(gmax, std::fabs(g1.eri.data()[n]));
            }
            check(dmax <= 1e-11 * gmax, "OZAKI W recomputation: backward within 1e-11 of the FP64 one", dmax, gmax);
        }

        // cancellation: forward returns 0.0 and leaves nothing to differentiate
        Stop stop;
        TriplesFunction tc;
        const double Ec = tc.forward(eri, e, o, t1, t2, &stop);
        check(Ec == 0.0, "cancelled forward returns 0.0", Ec, 0.0);
        bool threw = false;
        try {
            (void)tc.backward(1.0);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        check(threw, "backward without a successful forward throws", threw ? 1.0 : 0.0, 1.0);
    } catch (const std::exception &ex) {
        std::printf("FAIL exception: %s\n", ex.what());
        return 1;
    }
    std::printf("%s\n", failures ? "triples_function_check: FAILED" : "triples_function_check: all checks passed");
    return failures ? 1 : 0;
}
