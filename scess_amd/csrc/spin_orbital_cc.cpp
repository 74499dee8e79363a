// Spin-orbital CCSD and CCSD(T) and the closed-shell MP2 energy on the host: the reference's
// independent cross-checks of the closed-shell solvers (src/Quantum/CC/CC.h:44-75, CCSD.cpp, CCSD_T.cpp;
// HF::MP2_RHF_ClosedShellAO, src/Quantum/HF/HF.h:148-151). They serve the regression runner, which
// compares the closed-shell GPU path against them; sizes are tiny ((2n)^4 antisymmetrised integrals).
//
// Equations: CCSD from the intermediates of Stanton & Gauss, J. Chem. Phys. 94, 4334 (1991), Jacobi
// steps with DIIS on the flattened amplitudes (8 vectors from the third iteration, B scaled by its
// largest entry); (T) from the spin-orbital expressions
//   D t_ijk^abc(d) = P(i/jk) P(a/bc) t_i^a <jk||bc>
//   D t_ijk^abc(c) = P(i/jk) P(a/bc) [ sum_e t_jk^ae <ei||bc> - sum_m t_im^bc <ma||jk> ]
//   E(T) = 1/36 sum t(c) D (t(c) + t(d)),  P(i/jk) f(ijk) = f(ijk) - f(jik) - f(kji).
// Spin orbital P = 2p + spin, so the occupied spin orbitals (p < nOcc) come first.
#include "../../include/scess_cc.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace SCESS {
namespace {

// <PQ||RS> over N spin orbitals from <pq|rs> (spatial, ColMajor n^4)
struct SpinInts {
    int64_t N, O, V;
    std::vector<double> g, e;
    double operator()(int64_t p, int64_t q, int64_t r, int64_t s) const { return g[((p * N + q) * N + r) * N + s]; }
};

SpinInts spin_integrals(const CC::TEITensor &eri, const std::vector<double> &eps, int64_t nOcc) {
    const int64_t n = eri.dimension(0);
    SpinInts S;
    S.N = 2 * n;
    S.O = 2 * nOcc;
    S.V = S.N - S.O;
    S.g.assign((size_t)(S.N * S.N * S.N * S.N), 0.0);
    S.e.resize((size_t)S.N);
    for (int64_t P = 0; P < S.N; P++)
        S.e[(size_t)P] = eps[(size_t)(P / 2)];
    auto phys = [&](int64_t P, int64_t Q, int64_t R, int64_t T) {
        if ((P & 1) != (R & 1) || (Q & 1) != (T & 1))
            return 0.0;
        return eri(P / 2, Q / 2, R / 2, T / 2);
    };
    for (int64_t P = 0; P < S.N; P++)
        for (int64_t Q = 0; Q < S.N; Q++)
            for (int64_t R = 0; R < S.N; R++)
                for (int64_t T = 0; T < S.N; T++)
                    S.g[(size_t)(((P * S.N + Q) * S.N + R) * S.N + T)] = phys(P, Q, R, T) - phys(P, Q, T, R);
    return S;
}

// Dense solve of the (n+1) x (n+1) DIIS system by Gaussian elimination with partial pivoting
std::vector<double> solve(std::vector<double> A, std::vector<double> b, int n) {
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[(size_t)(r * n + c)]) > std::fabs(A[(size_t)(piv * n + c)]))
                piv = r;
        if (A[(size_t)(piv * n + c)] == 0.0)
            throw std::runtime_error("spin-orbital CCSD: singular DIIS system");
        for (int k = 0; k < n; k++)
            std::swap(A[(size_t)(c * n + k)], A[(size_t)(piv * n + k)]);
        std::swap(b[(size_t)c], b[(size_t)piv]);
        for (int r = c + 1; r < n; r++) {
            const double f = A[(size_t)(r * n + c)] / A[(size_t)(c * n + c)];
            for (int k = c; k < n; k++)
                A[(size_t)(r * n + k)] -= f * A[(size_t)(c * n + k)];
            b[(size_t)r] -= f * b[(size_t)c];
        }
    }
    std::vector<double> x((size_t)n);
    for (int r = n - 1; r >= 0; r--) {
        double s = b[(size_t)r];
        for (int k = r + 1; k < n; k++)
            s -= A[(size_t)(r * n + k)] * x[(size_t)k];
        x[(size_t)r] = s / A[(size_t)(r * n + r)];
    }
    return x;
}

struct SpinAmps {
    std::vector<double> t1, t2;  // t1[i][a] (O x V), t2[i][j][a][b]
};

// Converged spin-orbital CCSD amplitudes; E_CCSD on return
SpinAmps spin_ccsd(const SpinInts &g, double &E_CCSD, bool &converged) {
    const int64_t O = g.O, V = g.V;
    auto T1 = [&](int64_t i, int64_t a) -> size_t { return (size_t)(i * V + a); };
    auto T2 = [&](int64_t i, int64_t j, int64_t a, int64_t b) -> size_t {
        return (size_t)(((i * O + j) * V + a) * V + b);
    };
    auto Dia = [&](int64_t i, int64_t a) { return g.e[(size_t)i] - g.e[(size_t)(O + a)]; };
    auto Dijab = [&](int64_t i, int64_t j, int64_t a, int64_t b) {
        return g.e[(size_t)i] + g.e[(size_t)j] - g.e[(size_t)(O + a)] - g.e[(size_t)(O + b)];
    };
    const size_t n1 = (size_t)(O * V), n2 = (size_t)(O * O * V * V);
    SpinAmps t;
    t.t1.assign(n1, 0.0);
    t.t2.assign(n2, 0.0);
    for (int64_t i = 0; i < O; i++)
        for (int64_t j = 0; j < O; j++)
            for (int64_t a = 0; a < V; a++)
                for (int64_t b = 0; b < V; b++)
                    t.t2[T2(i, j, a, b)] = g(i, j, O + a, O + b) / Dijab(i, j, a, b);
    auto energy = [&](const SpinAmps &x) {
        double E = 0.0;
        for (int64_t i = 0; i < O; i++)
            for (int64_t j = 0; j < O; j++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++)
                        E += g(i, j, O + a, O + b) *
                             (0.25 * x.t2[T2(i, j, a, b)] + 0.5 * x.t1[T1(i, a)] * x.t1[T1(j, b)]);
        return E;
    };
    std::vector<std::vector<double>> hx, he;
    double Eold = 0.0;
    converged = false;
    for (int it = 0; it < 300; it++) {
        const auto &t1 = t.t1;
        const auto &t2 = t.t2;
        std::vector<double> tau(n2), taut(n2);
        for (int64_t i = 0; i < O; i++)
            for (int64_t j = 0; j < O; j++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++) {
                        const double x = t1[T1(i, a)] * t1[T1(j, b)] - t1[T1(i, b)] * t1[T1(j, a)];
                        tau[T2(i, j, a, b)] = t2[T2(i, j, a, b)] + x;
                        taut[T2(i, j, a, b)] = t2[T2(i, j, a, b)] + 0.5 * x;
                    }
        std::vector<double> Fae((size_t)(V * V), 0.0), Fmi((size_t)(O * O), 0.0), Fme(n1, 0.0);
        for (int64_t a = 0; a < V; a++)
            for (int64_t e = 0; e < V; e++) {
                double s = 0.0;
                for (int64_t m = 0; m < O; m++)
                    for (int64_t f = 0; f < V; f++) {
                        s += t1[T1(m, f)] * g(m, O + a, O + f, O + e);
                        for (int64_t n = 0; n < O; n++)
                            s -= 0.5 * taut[T2(m, n, a, f)] * g(m, n, O + e, O + f);
                    }
                Fae[(size_t)(a * V + e)] = s;
            }
        for (int64_t m = 0; m < O; m++)
            for (int64_t i = 0; i < O; i++) {
                double s = 0.0;
                for (int64_t n = 0; n < O; n++)
                    for (int64_t e = 0; e < V; e++) {
                        s += t1[T1(n, e)] * g(m, n, i, O + e);
                        for (int64_t f = 0; f < V; f++)
                            s += 0.5 * taut[T2(i, n, e, f)] * g(m, n, O + e, O + f);
                    }
                Fmi[(size_t)(m * O + i)] = s;
            }
        for (int64_t m = 0; m < O; m++)
            for (int64_t e = 0; e < V; e++) {
                double s = 0.0;
                for (int64_t n = 0; n < O; n++)
                    for (int64_t f = 0; f < V; f++)
                        s += t1[T1(n, f)] * g(m, n, O + e, O + f);
                Fme[T1(m, e)] = s;
            }
        auto W4 = [](int64_t a, int64_t b, int64_t c, int64_t d, int64_t B, int64_t C, int64_t D) {
            return (size_t)(((a * B + b) * C + c) * D + d);
        };
        std::vector<double> Wmnij((size_t)(O * O * O * O)), Wabef((size_t)(V * V * V * V)),
            Wmbej((size_t)(O * V * V * O));
        for (int64_t m = 0; m < O; m++)
            for (int64_t n = 0; n < O; n++)
                for (int64_t i = 0; i < O; i++)
                    for (int64_t j = 0; j < O; j++) {
                        double s = g(m, n, i, j);
                        for (int64_t e = 0; e < V; e++) {
                            s += t1[T1(j, e)] * g(m, n, i, O + e) - t1[T1(i, e)] * g(m, n, j, O + e);
                            for (int64_t f = 0; f < V; f++)
                                s += 0.25 * tau[T2(i, j, e, f)] * g(m, n, O + e, O + f);
                        }
                        Wmnij[W4(m, n, i, j, O, O, O)] = s;
                    }
        for (int64_t a = 0; a < V; a++)
            for (int64_t b = 0; b < V; b++)
                for (int64_t e = 0; e < V; e++)
                    for (int64_t f = 0; f < V; f++) {
                        double s = g(O + a, O + b, O + e, O + f);
                        for (int64_t m = 0; m < O; m++) {
                            s += -t1[T1(m, b)] * g(O + a, m, O + e, O + f) + t1[T1(m, a)] * g(O + b, m, O + e, O + f);
                            for (int64_t n = 0; n < O; n++)
                                s += 0.25 * tau[T2(m, n, a, b)] * g(m, n, O + e, O + f);
                        }
                        Wabef[W4(a, b, e, f, V, V, V)] = s;
                    }
        for (int64_t m = 0; m < O; m++)
            for (int64_t b = 0; b < V; b++)
                for (int64_t e = 0; e < V; e++)
                    for (int64_t j = 0; j < O; j++) {
                        double s = g(m, O + b, O + e, j);
                        for (int64_t f = 0; f < V; f++)
                            s += t1[T1(j, f)] * g(m, O + b, O + e, O + f);
                        for (int64_t n = 0; n < O; n++) {
                            s -= t1[T1(n, b)] * g(m, n, O + e, j);
                            for (int64_t f = 0; f < V; f++)
                                s -= (0.5 * t2[T2(j, n, f, b)] + t1[T1(j, f)] * t1[T1(n, b)]) * g(m, n, O + e, O + f);
                        }
                        Wmbej[W4(m, b, e, j, V, V, O)] = s;
                    }
        // T1 residual
        std::vector<double> r1(n1, 0.0), r2(n2, 0.0);
        for (int64_t i = 0; i < O; i++)
            for (int64_t a = 0; a < V; a++) {
                double s = 0.0;
                for (int64_t e = 0; e < V; e++)
                    s += t1[T1(i, e)] * Fae[(size_t)(a * V + e)];
                for (int64_t m = 0; m < O; m++) {
                    s -= t1[T1(m, a)] * Fmi[(size_t)(m * O + i)];
                    for (int64_t e = 0; e < V; e++) {
                        s += t2[T2(i, m, a, e)] * Fme[T1(m, e)];
                        for (int64_t f = 0; f < V; f++)
                            s -= 0.5 * t2[T2(i, m, e, f)] * g(m, O + a, O + e, O + f);
                        for (int64_t n = 0; n < O; n++)
                            s -= 0.5 * t2[T2(m, n, a, e)] * g(n, m, O + e, i);
                    }
                }
                for (int64_t n = 0; n < O; n++)
                    for (int64_t f = 0; f < V; f++)
                        s -= t1[T1(n, f)] * g(n, O + a, i, O + f);
                r1[T1(i, a)] = s;
            }
        // T2 residual: the permutation operators applied to the unsymmetrised pieces
        std::vector<double> pab(n2, 0.0), pij(n2, 0.0), pijab(n2, 0.0);
        for (int64_t i = 0; i < O; i++)
            for (int64_t j = 0; j < O; j++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++) {
                        double xab = 0.0, xij = 0.0, xijab = 0.0, base = g(i, j, O + a, O + b);
                        for (int64_t e = 0; e < V; e++) {
                            double fbe = Fae[(size_t)(b * V + e)];
                            for (int64_t m = 0; m < O; m++)
                                fbe -= 0.5 * t1[T1(m, b)] * Fme[T1(m, e)];
                            xab += t2[T2(i, j, a, e)] * fbe;
                            xij += t1[T1(i, e)] * g(O + a, O + b, O + e, j);
                        }
                        for (int64_t m = 0; m < O; m++) {
                            double fmj = Fmi[(size_t)(m * O + j)];
                            for (int64_t e = 0; e < V; e++)
                                fmj += 0.5 * t1[T1(j, e)] * Fme[T1(m, e)];
                            xij -= t2[T2(i, m, a, b)] * fmj;
                            xab -= t1[T1(m, a)] * g(m, O + b, i, j);
                            for (int64_t e = 0; e < V; e++)
                                xijab += t2[T2(i, m, a, e)] * Wmbej[W4(m, b, e, j, V, V, O)] -
                                         t1[T1(i, e)] * t1[T1(m, a)] * g(m, O + b, O + e, j);
                            for (int64_t n = 0; n < O; n++)
                                base += 0.5 * tau[T2(m, n, a, b)] * Wmnij[W4(m, n, i, j, O, O, O)];
                        }
                        for (int64_t e = 0; e < V; e++)
                            for (int64_t f = 0; f < V; f++)
                                base += 0.5 * tau[T2(i, j, e, f)] * Wabef[W4(a, b, e, f, V, V, V)];
                        r2[T2(i, j, a, b)] = base;
                        pab[T2(i, j, a, b)] = xab;
                        pij[T2(i, j, a, b)] = xij;
                        pijab[T2(i, j, a, b)] = xijab;
                    }
        for (int64_t i = 0; i < O; i++)
            for (int64_t j = 0; j < O; j++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++)
                        r2[T2(i, j, a, b)] += pab[T2(i, j, a, b)] - pab[T2(i, j, b, a)] + pij[T2(i, j, a, b)] -
                                              pij[T2(j, i, a, b)] + pijab[T2(i, j, a, b)] - pijab[T2(j, i, a, b)] -
                                              pijab[T2(i, j, b, a)] + pijab[T2(j, i, b, a)];
        // Jacobi step, then DIIS
        std::vector<double> x(n1 + n2), err(n1 + n2);
        for (int64_t i = 0; i < O; i++)
            for (int64_t a = 0; a < V; a++)
                x[T1(i, a)] = r1[T1(i, a)] / Dia(i, a);
        for (int64_t i = 0; i < O; i++)
            for (int64_t j = 0; j < O; j++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++)
                        x[n1 + T2(i, j, a, b)] = r2[T2(i, j, a, b)] / Dijab(i, j, a, b);
        double rms = 0.0;
        for (size_t k = 0; k < n1 + n2; k++) {
            err[k] = x[k] - (k < n1 ? t1[k] : t2[k - n1]);
            rms += err[k] * err[k];
        }
        rms = std::sqrt(rms);
        hx.push_back(x);
        he.push_back(err);
        if (hx.size() > 8) {
            hx.erase(hx.begin());
            he.erase(he.begin());
        }
        if (it >= 2) {
            const int n = (int)hx.size();
            std::vector<double> B((size_t)((n + 1) * (n + 1)), -1.0), rhs((size_t)(n + 1), 0.0);
            B[(size_t)(n * (n + 1) + n)] = 0.0;
            double mx = 0.0;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    double d = 0.0;
                    for (size_t k = 0; k < n1 + n2; k++)
                        d += he[(size_t)a][k] * he[(size_t)b][k];
                    B[(size_t)(a * (n + 1) + b)] = d;
                    mx = std::max(mx, std::fabs(d));
                }
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++)
                    B[(size_t)(a * (n + 1) + b)] /= mx;
            rhs[(size_t)n] = -1.0;
            const std::vector<double> c = solve(B, rhs, n + 1);
            std::fill(x.begin(), x.end(), 0.0);
            for (int a = 0; a < n; a++)
                for (size_t k = 0; k < n1 + n2; k++)
                    x[k] += c[(size_t)a] * hx[(size_t)a][k];
        }
        std::copy(x.begin(), x.begin() + (ptrdiff_t)n1, t.t1.begin());
        std::copy(x.begin() + (ptrdiff_t)n1, x.end(), t.t2.begin());
        const double E = energy(t);
        if (std::fabs(E - Eold) < 1e-11 && rms < 1e-9) {
            E_CCSD = E;
            converged = true;
            return t;
        }
        Eold = E;
    }
    E_CCSD = Eold;
    return t;
}

double spin_triples(const SpinInts &g, const SpinAmps &t) {
    const int64_t O = g.O, V = g.V;
    auto T1 = [&](int64_t i, int64_t a) { return t.t1[(size_t)(i * V + a)]; };
    auto T2 = [&](int64_t i, int64_t j, int64_t a, int64_t b) { return t.t2[(size_t)(((i * O + j) * V + a) * V + b)]; };
    // unpermuted disconnected and connected pieces for one (i, j, k, a, b, c)
    auto disc = [&](int64_t i, int64_t j, int64_t k, int64_t a, int64_t b, int64_t c) {
        return T1(i, a) * g(j, k, O + b, O + c);
    };
    auto conn = [&](int64_t i, int64_t j, int64_t k, int64_t a, int64_t b, int64_t c) {
        double s = 0.0;
        for (int64_t e = 0; e < V; e++)
            s += T2(j, k, a, e) * g(O + e, i, O + b, O + c);
        for (int64_t m = 0; m < O; m++)
            s -= T2(i, m, b, c) * g(m, O + a, j, k);
        return s;
    };
    // P(i/jk) P(a/bc) of a piece
    auto perm = [&](auto f, int64_t i, int64_t j, int64_t k, int64_t a, int64_t b, int64_t c) {
        auto pa = [&](int64_t x, int64_t y, int64_t z) {
            return f(x, y, z, a, b, c) - f(x, y, z, b, a, c) - f(x, y, z, c, b, a);
        };
        return pa(i, j, k) - pa(j, i, k) - pa(k, j, i);
    };
    double E = 0.0;
    for (int64_t i = 0; i < O; i++)
        for (int64_t j = 0; j < O; j++)
            for (int64_t k = 0; k < O; k++)
                for (int64_t a = 0; a < V; a++)
                    for (int64_t b = 0; b < V; b++)
                        for (int64_t c = 0; c < V; c++) {
                            const double D = g.e[(size_t)i] + g.e[(size_t)j] + g.e[(size_t)k] -
                                             g.e[(size_t)(O + a)] - g.e[(size_t)(O + b)] - g.e[(size_t)(O + c)];
                            const double tc = perm(conn, i, j, k, a, b, c) / D;
                            const double td = perm(disc, i, j, k, a, b, c) / D;
                            E += tc * D * (tc + td);
                        }
    return E / 36.0;
}

CC::TEITensor mo_eri(const HF::HF_Wfn &rhf, const Integrals::MolInts &mints) {
    return Integrals::TransformTEI(mints.eri, rhf.C_Alpha, rhf.nmo, true);
}

int64_t occupied(const HF::HF_Wfn &rhf) {
    int64_t o = 0;
    while (o < (int64_t)rhf.Occ_Alpha.size() && rhf.Occ_Alpha[(size_t)o] == 1)
        o++;
    return o;
}

}  // namespace

namespace CC {

CCSD_Wfn CCSD_DIIS_FromMO(const TEITensor &eri, const std::vector<double> &orbitalEnergies, int64_t nOcc,
                          bool withTriples) {
    const SpinInts g = spin_integrals(eri, orbitalEnergies, nOcc);
    CCSD_Wfn w;
    const SpinAmps t = spin_ccsd(g, w.E_CCSD, w.converged);
    if (withTriples && w.converged)
        w.E_T = spin_triples(g, t);
    return w;
}

CCSD_Wfn CCSD_DIIS(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints) {
    return CCSD_DIIS_FromMO(mo_eri(rhfWfn, mints), rhfWfn.Ep_Alpha, occupied(rhfWfn), false);
}

CCSD_Wfn CCSD_T(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints) {
    return CCSD_DIIS_FromMO(mo_eri(rhfWfn, mints), rhfWfn.Ep_Alpha, occupied(rhfWfn), true);
}

}  // namespace CC

namespace HF {

double MP2_RHF_ClosedShellAO(const HF_Wfn &wfn, const Integrals::TEITensor &aoERI) {
    const CC::TEITensor eri = Integrals::TransformTEI(aoERI, wfn.C_Alpha, wfn.nmo, true);  // <pq|rs>
    const int64_t o = occupied(wfn), n = wfn.nmo;
    const std::vector<double> &e = wfn.Ep_Alpha;
    double E = 0.0;
    for (int64_t i = 0; i < o; i++)
        for (int64_t j = 0; j < o; j++)
            for (int64_t a = o; a < n; a++)
                for (int64_t b = o; b < n; b++)
                    E += eri(i, j, a, b) * (2.0 * eri(i, j, a, b) - eri(i, j, b, a)) /
                         (e[(size_t)i] + e[(size_t)j] - e[(size_t)a] - e[(size_t)b]);
    return E;
}

}  // namespace HF
}  // namespace SCESS
