/*
 * Gaussian AO integrals (overlap, kinetic, nuclear attraction, two-electron)
 * over contracted Cartesian shells, by the McMurchie-Davidson scheme
 * (Hermite expansion coefficients E, Hermite Coulomb integrals R from the
 * Boys function). Written from the textbook method (T. Helgaker, P. Jorgensen,
 * J. Olsen, "Molecular Electronic-Structure Theory", ch. 9).
 *
 * TEST INFRASTRUCTURE ONLY: it produces the water inputs of the end-to-end
 * checks against the reference's published energies
 * (src/SCESS/Tests.cpp:41-53). The reference evaluates these integrals with
 * libint2 (src/Quantum/Integrals/Integrals.cpp:28-53), which is absent here.
 *
 * Shells are given as flat arrays; every contraction column is its own shell.
 * Cartesian components of a shell of angular momentum l are ordered
 * lx = l..0, ly = l-lx..0, lz = l-lx-ly; primitive coefficients already
 * include the normalisation of the x^l component.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define LMAX 2
#define TMAX (4 * LMAX + 2)

/* Boys function F_n(T), n = 0..nmax, to about 1e-15 relative */
static void boys(int nmax, double T, double *F) {
    if (T < 1e-13) {
        for (int n = 0; n <= nmax; n++)
            F[n] = 1.0 / (2 * n + 1);
        return;
    }
    if (T > 45.0) {
        /* upward recursion from F_0 = sqrt(pi/T)/2 erf(sqrt T); e^-T negligible */
        F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        const double et = exp(-T);
        for (int n = 0; n < nmax; n++)
            F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
        return;
    }
    /* series for the highest order, then downward recursion */
    const double et = exp(-T);
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; k++) {
        term *= 2.0 * T / (2 * nmax + 2 * k + 1);
        sum += term;
        if (term < 1e-17 * sum)
            break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; n--)
        F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
}

/* Hermite expansion E[i][j][t] of one direction for exponents a, b and X_AB = A - B */
#define EI (LMAX + 3)
#define EJ (LMAX + 3)
#define ET (EI + EJ)
static void hermite(int imax, int jmax, double a, double b, double XAB, double E[EI][EJ][ET]) {
    const double p = a + b, mu = a * b / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
    memset(E, 0, sizeof(double) * EI * EJ * ET);
    E[0][0][0] = exp(-mu * XAB * XAB);
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            if (i == 0 && j == 0)
                continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (i > 0) {
                    v = XPA * E[i - 1][j][t];
                    if (t > 0)
                        v += E[i - 1][j][t - 1] / (2 * p);
                    if (t + 1 <= i - 1 + j)
                        v += (t + 1) * E[i - 1][j][t + 1];
                } else {
                    v = XPB * E[i][j - 1][t];
                    if (t > 0)
                        v += E[i][j - 1][t - 1] / (2 * p);
                    if (t + 1 <= i + j - 1)
                        v += (t + 1) * E[i][j - 1][t + 1];
                }
                E[i][j][t] = v;
            }
        }
}

/* Hermite Coulomb integrals R[t][u][v] = R^0_tuv(alpha, X, Y, Z) for t+u+v <= L */
static void hermite_coulomb(int L, double alpha, double X, double Y, double Z, double R[TMAX][TMAX][TMAX]) {
    static __thread double Rn[TMAX][TMAX][TMAX][TMAX];
    double F[TMAX + 1];
    boys(L, alpha * (X * X + Y * Y + Z * Z), F);
    for (int n = 0; n <= L; n++)
        Rn[n][0][0][0] = pow(-2.0 * alpha, n) * F[n];
    for (int n = L - 1; n >= 0; n--) {
        const int m = L - n; /* total order reachable from level n */
        for (int t = 0; t <= m; t++)
            for (int u = 0; u <= m - t; u++)
                for (int v = 0; v <= m - t - u; v++) {
                    if (t + u + v == 0)
                        continue;
                    double val;
                    if (t > 0) {
                        val = X * Rn[n + 1][t - 1][u][v];
                        if (t > 1)
                            val += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = Y * Rn[n + 1][t][u - 1][v];
                        if (u > 1)
                            val += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        val = Z * Rn[n + 1][t][u][v - 1];
                        if (v > 1)
                            val += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; t++)
        for (int u = 0; u <= L - t; u++)
            for (int v = 0; v <= L - t - u; v++)
                R[t][u][v] = Rn[0][t][u][v];
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_powers(int l, int *px, int *py, int *pz) {
    int c = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            px[c] = lx;
            py[c] = ly;
            pz[c] = l - lx - ly;
            c++;
        }
}

typedef struct {
    int ns;
    const int32_t *l, *nprim, *off;
    const double *center, *exps, *coefs;
} shells_t;

int64_t oracle_ncart_total(int32_t ns, const int32_t *l) {
    int64_t n = 0;
    for (int s = 0; s < ns; s++)
        n += ncart(l[s]);
    return n;
}

/* One-electron matrices S, T, V (nbf x nbf, row-major) over Cartesian functions */
void oracle_one_electron(int32_t ns, const int32_t *l, const int32_t *nprim, const int32_t *off, const double *center,
                         const double *exps, const double *coefs, int32_t natom, const double *atom_xyz,
                         const double *atom_z, double *S, double *T, double *V) {
    int64_t *f0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t nbf = 0;
    for (int s = 0; s < ns; s++) {
        f0[s] = nbf;
        nbf += ncart(l[s]);
    }
    for (int sa = 0; sa < ns; sa++)
        for (int sb = 0; sb < ns; sb++) {
            const int la = l[sa], lb = l[sb];
            int ax[6], ay[6], az[6], bx[6], by[6], bz[6];
            cart_powers(la, ax, ay, az);
            cart_powers(lb, bx, by, bz);
            const double *A = center + 3 * sa, *B = center + 3 * sb;
            for (int ia = 0; ia < ncart(la); ia++)
                for (int ib = 0; ib < ncart(lb); ib++) {
                    double s_ = 0.0, t_ = 0.0, v_ = 0.0;
                    for (int pa = 0; pa < nprim[sa]; pa++)
                        for (int pb = 0; pb < nprim[sb]; pb++) {
                            const double a = exps[off[sa] + pa], b = exps[off[sb] + pb];
                            const double ca = coefs[off[sa] + pa], cb = coefs[off[sb] + pb];
                            const double p = a + b;
                            double Ex[EI][EJ][ET], Ey[EI][EJ][ET], Ez[EI][EJ][ET];
                            hermite(la, lb + 2, a, b, A[0] - B[0], Ex);
                            hermite(la, lb + 2, a, b, A[1] - B[1], Ey);
                            hermite(la, lb + 2, a, b, A[2] - B[2], Ez);
                            const double sp = sqrt(M_PI / p);
                            const double sx = Ex[ax[ia]][bx[ib]][0] * sp, sy = Ey[ay[ia]][by[ib]][0] * sp,
                                         sz = Ez[az[ia]][bz[ib]][0] * sp;
                            /* 1D kinetic pieces: -1/2 d^2/dx^2 acting on the ket */
                            const int jx = bx[ib], jy = by[ib], jz = bz[ib];
                            const double kx = (4 * b * b * Ex[ax[ia]][jx + 2][0] - 2 * b * (2 * jx + 1) * Ex[ax[ia]][jx][0] +
                                               (jx > 1 ? jx * (jx - 1) * Ex[ax[ia]][jx - 2][0] : 0.0)) * sp;
                            const double ky = (4 * b * b * Ey[ay[ia]][jy + 2][0] - 2 * b * (2 * jy + 1) * Ey[ay[ia]][jy][0] +
                                               (jy > 1 ? jy * (jy - 1) * Ey[ay[ia]][jy - 2][0] : 0.0)) * sp;
                            const double kz = (4 * b * b * Ez[az[ia]][jz + 2][0] - 2 * b * (2 * jz + 1) * Ez[az[ia]][jz][0] +
                                               (jz > 1 ? jz * (jz - 1) * Ez[az[ia]][jz - 2][0] : 0.0)) * sp;
                            const double cc = ca * cb;
                            s_ += cc * sx * sy * sz;
                            t_ += cc * (-0.5) * (kx * sy * sz + sx * ky * sz + sx * sy * kz);
                            /* nuclear attraction */
                            const double P[3] = {(a * A[0] + b * B[0]) / p, (a * A[1] + b * B[1]) / p,
                                                 (a * A[2] + b * B[2]) / p};
                            const int L = la + lb;
                            for (int c = 0; c < natom; c++) {
                                double R[TMAX][TMAX][TMAX];
                                hermite_coulomb(L, p, P[0] - atom_xyz[3 * c], P[1] - atom_xyz[3 * c + 1],
                                                P[2] - atom_xyz[3 * c + 2], R);
                                double sum = 0.0;
                                for (int t = 0; t <= ax[ia] + bx[ib]; t++)
                                    for (int u = 0; u <= ay[ia] + by[ib]; u++)
                                        for (int w = 0; w <= az[ia] + bz[ib]; w++)
                                            sum += Ex[ax[ia]][bx[ib]][t] * Ey[ay[ia]][by[ib]][u] *
                                                   Ez[az[ia]][bz[ib]][w] * R[t][u][w];
                                v_ += cc * (-atom_z[c]) * 2.0 * M_PI / p * sum;
                            }
                        }
                    const int64_t r = f0[sa] + ia, q = f0[sb] + ib;
                    S[r * nbf + q] = s_;
                    T[r * nbf + q] = t_;
                    V[r * nbf + q] = v_;
                }
        }
    free(f0);
}

/* Two-electron integrals (ab|cd) over Cartesian functions, full nbf^4 array (row-major), 8-fold symmetry used */
void oracle_two_electron(int32_t ns, const int32_t *l, const int32_t *nprim, const int32_t *off, const double *center,
                         const double *exps, const double *coefs, double *G) {
    int64_t *f0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)ns);
    int64_t nbf = 0;
    for (int s = 0; s < ns; s++) {
        f0[s] = nbf;
        nbf += ncart(l[s]);
    }
    const int64_t n2 = nbf * nbf, n3 = n2 * nbf;
    for (int sa = 0; sa < ns; sa++)
        for (int sb = 0; sb <= sa; sb++)
            for (int sc = 0; sc < ns; sc++)
                for (int sd = 0; sd <= sc; sd++) {
                    if (sa * (sa + 1) / 2 + sb < sc * (sc + 1) / 2 + sd)
                        continue;
                    const int la = l[sa], lb = l[sb], lc = l[sc], ld = l[sd];
                    int ax[6], ay[6], az[6], bx[6], by[6], bz[6], cx[6], cy[6], cz[6], dx[6], dy[6], dz[6];
                    cart_powers(la, ax, ay, az);
                    cart_powers(lb, bx, by, bz);
                    cart_powers(lc, cx, cy, cz);
                    cart_powers(ld, dx, dy, dz);
                    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
                    double blk[6 * 6 * 6 * 6];
                    memset(blk, 0, sizeof(blk));
                    const double *A = center + 3 * sa, *B = center + 3 * sb, *C = center + 3 * sc,
                                 *D = center + 3 * sd;
                    for (int pa = 0; pa < nprim[sa]; pa++)
                        for (int pb = 0; pb < nprim[sb]; pb++) {
                            const double a = exps[off[sa] + pa], b = exps[off[sb] + pb];
                            const double p = a + b;
                            double Eab[3][EI][EJ][ET];
                            hermite(la, lb, a, b, A[0] - B[0], Eab[0]);
                            hermite(la, lb, a, b, A[1] - B[1], Eab[1]);
                            hermite(la, lb, a, b, A[2] - B[2], Eab[2]);
                            const double P[3] = {(a * A[0] + b * B[0]) / p, (a * A[1] + b * B[1]) / p,
                                                 (a * A[2] + b * B[2]) / p};
                            const double cab = coefs[off[sa] + pa] * coefs[off[sb] + pb];
                            for (int pc = 0; pc < nprim[sc]; pc++)
                                for (int pd = 0; pd < nprim[sd]; pd++) {
                                    const double c = exps[off[sc] + pc], d = exps[off[sd] + pd];
                                    const double q = c + d;
                                    double Ecd[3][EI][EJ][ET];
                                    hermite(lc, ld, c, d, C[0] - D[0], Ecd[0]);
                                    hermite(lc, ld, c, d, C[1] - D[1], Ecd[1]);
                                    hermite(lc, ld, c, d, C[2] - D[2], Ecd[2]);
                                    const double Q[3] = {(c * C[0] + d * D[0]) / q, (c * C[1] + d * D[1]) / q,
                                                         (c * C[2] + d * D[2]) / q};
                                    const double alpha = p * q / (p + q);
                                    const int L = la + lb + lc + ld;
                                    double R[TMAX][TMAX][TMAX];
                                    hermite_coulomb(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
                                    const double pref = cab * coefs[off[sc] + pc] * coefs[off[sd] + pd] * 2.0 *
                                                        pow(M_PI, 2.5) / (p * q * sqrt(p + q));
                                    for (int ia = 0; ia < na; ia++)
                                        for (int ib = 0; ib < nb; ib++)
                                            for (int ic = 0; ic < nc; ic++)
                                                for (int id = 0; id < nd; id++) {
                                                    double sum = 0.0;
                                                    const int mx1 = ax[ia] + bx[ib], my1 = ay[ia] + by[ib],
                                                              mz1 = az[ia] + bz[ib];
                                                    const int mx2 = cx[ic] + dx[id], my2 = cy[ic] + dy[id],
                                                              mz2 = cz[ic] + dz[id];
                                                    for (int t = 0; t <= mx1; t++)
                                                        for (int u = 0; u <= my1; u++)
                                                            for (int v = 0; v <= mz1; v++) {
                                                                const double e1 = Eab[0][ax[ia]][bx[ib]][t] *
                                                                                  Eab[1][ay[ia]][by[ib]][u] *
                                                                                  Eab[2][az[ia]][bz[ib]][v];
                                                                if (e1 == 0.0)
                                                                    continue;
                                                                for (int tt = 0; tt <= mx2; tt++)
                                                                    for (int uu = 0; uu <= my2; uu++)
                                                                        for (int vv = 0; vv <= mz2; vv++) {
                                                                            const double e2 =
                                                                                Ecd[0][cx[ic]][dx[id]][tt] *
                                                                                Ecd[1][cy[ic]][dy[id]][uu] *
                                                                                Ecd[2][cz[ic]][dz[id]][vv];
                                                                            const double sg =
                                                                                ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                                                            sum += e1 * sg * e2 * R[t + tt][u + uu][v + vv];
                                                                        }
                                                            }
                                                    blk[((ia * nb + ib) * nc + ic) * nd + id] += pref * sum;
                                                }
                                }
                        }
                    for (int ia = 0; ia < na; ia++)
                        for (int ib = 0; ib < nb; ib++)
                            for (int ic = 0; ic < nc; ic++)
                                for (int id = 0; id < nd; id++) {
                                    const double x = blk[((ia * nb + ib) * nc + ic) * nd + id];
                                    const int64_t i = f0[sa] + ia, j = f0[sb] + ib, k = f0[sc] + ic, m = f0[sd] + id;
                                    G[i * n3 + j * n2 + k * nbf + m] = x;
                                    G[j * n3 + i * n2 + k * nbf + m] = x;
                                    G[i * n3 + j * n2 + m * nbf + k] = x;
                                    G[j * n3 + i * n2 + m * nbf + k] = x;
                                    G[k * n3 + m * n2 + i * nbf + j] = x;
                                    G[m * n3 + k * n2 + i * nbf + j] = x;
                                    G[k * n3 + m * n2 + j * nbf + i] = x;
                                    G[m * n3 + k * n2 + j * nbf + i] = x;
                                }
                }
    free(f0);
}
