/*
 * CPU restatement of the closed-shell perturbative triples (T) of SCESS.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product path links or calls this
 * file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, as the checker and as the reported CPU baseline.
 *
 * Every function takes the integral and amplitude blocks in Eigen's default
 * ColMajor layout (first index fastest), the layout of the reference's
 * tensors:
 *   e[n]              orbital energies, occupied first (n = o + v)
 *   t1(i,a)           i + o*a
 *   t2(i,j,a,b)       i + o*(j + o*(a + v*b)),  t_ij^ab = t_ji^ba
 *   ovvv(i,a,b,f)     (ia|bf)   i + o*(a + v*(b + v*f))
 *   ovoo(i,a,m,j)     (ia|mj)   i + o*(a + v*(m + o*j))
 *   ovov(i,a,j,b)     (ia|jb)   i + o*(a + v*(j + o*b))
 * These are the chemists'-notation blocks named as in
 * src/Quantum/CC/RCCSD.cpp:31-44 (ovvv(k,d,a,c) = (kd|ac) etc.).
 *
 * Functions:
 *   oracle_t_abc      abc-driven restatement of RCCSD_TriplesCorrection
 *                     (src/Quantum/CC/RCCSD_T.cpp:17-168): repack into P, R,
 *                     S, Q (:27-56), GetW as two matrix products (:61-68),
 *                     GetV (:69-74), Symmetrise (:76-106), unique a>=b>=c
 *                     with degeneracy weights (:108-114, :134-139), r3 and
 *                     the denominator (:141-157), E = 2 * sum (:164-167).
 *                     Threads pull triples from one atomic counter like
 *                     ParallelForDynamic (src/Tensor/TensorParallel.h:58-87);
 *                     unlike the reference, per-triple energies are summed in
 *                     triple order with a Neumaier sum, so the result does
 *                     not depend on the thread count.
 *   oracle_t_ijk      ijk-driven form of the same functional: for each unique
 *                     occupied triple i>=j>=k the full v^3 block of W, V and
 *                     r3 acting on the virtual indices; E = sum 2/deg_ijk F_ijk.
 *   oracle_t_brute    brute-force 6-index sum E = 1/3 sum_{ijkabc} W r3(Y)/D
 *                     over all o^3 v^3 tuples (tiny sizes only).
 *   oracle_t_grad     FP64 backward of the ijk-driven form: dE/d(ovvv),
 *                     dE/d(ovoo), dE/d(ovov), dE/dt1, dE/dt2, treating every
 *                     array entry as an independent variable read exactly as
 *                     the forward reads it.
 *
 * Parity pinning (tests/test_oracle.py): the three forward forms against one
 * another, against an independent spin-orbital (T) written from the textbook
 * formula, E = 0 for o = 1 or v = 1, invariance under a uniform shift of the
 * orbital energies and under relabelling, E proportional to s^2 when the
 * integrals are scaled by s with t1 = 0, and oracle_t_grad against central
 * differences of oracle_t_ijk.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t idx_t;

/* ---------------------------------------------------------------- sums */
typedef struct {
    double s, c;
} nsum_t;

static inline void nsum_add(nsum_t *k, double x) {
    const double t = k->s + x;
    if (fabs(k->s) >= fabs(x))
        k->c += (k->s - t) + x;
    else
        k->c += (x - t) + k->s;
    k->s = t;
}

static double nsum_array(const double *x, idx_t n) {
    nsum_t k = {0.0, 0.0};
    for (idx_t i = 0; i < n; i++)
        nsum_add(&k, x[i]);
    return k.s + k.c;
}

/* ------------------------------------------------ abc-driven restatement */
typedef struct {
    idx_t o, v;
    const double *e, *t1, *t2;
    double *P, *R, *S, *Q;
    const idx_t *items; /* indices into the a>=b>=c enumeration */
    idx_t nitems;
    const int32_t *trip; /* 3 * ntriples: (a, b, c) of each unique triple */
    double *item_energy;
    atomic_llong next;
} abc_job_t;

/*
 * Optional BLAS dgemm for GetW's two products: the CBLAS interface with 64-bit integers (as exported
 * by an ILP64 OpenBLAS), set by oracle_set_dgemm. The reference runs these products through BLAS
 * (Eigen with EIGEN_USE_BLAS, cmake/QuantumSCESS.cmake:79; RCCSD_T.cpp:64-67); NULL keeps the plain
 * loops below. The caller makes the BLAS single-threaded: the triples are spread over threads here.
 */
typedef void (*cblas_dgemm64_t)(int layout, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                                const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                                int64_t ldc);
static cblas_dgemm64_t g_dgemm = NULL;
enum { CBLAS_COL_MAJOR = 102, CBLAS_NO_TRANS = 111 };

void oracle_set_dgemm(void *fn) { g_dgemm = (cblas_dgemm64_t)fn; }

/* C(o x o^2) = A(o x v) B(v x o^2), all ColMajor (RCCSD_T.cpp:63-65) */
static void gemm_w1(idx_t o, idx_t v, const double *restrict A, const double *restrict B, double *restrict C) {
    const idx_t o2 = o * o;
    if (g_dgemm) {
        g_dgemm(CBLAS_COL_MAJOR, CBLAS_NO_TRANS, CBLAS_NO_TRANS, o, o2, v, 1.0, A, o, B, v, 0.0, C, o);
        return;
    }
    idx_t n = 0;
    for (; n + 4 <= o2; n += 4) {
        double *c0 = C + o * n, *c1 = c0 + o, *c2 = c1 + o, *c3 = c2 + o;
        for (idx_t i = 0; i < o; i++)
            c0[i] = c1[i] = c2[i] = c3[i] = 0.0;
        const double *b0 = B + v * n, *b1 = b0 + v, *b2 = b1 + v, *b3 = b2 + v;
        for (idx_t f = 0; f < v; f++) {
            const double *a = A + o * f;
            const double x0 = b0[f], x1 = b1[f], x2 = b2[f], x3 = b3[f];
            for (idx_t i = 0; i < o; i++) {
                const double ai = a[i];
                c0[i] += ai * x0;
                c1[i] += ai * x1;
                c2[i] += ai * x2;
                c3[i] += ai * x3;
            }
        }
    }
    for (; n < o2; n++) {
        double *c0 = C + o * n;
        for (idx_t i = 0; i < o; i++)
            c0[i] = 0.0;
        const double *b0 = B + v * n;
        for (idx_t f = 0; f < v; f++) {
            const double *a = A + o * f;
            for (idx_t i = 0; i < o; i++)
                c0[i] += a[i] * b0[f];
        }
    }
}

/* C(o^2 x o) -= A(o^2 x o) B(o x o), all ColMajor (RCCSD_T.cpp:66-67) */
static void gemm_w2(idx_t o, const double *restrict A, const double *restrict B, double *restrict C) {
    const idx_t o2 = o * o;
    if (g_dgemm) {
        g_dgemm(CBLAS_COL_MAJOR, CBLAS_NO_TRANS, CBLAS_NO_TRANS, o2, o, o, -1.0, A, o2, B, o, 1.0, C, o2);
        return;
    }
    for (idx_t k = 0; k < o; k++) {
        double *c = C + o2 * k;
        for (idx_t m = 0; m < o; m++) {
            const double b = B[m + o * k];
            const double *a = A + o2 * m;
            for (idx_t r = 0; r < o2; r++)
                c[r] -= a[r] * b;
        }
    }
}

static void get_w(const abc_job_t *J, idx_t a, idx_t b, idx_t c, double *out) {
    const idx_t o = J->o, v = J->v, o2 = o * o, o3 = o2 * o;
    gemm_w1(o, v, J->P + o * v * (a + v * b), J->Q + v * o2 * c, out);
    gemm_w2(o, J->R + o3 * a, J->t2 + o2 * (b + v * c), out);
}

static void get_v(const abc_job_t *J, idx_t a, idx_t b, idx_t c, double *out) {
    const idx_t o = J->o, v = J->v, o2 = o * o;
    const double *Sab = J->S + o2 * (a + v * b);
    for (idx_t k = 0; k < o; k++) {
        const double s = 0.5 * J->t1[k + o * c];
        for (idx_t r = 0; r < o2; r++)
            out[r + o2 * k] = s * Sab[r];
    }
}

typedef void (*getter_t)(const abc_job_t *, idx_t, idx_t, idx_t, double *);

/* X_ijk^abc + X_ikj^acb + X_jik^bac + X_jki^bca + X_kij^cab + X_kji^cba (RCCSD_T.cpp:76-106) */
static void symmetrise(const abc_job_t *J, idx_t a, idx_t b, idx_t c, getter_t get, double *out, double *scr) {
    const idx_t o = J->o, o2 = o * o;
#define ID3(i, j, k) ((i) + o * (j) + o2 * (k))
    get(J, a, b, c, out);
    get(J, a, c, b, scr);
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++)
                out[ID3(i, j, k)] += scr[ID3(i, k, j)];
    get(J, b, a, c, scr);
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++)
                out[ID3(i, j, k)] += scr[ID3(j, i, k)];
    get(J, b, c, a, scr);
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++)
                out[ID3(i, j, k)] += scr[ID3(j, k, i)];
    get(J, c, a, b, scr);
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++)
                out[ID3(i, j, k)] += scr[ID3(k, i, j)];
    get(J, c, b, a, scr);
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++)
                out[ID3(i, j, k)] += scr[ID3(k, j, i)];
}

static double abc_triple_energy(const abc_job_t *J, idx_t a, idx_t b, idx_t c, double *buf) {
    const idx_t o = J->o, o2 = o * o, o3 = o2 * o;
    double *W = buf, *V = W + o3, *Y = V + o3, *scr = Y + o3;
    symmetrise(J, a, b, c, get_w, W, scr);
    symmetrise(J, a, b, c, get_v, V, scr);
    double deg = 1.0;
    if (a == c)
        deg = 6.0;
    else if (a == b || b == c)
        deg = 2.0;
    for (idx_t n = 0; n < o3; n++)
        Y[n] = W[n] + V[n];
    const double *e = J->e;
    const double eVirt = e[o + a] + e[o + b] + e[o + c];
    double energy = 0.0;
    for (idx_t k = 0; k < o; k++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t i = 0; i < o; i++) {
                const double r3 = 4.0 * Y[ID3(i, j, k)] + Y[ID3(k, i, j)] + Y[ID3(j, k, i)] -
                                  2.0 * Y[ID3(k, j, i)] - 2.0 * Y[ID3(i, k, j)] - 2.0 * Y[ID3(j, i, k)];
                const double d3 = (e[i] + e[j] + e[k] - eVirt) * deg;
                energy += W[ID3(i, j, k)] * r3 / d3;
            }
#undef ID3
    return energy;
}

static void *abc_worker(void *arg) {
    abc_job_t *J = (abc_job_t *)arg;
    const idx_t o3 = J->o * J->o * J->o;
    double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)(o3 > 0 ? o3 : 1));
    for (;;) {
        const idx_t it = (idx_t)atomic_fetch_add(&J->next, 1);
        if (it >= J->nitems)
            break;
        const idx_t t = J->items ? J->items[it] : it;
        J->item_energy[it] =
            abc_triple_energy(J, J->trip[3 * t], J->trip[3 * t + 1], J->trip[3 * t + 2], buf);
    }
    free(buf);
    return NULL;
}

int64_t oracle_num_unique_triples(int64_t v) { return v * (v + 1) * (v + 2) / 6; }

/* Repacked inputs of the abc-driven restatement (RCCSD_T.cpp:27-56), reusable across evaluations. */
typedef struct {
    abc_job_t J;
    int32_t *trip;
    idx_t ntrip;
} abc_handle_t;

void *oracle_t_abc_prepare(int64_t o, int64_t v, const double *e, const double *t1, const double *t2,
                           const double *ovvv, const double *ovoo, const double *ovov) {
    const idx_t o2 = o * o, o3 = o2 * o;
    abc_handle_t *H = (abc_handle_t *)calloc(1, sizeof(abc_handle_t));
    abc_job_t *J = &H->J;
    J->o = o;
    J->v = v;
    J->e = e;
    J->t1 = t1;
    J->t2 = t2;
    J->P = (double *)malloc(sizeof(double) * (size_t)(o * v * v * v));
    J->R = (double *)malloc(sizeof(double) * (size_t)(o3 * v));
    J->S = (double *)malloc(sizeof(double) * (size_t)(o2 * v * v));
    J->Q = (double *)malloc(sizeof(double) * (size_t)(v * o2 * v));
    /* P(i,f,a,b) = (ia|bf), R(i,j,m,a) = (ia|mj), S(i,j,a,b) = (ia|jb), Q(f,j,k,c) = t_kj^cf (RCCSD_T.cpp:27-56) */
    for (idx_t b = 0; b < v; b++)
        for (idx_t a = 0; a < v; a++)
            for (idx_t f = 0; f < v; f++)
                for (idx_t i = 0; i < o; i++)
                    J->P[i + o * (f + v * (a + v * b))] = ovvv[i + o * (a + v * (b + v * f))];
    for (idx_t a = 0; a < v; a++)
        for (idx_t m = 0; m < o; m++)
            for (idx_t j = 0; j < o; j++)
                for (idx_t i = 0; i < o; i++)
                    J->R[i + o * (j + o * (m + o * a))] = ovoo[i + o * (a + v * (m + o * j))];
    for (idx_t b = 0; b < v; b++)
        for (idx_t a = 0; a < v; a++)
            for (idx_t j = 0; j < o; j++)
                for (idx_t i = 0; i < o; i++)
                    J->S[i + o * (j + o * (a + v * b))] = ovov[i + o * (a + v * (j + o * b))];
    for (idx_t c = 0; c < v; c++)
        for (idx_t k = 0; k < o; k++)
            for (idx_t j = 0; j < o; j++)
                for (idx_t f = 0; f < v; f++)
                    J->Q[f + v * (j + o * (k + o * c))] = t2[k + o * (j + o * (c + v * f))];
    H->ntrip = oracle_num_unique_triples(v);
    H->trip = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(H->ntrip > 0 ? H->ntrip : 1));
    idx_t t = 0;
    for (idx_t a = 0; a < v; a++)
        for (idx_t b = 0; b <= a; b++)
            for (idx_t c = 0; c <= b; c++) {
                H->trip[3 * t] = (int32_t)a;
                H->trip[3 * t + 1] = (int32_t)b;
                H->trip[3 * t + 2] = (int32_t)c;
                t++;
            }
    J->trip = H->trip;
    return H;
}

void oracle_t_abc_free(void *h) {
    abc_handle_t *H = (abc_handle_t *)h;
    if (!H)
        return;
    free(H->J.P);
    free(H->J.R);
    free(H->J.S);
    free(H->J.Q);
    free(H->trip);
    free(H);
}

/*
 * E(T) of the abc-driven restatement from prepared inputs. items == NULL
 * evaluates every unique triple; otherwise only the listed ones (indices into
 * the enumeration a = 0..v-1, b = 0..a, c = 0..b of RCCSD_T.cpp:109-114), and
 * the return value is 2 * (sum over those triples). item_energy (optional)
 * receives the per-triple energies before the factor 2.
 */
double oracle_t_abc_eval(void *h, const int64_t *items, int64_t nitems, int nthreads, double *item_energy) {
    abc_handle_t *H = (abc_handle_t *)h;
    abc_job_t *J = &H->J;
    J->items = items;
    J->nitems = items ? nitems : H->ntrip;
    double *energies = item_energy ? item_energy : (double *)malloc(sizeof(double) * (size_t)(J->nitems + 1));
    J->item_energy = energies;
    atomic_init(&J->next, 0);
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > J->nitems)
        nthreads = (int)(J->nitems > 0 ? J->nitems : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int w = 1; w < nthreads; w++)
        pthread_create(&th[w], NULL, abc_worker, J);
    abc_worker(J);
    for (int w = 1; w < nthreads; w++)
        pthread_join(th[w], NULL);
    free(th);
    const double E = 2.0 * nsum_array(energies, J->nitems);
    if (!item_energy)
        free(energies);
    return E;
}

double oracle_t_abc(int64_t o, int64_t v, const double *e, const double *t1, const double *t2,
                    const double *ovvv, const double *ovoo, const double *ovov, const int64_t *items,
                    int64_t nitems, int nthreads, double *item_energy) {
    void *h = oracle_t_abc_prepare(o, v, e, t1, t2, ovvv, ovoo, ovov);
    const double E = oracle_t_abc_eval(h, items, nitems, nthreads, item_energy);
    oracle_t_abc_free(h);
    return E;
}

/* ------------------------------------------------------ ijk-driven form */
/*
 * The six simultaneous permutations of the pairs (ia), (jb), (kc): term pi
 * takes x = abc[pi0], y = abc[pi1], z = abc[pi2] and p = ijk[pi0],
 * q = ijk[pi1], r = ijk[pi2], contributing
 *   w(x,y,z; p,q,r) = sum_f (px|yf) t_rq^zf - sum_m (px|mq) t_mr^yz
 *   v(x,y,z; p,q,r) = 1/2 (px|qy) t_r^z
 * (the identity term is GetW / GetV of RCCSD_T.cpp:61-74; the others are the
 * permuted calls of Symmetrise, RCCSD_T.cpp:80-105).
 */
static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

typedef struct {
    idx_t o, v;
    const double *e, *t1, *t2, *ovvv, *ovoo, *ovov;
    double *Xr; /* Xr[p][x][y][f] = (px|yf), f fastest */
    double *Tr; /* Tr[q][r][z][f] = t_qr^zf, f fastest */
} ijk_ctx_t;

static void ijk_ctx_init(ijk_ctx_t *C, idx_t o, idx_t v, const double *e, const double *t1, const double *t2,
                         const double *ovvv, const double *ovoo, const double *ovov) {
    C->o = o;
    C->v = v;
    C->e = e;
    C->t1 = t1;
    C->t2 = t2;
    C->ovvv = ovvv;
    C->ovoo = ovoo;
    C->ovov = ovov;
    C->Xr = (double *)malloc(sizeof(double) * (size_t)(o * v * v * v));
    C->Tr = (double *)malloc(sizeof(double) * (size_t)(o * o * v * v));
    for (idx_t p = 0; p < o; p++)
        for (idx_t x = 0; x < v; x++)
            for (idx_t y = 0; y < v; y++)
                for (idx_t f = 0; f < v; f++)
                    C->Xr[((p * v + x) * v + y) * v + f] = ovvv[p + o * (x + v * (y + v * f))];
    for (idx_t q = 0; q < o; q++)
        for (idx_t r = 0; r < o; r++)
            for (idx_t z = 0; z < v; z++)
                for (idx_t f = 0; f < v; f++)
                    C->Tr[((q * o + r) * v + z) * v + f] = t2[q + o * (r + o * (z + v * f))];
}

static void ijk_ctx_free(ijk_ctx_t *C) {
    free(C->Xr);
    free(C->Tr);
}

/* W and V (v^3, index a*v*v + b*v + c) of one occupied triple */
static void ijk_build_wv(const ijk_ctx_t *C, idx_t i, idx_t j, idx_t k, double *W, double *V) {
    const idx_t o = C->o, v = C->v, v3 = v * v * v;
    const idx_t occ[3] = {i, j, k};
    memset(W, 0, sizeof(double) * (size_t)v3);
    memset(V, 0, sizeof(double) * (size_t)v3);
    for (int s = 0; s < 6; s++) {
        const idx_t p = occ[PERM[s][0]], q = occ[PERM[s][1]], r = occ[PERM[s][2]];
        for (idx_t a = 0; a < v; a++)
            for (idx_t b = 0; b < v; b++)
                for (idx_t c = 0; c < v; c++) {
                    const idx_t abc[3] = {a, b, c};
                    const idx_t x = abc[PERM[s][0]], y = abc[PERM[s][1]], z = abc[PERM[s][2]];
                    const double *X = C->Xr + ((p * v + x) * v + y) * v;
                    const double *T = C->Tr + ((r * o + q) * v + z) * v;
                    double w = 0.0;
                    for (idx_t f = 0; f < v; f++)
                        w += X[f] * T[f];
                    for (idx_t m = 0; m < o; m++)
                        w -= C->ovoo[p + o * (x + v * (m + o * q))] * C->t2[m + o * (r + o * (y + v * z))];
                    W[(a * v + b) * v + c] += w;
                    V[(a * v + b) * v + c] += 0.5 * C->ovov[p + o * (x + v * (q + o * y))] * C->t1[r + o * z];
                }
    }
}

/* r3 acting on the virtual indices at fixed ijk: 4Y^abc + Y^bca + Y^cab - 2Y^cba - 2Y^acb - 2Y^bac */
static inline double r3v(const double *Y, idx_t v, idx_t a, idx_t b, idx_t c) {
#define YI(x, y, z) Y[((x) * v + (y)) * v + (z)]
    return 4.0 * YI(a, b, c) + YI(b, c, a) + YI(c, a, b) - 2.0 * YI(c, b, a) - 2.0 * YI(a, c, b) -
           2.0 * YI(b, a, c);
#undef YI
}

static double ijk_block_F(const ijk_ctx_t *C, idx_t i, idx_t j, idx_t k, double *W, double *V, double *Y) {
    const idx_t o = C->o, v = C->v, v3 = v * v * v;
    ijk_build_wv(C, i, j, k, W, V);
    for (idx_t n = 0; n < v3; n++)
        Y[n] = W[n] + V[n];
    const double *e = C->e;
    const double eocc = e[i] + e[j] + e[k];
    nsum_t F = {0.0, 0.0};
    for (idx_t a = 0; a < v; a++)
        for (idx_t b = 0; b < v; b++)
            for (idx_t c = 0; c < v; c++) {
                const double D = eocc - e[o + a] - e[o + b] - e[o + c];
                nsum_add(&F, W[(a * v + b) * v + c] * r3v(Y, v, a, b, c) / D);
            }
    return F.s + F.c;
}

int64_t oracle_num_unique_ijk(int64_t o) { return o * (o + 1) * (o + 2) / 6; }

/* The unique occupied triples i>=j>=k in enumeration order (i outer, k inner), 3 ints each */
void oracle_unique_ijk(int64_t o, int32_t *out) {
    idx_t t = 0;
    for (idx_t i = 0; i < o; i++)
        for (idx_t j = 0; j <= i; j++)
            for (idx_t k = 0; k <= j; k++) {
                out[3 * t] = (int32_t)i;
                out[3 * t + 1] = (int32_t)j;
                out[3 * t + 2] = (int32_t)k;
                t++;
            }
}

typedef struct {
    const ijk_ctx_t *C;
    const int32_t *ijk;
    const int64_t *items;
    idx_t nitems;
    double *F;
    atomic_llong next;
} ijk_job_t;

static void *ijk_worker(void *arg) {
    ijk_job_t *J = (ijk_job_t *)arg;
    const idx_t v3 = J->C->v * J->C->v * J->C->v;
    double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)(v3 > 0 ? v3 : 1));
    for (;;) {
        const idx_t it = (idx_t)atomic_fetch_add(&J->next, 1);
        if (it >= J->nitems)
            break;
        const idx_t t = J->items ? J->items[it] : it;
        const int32_t *q = J->ijk + 3 * t;
        J->F[it] = (q[0] == q[2]) ? 0.0 : ijk_block_F(J->C, q[0], q[1], q[2], buf, buf + v3, buf + 2 * v3);
    }
    free(buf);
    return NULL;
}

/*
 * F_ijk = sum_abc W r3(Y) / D for the listed unique occupied triples (items ==
 * NULL: all of them, in the order of oracle_unique_ijk). Returns
 * sum 2/deg_ijk F_ijk over the listed triples, which is E(T) when all are listed.
 */
double oracle_t_ijk(int64_t o, int64_t v, const double *e, const double *t1, const double *t2, const double *ovvv,
                    const double *ovoo, const double *ovov, const int64_t *items, int64_t nitems, int nthreads,
                    double *F_out) {
    ijk_ctx_t C;
    ijk_ctx_init(&C, o, v, e, t1, t2, ovvv, ovoo, ovov);
    const idx_t nijk = oracle_num_unique_ijk(o);
    int32_t *ijk = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nijk > 0 ? nijk : 1));
    oracle_unique_ijk(o, ijk);
    ijk_job_t J;
    J.C = &C;
    J.ijk = ijk;
    J.items = items;
    J.nitems = items ? nitems : nijk;
    J.F = F_out ? F_out : (double *)malloc(sizeof(double) * (size_t)(J.nitems + 1));
    atomic_init(&J.next, 0);
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > J.nitems)
        nthreads = (int)(J.nitems > 0 ? J.nitems : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int w = 1; w < nthreads; w++)
        pthread_create(&th[w], NULL, ijk_worker, &J);
    ijk_worker(&J);
    for (int w = 1; w < nthreads; w++)
        pthread_join(th[w], NULL);
    free(th);
    nsum_t E = {0.0, 0.0};
    for (idx_t it = 0; it < J.nitems; it++) {
        const idx_t t = items ? items[it] : it;
        const int32_t *q = ijk + 3 * t;
        const double deg = (q[0] == q[2]) ? 6.0 : ((q[0] == q[1] || q[1] == q[2]) ? 2.0 : 1.0);
        nsum_add(&E, 2.0 / deg * J.F[it]);
    }
    if (!F_out)
        free(J.F);
    free(ijk);
    ijk_ctx_free(&C);
    return E.s + E.c;
}

/* ------------------------------------------------------- brute force */
double oracle_t_brute(int64_t o, int64_t v, const double *e, const double *t1, const double *t2, const double *ovvv,
                      const double *ovoo, const double *ovov) {
    const idx_t N = o * o * o * v * v * v;
#define I6(i, j, k, a, b, c) (((((i) * o + (j)) * o + (k)) * v + (a)) * v * v + (b) * v + (c))
    double *w = (double *)calloc((size_t)N, sizeof(double));
    double *vv = (double *)calloc((size_t)N, sizeof(double));
    double *W = (double *)calloc((size_t)N, sizeof(double));
    double *Y = (double *)calloc((size_t)N, sizeof(double));
    /* unsymmetrised w_ijk^abc = sum_f (ia|bf) t_kj^cf - sum_m (ia|mj) t_mk^bc and v = 1/2 (ia|jb) t_k^c */
    for (idx_t i = 0; i < o; i++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t k = 0; k < o; k++)
                for (idx_t a = 0; a < v; a++)
                    for (idx_t b = 0; b < v; b++)
                        for (idx_t c = 0; c < v; c++) {
                            double s = 0.0;
                            for (idx_t f = 0; f < v; f++)
                                s += ovvv[i + o * (a + v * (b + v * f))] * t2[k + o * (j + o * (c + v * f))];
                            for (idx_t m = 0; m < o; m++)
                                s -= ovoo[i + o * (a + v * (m + o * j))] * t2[m + o * (k + o * (b + v * c))];
                            w[I6(i, j, k, a, b, c)] = s;
                            vv[I6(i, j, k, a, b, c)] = 0.5 * ovov[i + o * (a + v * (j + o * b))] * t1[k + o * c];
                        }
    /* simultaneous permutations of the pairs (ia), (jb), (kc) */
    for (idx_t i = 0; i < o; i++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t k = 0; k < o; k++)
                for (idx_t a = 0; a < v; a++)
                    for (idx_t b = 0; b < v; b++)
                        for (idx_t c = 0; c < v; c++) {
                            const idx_t n0 = I6(i, j, k, a, b, c), n1 = I6(i, k, j, a, c, b),
                                        n2 = I6(j, i, k, b, a, c), n3 = I6(j, k, i, b, c, a),
                                        n4 = I6(k, i, j, c, a, b), n5 = I6(k, j, i, c, b, a);
                            W[n0] = w[n0] + w[n1] + w[n2] + w[n3] + w[n4] + w[n5];
                            Y[n0] = W[n0] + vv[n0] + vv[n1] + vv[n2] + vv[n3] + vv[n4] + vv[n5];
                        }
    /* E = 1/3 sum over all tuples of W r3(Y) / D with r3 on the occupied indices */
    nsum_t E = {0.0, 0.0};
    for (idx_t i = 0; i < o; i++)
        for (idx_t j = 0; j < o; j++)
            for (idx_t k = 0; k < o; k++)
                for (idx_t a = 0; a < v; a++)
                    for (idx_t b = 0; b < v; b++)
                        for (idx_t c = 0; c < v; c++) {
                            const double r3 = 4.0 * Y[I6(i, j, k, a, b, c)] + Y[I6(k, i, j, a, b, c)] +
                                              Y[I6(j, k, i, a, b, c)] - 2.0 * Y[I6(k, j, i, a, b, c)] -
                                              2.0 * Y[I6(i, k, j, a, b, c)] - 2.0 * Y[I6(j, i, k, a, b, c)];
                            const double D = e[i] + e[j] + e[k] - e[o + a] - e[o + b] - e[o + c];
                            nsum_add(&E, W[I6(i, j, k, a, b, c)] * r3 / D);
                        }
#undef I6
    free(w);
    free(vv);
    free(W);
    free(Y);
    return (E.s + E.c) / 3.0;
}

/* ------------------------------------------------------------ backward */
/*
 * dE/d(input) for E = sum_{i>=j>=k} 2/deg_ijk sum_abc W r3(Y)/D, Y = W + V.
 * With r3 symmetric: dE/dW = w (r3(Y)/D + r3(W/D)) and dE/dV = w r3(W/D),
 * w = 2/deg_ijk, then through the six permutation terms to the array entries
 * each of them reads. Gradients are accumulated (+=) into the caller's arrays,
 * which use the input layouts. items (nitems entries) selects unique occupied
 * triples by their index in oracle_unique_ijk's order (NULL: all of them), so
 * the result is the gradient of the F-weighted sum over those triples only.
 * The f-sums run on the f-contiguous copies Xr, Tr with f-contiguous gradient
 * buffers dXr, dTr, added into d_ovvv and d_t2 at the end. Single-threaded.
 */
void oracle_t_grad_items(int64_t o, int64_t v, const double *e, const double *t1, const double *t2,
                         const double *ovvv, const double *ovoo, const double *ovov, const int64_t *items,
                         int64_t nitems, double *d_ovvv, double *d_ovoo, double *d_ovov, double *d_t1,
                         double *d_t2) {
    ijk_ctx_t C;
    ijk_ctx_init(&C, o, v, e, t1, t2, ovvv, ovoo, ovov);
    const idx_t v3 = v * v * v;
    double *W = (double *)malloc(sizeof(double) * (size_t)v3);
    double *V = (double *)malloc(sizeof(double) * (size_t)v3);
    double *Y = (double *)malloc(sizeof(double) * (size_t)v3);
    double *WD = (double *)malloc(sizeof(double) * (size_t)v3);
    double *GW = (double *)malloc(sizeof(double) * (size_t)v3);
    double *GV = (double *)malloc(sizeof(double) * (size_t)v3);
    double *dXr = (double *)calloc((size_t)(o * v3), sizeof(double));    /* [p][x][y][f] */
    double *dTr = (double *)calloc((size_t)(o * o * v * v), sizeof(double)); /* [q][r][z][f] */
    const idx_t nijk = oracle_num_unique_ijk(o);
    int32_t *ijk = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nijk > 0 ? nijk : 1));
    oracle_unique_ijk(o, ijk);
    const idx_t nit = items ? nitems : nijk;
    for (idx_t it = 0; it < nit; it++) {
        const int32_t *qq = ijk + 3 * (items ? items[it] : it);
        const idx_t i = qq[0], j = qq[1], k = qq[2];
        if (i == k)
            continue;
        const double wgt = (i == j || j == k) ? 1.0 : 2.0;
        ijk_build_wv(&C, i, j, k, W, V);
        const double eocc = e[i] + e[j] + e[k];
        for (idx_t a = 0; a < v; a++)
            for (idx_t b = 0; b < v; b++)
                for (idx_t c = 0; c < v; c++) {
                    const idx_t n = (a * v + b) * v + c;
                    const double D = eocc - e[o + a] - e[o + b] - e[o + c];
                    Y[n] = W[n] + V[n];
                    WD[n] = W[n] / D;
                }
        for (idx_t a = 0; a < v; a++)
            for (idx_t b = 0; b < v; b++)
                for (idx_t c = 0; c < v; c++) {
                    const idx_t n = (a * v + b) * v + c;
                    const double D = eocc - e[o + a] - e[o + b] - e[o + c];
                    const double rw = r3v(WD, v, a, b, c);
                    GW[n] = wgt * (r3v(Y, v, a, b, c) / D + rw);
                    GV[n] = wgt * rw;
                }
        const idx_t occ[3] = {i, j, k};
        for (int s = 0; s < 6; s++) {
            const idx_t p = occ[PERM[s][0]], q = occ[PERM[s][1]], r = occ[PERM[s][2]];
            /* term s reads (x, y, z) = (abc[PERM[s][0]], abc[PERM[s][1]], abc[PERM[s][2]]) */
            for (idx_t x = 0; x < v; x++)
                for (idx_t y = 0; y < v; y++) {
                    const double *X = C.Xr + ((p * v + x) * v + y) * v;
                    double *dX = dXr + ((p * v + x) * v + y) * v;
                    for (idx_t z = 0; z < v; z++) {
                        idx_t abc[3];
                        abc[PERM[s][0]] = x;
                        abc[PERM[s][1]] = y;
                        abc[PERM[s][2]] = z;
                        const idx_t n = (abc[0] * v + abc[1]) * v + abc[2];
                        const double g = GW[n];
                        const double *T = C.Tr + ((r * o + q) * v + z) * v;
                        double *dT = dTr + ((r * o + q) * v + z) * v;
                        for (idx_t f = 0; f < v; f++) {
                            dX[f] += g * T[f];
                            dT[f] += g * X[f];
                        }
                        for (idx_t m = 0; m < o; m++) {
                            const idx_t iR = p + o * (x + v * (m + o * q));
                            const idx_t iT = m + o * (r + o * (y + v * z));
                            d_ovoo[iR] -= g * t2[iT];
                            d_t2[iT] -= g * ovoo[iR];
                        }
                        const double gv = 0.5 * GV[n];
                        const idx_t iS = p + o * (x + v * (q + o * y));
                        d_ovov[iS] += gv * t1[r + o * z];
                        d_t1[r + o * z] += gv * ovov[iS];
                    }
                }
        }
    }
    for (idx_t p = 0; p < o; p++)
        for (idx_t x = 0; x < v; x++)
            for (idx_t y = 0; y < v; y++)
                for (idx_t f = 0; f < v; f++)
                    d_ovvv[p + o * (x + v * (y + v * f))] += dXr[((p * v + x) * v + y) * v + f];
    for (idx_t q = 0; q < o; q++)
        for (idx_t r = 0; r < o; r++)
            for (idx_t z = 0; z < v; z++)
                for (idx_t f = 0; f < v; f++)
                    d_t2[q + o * (r + o * (z + v * f))] += dTr[((q * o + r) * v + z) * v + f];
    free(ijk);
    free(dXr);
    free(dTr);
    free(W);
    free(V);
    free(Y);
    free(WD);
    free(GW);
    free(GV);
    ijk_ctx_free(&C);
}

void oracle_t_grad(int64_t o, int64_t v, const double *e, const double *t1, const double *t2, const double *ovvv,
                   const double *ovoo, const double *ovov, double *d_ovvv, double *d_ovoo, double *d_ovov,
                   double *d_t1, double *d_t2) {
    oracle_t_grad_items(o, v, e, t1, t2, ovvv, ovoo, ovov, NULL, 0, d_ovvv, d_ovoo, d_ovov, d_t1, d_t2);
}
