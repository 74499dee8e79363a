"""CPU restatement of the closed-shell CCSD solver of SCESS (numpy).

TEST INFRASTRUCTURE ONLY (the checker of the GPU residual engine); the product
path never imports it.

Follows src/Quantum/CC/RCCSD.cpp:
  make_blocks             MakeIntegralBlocks        :47-72
  plus_transpose          PlusTranspose             :75-77
  make_tau                MakeTau                   :80-83
  correlation_energy      CorrelationEnergy         :85-89
  coupling_energy         CouplingEnergy            :91-98
  update_amplitudes       UpdateAmplitudes          :101-237 (every einsum string
                          and sign as in the reference's comments and Einsum calls)
  AmplitudeDIIS           AmplitudeDIIS             :240-279 (fullPivLu -> an LU solve
                          of the same bordered system)
  solve_rccsd_diis        SolveRCCSD_DIIS           :295-394 (thresholds :302-307, MP2
                          guess :333-339, DIIS from iteration 2 :354-359)

Integrals come in chemists' notation as a full (p,q,r,s) array over all n
orbitals (occupied first), C order; tiny sizes only (the v^4 block is formed).
"""

from __future__ import annotations

import numpy as np

ENERGY_CONVERGENCE = 1.0e-10
AMPLITUDE_CONVERGENCE = 1.0e-8
MAX_ITERATIONS = 100
DIIS_START = 2
MAX_DIIS_VECS = 8


def make_blocks(eri_chem: np.ndarray, o: int) -> dict:
    """The blocks of RCCSD.cpp:47-72 from (pq|rs) (C-order array indexed p,q,r,s)."""
    O = slice(0, o)
    V = slice(o, None)
    g = eri_chem
    ints = {}
    ints["ovov"] = np.ascontiguousarray(g[O, V, O, V])          # (kc|ld) indexed (k,c,l,d)
    ints["oovv"] = np.ascontiguousarray(g[O, O, V, V])          # (ki|ac) indexed (k,i,a,c)
    ints["ovvo"] = np.ascontiguousarray(g[O, V, V, O])          # (kc|ai) indexed (k,c,a,i)
    ints["ovvv"] = np.ascontiguousarray(g[O, V, V, V])          # (kd|ac) indexed (k,d,a,c)
    ints["ovoo"] = np.ascontiguousarray(g[O, V, O, O])          # (lc|ki) indexed (l,c,k,i)
    ints["oooo_phys"] = np.ascontiguousarray(g[O, O, O, O].transpose(0, 2, 1, 3))  # <kl|ij> = (ki|lj)
    ints["vvvv_phys"] = np.ascontiguousarray(g[V, V, V, V].transpose(0, 2, 1, 3))  # <ab|cd> = (ac|bd)
    ints["oovv_phys"] = np.ascontiguousarray(g[O, V, O, V].transpose(0, 2, 1, 3))  # <ij|ab> = (ia|jb)
    ints["L_oovv_phys"] = 2.0 * ints["oovv_phys"] - ints["oovv_phys"].transpose(0, 1, 3, 2)
    ints["L_ovov"] = 2.0 * ints["ovov"] - ints["ovov"].transpose(0, 3, 2, 1)
    return ints


def plus_transpose(x):
    """X_ij^ab + X_ji^ba"""
    return x + x.transpose(1, 0, 3, 2)


def make_tau(t1, t2):
    """tau_ij^ab = t_ij^ab + t_i^a t_j^b"""
    return t2 + np.einsum("ia,jb->ijab", t1, t1)


def correlation_energy(ints, t1, t2):
    return float(np.sum(ints["L_oovv_phys"] * make_tau(t1, t2)))


def coupling_energy(f12, t2):
    if f12 is None:
        return 0.0
    e = f12["energy"]
    return float(np.sum((2.0 * e - e.transpose(0, 1, 3, 2)) * t2))


def update_amplitudes(ints, eia, eijab, t2driver, t1, t2):
    """One Jacobi update t <- R(t)/D (RCCSD.cpp:101-237). Returns (t1new, t2new)."""
    es = np.einsum
    tau = make_tau(t1, t2)
    L_ovov, ovov, oovv, ovvo, ovvv, ovoo = (ints[k] for k in ("L_ovov", "ovov", "oovv", "ovvo", "ovvv", "ovoo"))
    Foo = es("kcld,ilcd->ki", L_ovov, tau)
    Fvv = -es("kcld,klad->ac", L_ovov, tau)
    Fov = es("kcld,ld->kc", L_ovov, t1)

    t1new = es("ac,ic->ia", Fvv, t1)
    t1new -= es("ki,ka->ia", Foo, t1)
    t1new += 2.0 * es("kc,kica->ia", Fov, t2)
    t1new -= es("kc,ikca->ia", Fov, t2)
    Fov_t1 = es("kc,ic->ki", Fov, t1)
    t1new += es("ki,ka->ia", Fov_t1, t1)
    t1new += 2.0 * es("kcai,kc->ia", ovvo, t1)
    t1new -= es("kiac,kc->ia", oovv, t1)
    t1new += 2.0 * es("kdac,ikcd->ia", ovvv, tau)
    t1new -= es("kcad,ikcd->ia", ovvv, tau)
    t1new -= 2.0 * es("lcki,klac->ia", ovoo, tau)
    t1new += es("kcli,klac->ia", ovoo, tau)

    tmp2 = -es("kibc,ka->abic", oovv, t1)
    tmp2 += es("iacb->abic", ovvv)
    t2new = plus_transpose(es("abic,jc->ijab", tmp2, t1))
    tmp2 = es("kcai,jc->akij", ovvo, t1)
    tmp2 += es("iajk->akij", ovoo)
    t2new -= plus_transpose(es("akij,kb->ijab", tmp2, t1))
    t2new += ints["oovv_phys"]
    if t2driver is not None:
        t2new += t2driver

    Loo = Foo + 2.0 * es("lcki,lc->ki", ovoo, t1) - es("kcli,lc->ki", ovoo, t1)
    Lvv = Fvv + 2.0 * es("kdac,kd->ac", ovvv, t1) - es("kcad,kd->ac", ovvv, t1)

    Woooo = ints["oooo_phys"].copy()
    Woooo += es("lcki,jc->klij", ovoo, t1)
    Woooo += es("kclj,ic->klij", ovoo, t1)
    Woooo += es("kcld,ijcd->klij", ovov, tau)

    Wvvvv = ints["vvvv_phys"].copy()
    Wvvvv -= es("kdac,kb->abcd", ovvv, t1)
    Wvvvv -= es("kcbd,ka->abcd", ovvv, t1)

    Wvoov = es("kcai->akic", ovvo).copy()
    Wvoov += es("kcad,id->akic", ovvv, t1)
    Wvoov -= es("kcli,la->akic", ovoo, t1)
    Wvoov -= 0.5 * es("ldkc,ilda->akic", ovov, t2)
    Wvoov -= 0.5 * es("lckd,ilad->akic", ovov, t2)
    Wvoov += es("ldkc,ilad->akic", ovov, t2)
    tmp = es("ldkc,id->lkci", ovov, t1)
    Wvoov -= es("lkci,la->akic", tmp, t1)

    Wvovo = es("kiac->akci", oovv).copy()
    Wvovo += es("kdac,id->akci", ovvv, t1)
    Wvovo -= es("lcki,la->akci", ovoo, t1)
    Wvovo -= 0.5 * es("lckd,ilda->akci", ovov, t2)
    tmp = es("lckd,id->lcki", ovov, t1)
    Wvovo -= es("lcki,la->akci", tmp, t1)

    t2new += es("klij,klab->ijab", Woooo, tau)
    t2new += es("abcd,ijcd->ijab", Wvvvv, tau)
    t2new += plus_transpose(es("ac,ijcb->ijab", Lvv, t2))
    t2new -= plus_transpose(es("ki,kjab->ijab", Loo, t2))
    tmp = 2.0 * es("akic,kjcb->ijab", Wvoov, t2)
    tmp -= es("akci,kjcb->ijab", Wvovo, t2)
    t2new += plus_transpose(tmp)
    tmp = es("akic,kjbc->ijab", Wvoov, t2)
    t2new -= plus_transpose(tmp)
    tmp = es("bkci,kjac->ijab", Wvovo, t2)
    t2new -= plus_transpose(tmp)

    return t1new / eia, t2new / eijab


def solve_full_pivot(A, b):
    """LU with full pivoting (Eigen's fullPivLu): zero pivots give zero components, so
    rank-deficient but consistent systems (e.g. identical DIIS error vectors) still solve."""
    A = np.array(A, dtype=float)
    b = np.array(b, dtype=float)
    n = A.shape[0]
    cp = np.arange(n)
    for k in range(n):
        sub = np.abs(A[k:, k:])
        if sub.max() == 0.0:
            break
        pr, pc = np.unravel_index(np.argmax(sub), sub.shape)
        pr += k
        pc += k
        A[[k, pr]] = A[[pr, k]]
        b[[k, pr]] = b[[pr, k]]
        A[:, [k, pc]] = A[:, [pc, k]]
        cp[[k, pc]] = cp[[pc, k]]
        for i in range(k + 1, n):
            f = A[i, k] / A[k, k]
            A[i, k:] -= f * A[k, k:]
            b[i] -= f * b[k]
    y = np.zeros(n)
    for k in range(n - 1, -1, -1):
        if A[k, k] == 0.0:
            continue
        y[k] = (b[k] - A[k, k + 1:] @ y[k + 1:]) / A[k, k]
    x = np.zeros(n)
    x[cp] = y
    return x


class AmplitudeDIIS:
    """RCCSD.cpp:240-279."""

    def __init__(self, max_vectors=MAX_DIIS_VECS):
        self.max_vectors = max_vectors
        self.amps, self.errs = [], []

    def extrapolate(self, amplitudes, error):
        self.amps.append(amplitudes.copy())
        self.errs.append(error.copy())
        if len(self.errs) > self.max_vectors:
            self.amps.pop(0)
            self.errs.pop(0)
        n = len(self.errs)
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for i in range(n):
            for j in range(n):
                B[i, j] = float(np.dot(self.errs[i], self.errs[j]))
        mx = np.abs(B[:n, :n]).max()
        if mx > 0.0:
            B[:n, :n] /= mx
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        coef = solve_full_pivot(B, rhs)
        out = np.zeros_like(amplitudes)
        for i in range(n):
            out += coef[i] * self.amps[i]
        return out


def denominators(e, o):
    eo, ev = e[:o], e[o:]
    eia = eo[:, None] - ev[None, :]
    eijab = eia[:, None, :, None] + eia[None, :, None, :]
    return eia, eijab


def solve_rccsd_diis(eri_chem, e, o, f12=None, max_iter=MAX_ITERATIONS, callback=None):
    """SolveRCCSD_DIIS (RCCSD.cpp:295-394). Returns dict(E_CCSD, t1, t2, converged, iterations, history)."""
    ints = make_blocks(eri_chem, o)
    v = eri_chem.shape[0] - o
    eia, eijab = denominators(e, o)
    t2driver = f12["residual"] if f12 is not None else None
    t1 = np.zeros((o, v))
    t2 = ints["oovv_phys"].copy()
    if t2driver is not None:
        t2 = t2 + t2driver
    t2 = t2 / eijab
    diis = AmplitudeDIIS()
    E_last = np.finfo(float).max
    hist = []
    converged = False
    it = 0
    for it in range(max_iter):
        prev = np.concatenate([t1.ravel(), t2.ravel()])
        t1, t2 = update_amplitudes(ints, eia, eijab, t2driver, t1, t2)
        upd = np.concatenate([t1.ravel(), t2.ravel()])
        change = upd - prev
        if it >= DIIS_START:
            x = diis.extrapolate(upd, change)
            t1 = x[: t1.size].reshape(t1.shape)
            t2 = x[t1.size:].reshape(t2.shape)
        else:
            diis.extrapolate(upd, change)
        E = correlation_energy(ints, t1, t2) + coupling_energy(f12, t2)
        dE = abs(E - E_last)
        rms = float(np.linalg.norm(change) / np.sqrt(change.size))
        E_last = E
        hist.append((E, dE, rms))
        if callback:
            callback(it, E, dE, rms)
        if dE < ENERGY_CONVERGENCE and rms < AMPLITUDE_CONVERGENCE:
            converged = True
            break
    return dict(E_CCSD=E_last, t1=t1, t2=t2, converged=converged, iterations=it + 1, history=hist)
