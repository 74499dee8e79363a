"""Independent spin-orbital CCSD from the textbook equations (test helper).

Stanton & Gauss, J. Chem. Phys. 94, 4334 (1991) intermediates, as used in
T. D. Crawford's programming projects and in the reference's spin-orbital
path (src/Quantum/CC/CCSD.cpp). Canonical orbitals (diagonal Fock), plain
Jacobi iterations with DIIS on the flattened amplitudes. Tiny sizes only.
"""

import numpy as np


def spin_integrals(eri_chem, e, o):
    n = eri_chem.shape[0]
    N = 2 * n
    sp = np.arange(N) // 2
    sg = np.arange(N) % 2
    same = (sg[:, None] == sg[None, :]).astype(float)
    g = eri_chem[np.ix_(sp, sp, sp, sp)]
    phys = g.transpose(0, 2, 1, 3) * same[:, None, :, None] * same[None, :, None, :]
    asym = phys - phys.transpose(0, 1, 3, 2)
    # occupied spin orbitals first
    order = np.concatenate([np.arange(2 * o), np.arange(2 * o, N)])
    return asym[np.ix_(order, order, order, order)], e[sp][order]


def ccsd_energy(g, e, O, tol=1e-11, maxit=300):
    N = g.shape[0]
    o, v = slice(0, O), slice(O, N)
    eo, ev = e[:O], e[O:]
    Dia = eo[:, None] - ev[None, :]
    Dijab = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    t1 = np.zeros((O, N - O))
    t2 = g[o, o, v, v] / Dijab
    es = np.einsum
    hist_x, hist_e = [], []
    Eold = 0.0
    for it in range(maxit):
        tau_t = t2 + 0.5 * (es("ia,jb->ijab", t1, t1) - es("ib,ja->ijab", t1, t1))
        tau = t2 + es("ia,jb->ijab", t1, t1) - es("ib,ja->ijab", t1, t1)
        Fae = es("mf,mafe->ae", t1, g[o, v, v, v]) - 0.5 * es("mnaf,mnef->ae", tau_t, g[o, o, v, v])
        Fmi = es("ne,mnie->mi", t1, g[o, o, o, v]) + 0.5 * es("inef,mnef->mi", tau_t, g[o, o, v, v])
        Fme = es("nf,mnef->me", t1, g[o, o, v, v])
        Wmnij = g[o, o, o, o] + es("je,mnie->mnij", t1, g[o, o, o, v]) - es("ie,mnje->mnij", t1, g[o, o, o, v]) \
            + 0.25 * es("ijef,mnef->mnij", tau, g[o, o, v, v])
        Wabef = g[v, v, v, v] - es("mb,amef->abef", t1, g[v, o, v, v]) + es("ma,bmef->abef", t1, g[v, o, v, v]) \
            + 0.25 * es("mnab,mnef->abef", tau, g[o, o, v, v])
        Wmbej = g[o, v, v, o] + es("jf,mbef->mbej", t1, g[o, v, v, v]) - es("nb,mnej->mbej", t1, g[o, o, v, o]) \
            - es("jnfb,mnef->mbej", 0.5 * t2 + es("jf,nb->jnfb", t1, t1), g[o, o, v, v])
        r1 = es("ie,ae->ia", t1, Fae) - es("ma,mi->ia", t1, Fmi) + es("imae,me->ia", t2, Fme) \
            - es("nf,naif->ia", t1, g[o, v, o, v]) - 0.5 * es("imef,maef->ia", t2, g[o, v, v, v]) \
            - 0.5 * es("mnae,nmei->ia", t2, g[o, o, v, o])
        tmp = es("ijae,be->ijab", t2, Fae - 0.5 * es("mb,me->be", t1, Fme))
        r2 = g[o, o, v, v] + tmp - tmp.transpose(0, 1, 3, 2)
        tmp = es("imab,mj->ijab", t2, Fmi + 0.5 * es("je,me->mj", t1, Fme))
        r2 -= tmp - tmp.transpose(1, 0, 2, 3)
        r2 += 0.5 * es("mnab,mnij->ijab", tau, Wmnij) + 0.5 * es("ijef,abef->ijab", tau, Wabef)
        tmp = es("imae,mbej->ijab", t2, Wmbej) - es("ie,ma,mbej->ijab", t1, t1, g[o, v, v, o])
        r2 += tmp - tmp.transpose(1, 0, 2, 3) - tmp.transpose(0, 1, 3, 2) + tmp.transpose(1, 0, 3, 2)
        tmp = es("ie,abej->ijab", t1, g[v, v, v, o])
        r2 += tmp - tmp.transpose(1, 0, 2, 3)
        tmp = es("ma,mbij->ijab", t1, g[o, v, o, o])
        r2 -= tmp - tmp.transpose(0, 1, 3, 2)
        t1n, t2n = r1 / Dia, r2 / Dijab
        x = np.concatenate([t1n.ravel(), t2n.ravel()])
        err = x - np.concatenate([t1.ravel(), t2.ravel()])
        hist_x.append(x)
        hist_e.append(err)
        if len(hist_x) > 8:
            hist_x.pop(0)
            hist_e.pop(0)
        if it >= 2:
            n = len(hist_x)
            B = -np.ones((n + 1, n + 1))
            B[n, n] = 0.0
            for a_ in range(n):
                for b_ in range(n):
                    B[a_, b_] = hist_e[a_] @ hist_e[b_]
            B[:n, :n] /= np.abs(B[:n, :n]).max()
            rhs = np.zeros(n + 1)
            rhs[n] = -1.0
            c = np.linalg.solve(B, rhs)
            x = sum(c[k] * hist_x[k] for k in range(n))
        t1 = x[: t1.size].reshape(t1.shape)
        t2 = x[t1.size:].reshape(t2.shape)
        E = 0.25 * es("ijab,ijab->", g[o, o, v, v], t2) + 0.5 * es("ijab,ia,jb->", g[o, o, v, v], t1, t1)
        if abs(E - Eold) < tol and np.linalg.norm(err) < 1e-9:
            return E
        Eold = E
    raise RuntimeError("spin-orbital CCSD did not converge")
