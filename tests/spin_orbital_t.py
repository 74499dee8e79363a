"""Independent spin-orbital (T) from the textbook formula (test helper).

Written from the standard spin-orbital expressions (e.g. T. D. Crawford's
programming projects, and the spin-orbital path of the reference,
src/Quantum/CC/CCSD_T.cpp:89-228):

  D t_ijk^abc(d) = P(i/jk) P(a/bc) t_i^a <jk||bc>
  D t_ijk^abc(c) = P(i/jk) P(a/bc) [ sum_e t_jk^ae <ei||bc> - sum_m t_im^bc <ma||jk> ]
  E(T) = 1/36 sum_ijkabc t(c) D (t(c) + t(d))

with P(i/jk) f(ijk) = f(ijk) - f(jik) - f(kji). The closed-shell amplitudes map
to spin orbitals as t_IJ^AB = t_ij^ab d(IA) d(JB) - t_ij^ba d(IB) d(JA) and
t_I^A = t_i^a d(IA). Tiny sizes only (6-index spin-orbital arrays).
"""

import numpy as np


def spin_orbital_triples(e, t1, t2, eri_chem, o):
    n = eri_chem.shape[0]
    v = n - o
    # spin-orbital index P = 2 p + s
    N = 2 * n
    sp = np.arange(N) // 2
    sg = np.arange(N) % 2
    same = (sg[:, None] == sg[None, :]).astype(float)
    # <PQ|RS> = (pr|qs) d(sP,sR) d(sQ,sS)
    g = eri_chem[np.ix_(sp, sp, sp, sp)]  # (pq|rs) indexed P,Q,R,S
    phys = g.transpose(0, 2, 1, 3) * same[:, None, :, None] * same[None, :, None, :]
    asym = phys - phys.transpose(0, 1, 3, 2)
    O, V = 2 * o, 2 * v
    occ = np.arange(O)
    vir = np.arange(O, N)
    eso = e[sp]
    eo, ev = eso[:O], eso[O:]
    # amplitudes
    so = sp[:O]
    sv = sp[O:] - o
    go, gv = sg[:O], sg[O:]
    d_ov = (go[:, None] == gv[None, :]).astype(float)
    T1 = t1[np.ix_(so, sv)] * d_ov
    tab = t2[np.ix_(so, so, sv, sv)]
    tba = t2.transpose(0, 1, 3, 2)[np.ix_(so, so, sv, sv)]
    T2 = tab * d_ov[:, None, :, None] * d_ov[None, :, None, :] - tba * d_ov[:, None, None, :] * d_ov[None, :, :, None]
    g_vovv = asym[np.ix_(vir, occ, vir, vir)]  # <ei||bc>
    g_ovoo = asym[np.ix_(occ, vir, occ, occ)]  # <ma||jk>
    g_oovv = asym[np.ix_(occ, occ, vir, vir)]  # <jk||bc>

    def P_i_jk(x):  # x[i,j,k,...]
        return x - x.transpose(1, 0, 2, 3, 4, 5) - x.transpose(2, 1, 0, 3, 4, 5)

    def P_a_bc(x):
        return x - x.transpose(0, 1, 2, 4, 3, 5) - x.transpose(0, 1, 2, 5, 4, 3)

    disc = np.einsum("ia,jkbc->ijkabc", T1, g_oovv)
    td = P_i_jk(P_a_bc(disc))
    conn = np.einsum("jkae,eibc->ijkabc", T2, g_vovv) - np.einsum("imbc,majk->ijkabc", T2, g_ovoo)
    tc = P_i_jk(P_a_bc(conn))
    D = (eo[:, None, None, None, None, None] + eo[None, :, None, None, None, None]
         + eo[None, None, :, None, None, None] - ev[None, None, None, :, None, None]
         - ev[None, None, None, None, :, None] - ev[None, None, None, None, None, :])
    return float(np.sum(tc * (tc + td) / D) / 36.0)
