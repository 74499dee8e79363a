"""F_ijk of single occupied triples with numpy/BLAS (TEST INFRASTRUCTURE ONLY).

The same ijk-driven form as oracle_t_ijk in triples_oracle.c, written with dense BLAS contractions so
that one triple at v = 500 takes seconds: for the occupied triple (i, j, k),

  W^abc = sum over the six simultaneous permutations of the pairs (ia), (jb), (kc) of
          w(x,y,z; p,q,r) = sum_f (px|yf) t_rq^zf - sum_m (px|mq) t_mr^yz
  V^abc = the same sum of 1/2 (px|qy) t_r^z
  F_ijk = sum_abc W r3(W + V) / D,  r3(Y) = 4Y^abc + Y^bca + Y^cab - 2Y^cba - 2Y^acb - 2Y^bac,
  D = e_i + e_j + e_k - e_a - e_b - e_c

(GetW / GetV and Symmetrise of src/Quantum/CC/RCCSD_T.cpp:61-106, the r3/D epilogue of :141-158 with r3
acting on the virtual indices). Pinned against oracle.t_ijk at small sizes in tests/test_oracle.py.
"""

from __future__ import annotations

import numpy as np

PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def _abc_from_xyz(w: np.ndarray, perm) -> np.ndarray:
    """A[a,b,c] = w[abc[perm[0]], abc[perm[1]], abc[perm[2]]]."""
    return np.transpose(w, np.argsort(perm))


def t_ijk_blas(inp, i: int, j: int, k: int) -> float:
    o, v = int(inp.o), int(inp.v)
    occ = (i, j, k)
    W = np.zeros((v, v, v))
    V = np.zeros((v, v, v))
    for perm in PERMS:
        p, q, r = occ[perm[0]], occ[perm[1]], occ[perm[2]]
        X = np.ascontiguousarray(inp.ovvv[p])                    # (x, y, f) = (px|yf)
        T = np.ascontiguousarray(inp.t2[r, q])                   # (z, f) = t_rq^zf
        w = (X.reshape(v * v, v) @ T.T).reshape(v, v, v)         # sum_f -> (x, y, z)
        R = np.ascontiguousarray(inp.ovoo[p, :, :, q])           # (x, m) = (px|mq)
        TM = np.ascontiguousarray(inp.t2[:, r]).reshape(o, v * v)  # (m, (y, z)) = t_mr^yz
        w -= (R @ TM).reshape(v, v, v)
        W += _abc_from_xyz(w, perm)
        S = inp.ovov[p, :, q, :]                                 # (x, y) = (px|qy)
        vv = 0.5 * S[:, :, None] * inp.t1[r][None, None, :]
        V += _abc_from_xyz(vv, perm)
    Y = W + V
    r3 = (4.0 * Y + np.einsum("bca->abc", Y) + np.einsum("cab->abc", Y)
          - 2.0 * (np.einsum("cba->abc", Y) + np.einsum("acb->abc", Y) + np.einsum("bac->abc", Y)))
    e = np.asarray(inp.e)
    ev = e[o:]
    D = (e[i] + e[j] + e[k]) - ev[:, None, None] - ev[None, :, None] - ev[None, None, :]
    return float(np.sum(W * r3 / D))
