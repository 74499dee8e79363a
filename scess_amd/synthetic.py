"""Seeded synthetic inputs for the closed-shell (T) and CCSD paths.

The generator, with every convention fixed here:

* ``rng = numpy.random.Generator(numpy.random.PCG64(seed))``; draws happen in the
  order e_occ, e_vir, L, t1.
* ``e_occ ~ U[-1.5, -0.3]`` and ``e_vir ~ U[0.1, 3.0]`` (hartree), each sorted
  ascending; occupied orbitals first, as the reference orders orbitals.
* A symmetric Cholesky-like factor ``L[P, p, q]`` over all ``n = o + v`` orbitals,
  ``P < naux = 3 v``.  Every entry is drawn ``N(0, sigma)`` with standard
  deviation ``sigma = 1/sqrt(naux)``; the symmetry ``L_pq = L_qp`` is imposed by
  MIRRORING the upper triangle (p <= q) onto the lower one, so each entry keeps
  the stated distribution.  (Averaging L and its transpose instead would shrink
  the off-diagonal entries by sqrt(2) and E(T) by about 16x.)
* ``(pq|rs) = sum_P L[P,p,q] L[P,r,s]``; the blocks returned are
  ``ovvv(i,a,b,f) = (ia|bf)``, ``ovoo(i,a,m,j) = (ia|mj)`` and
  ``ovov(i,a,j,b) = (ia|jb)`` (chemists' notation, named as in
  src/Quantum/CC/RCCSD.cpp:31-44).
* ``t2(i,j,a,b) = (ia|jb) / (e_i + e_j - e_a - e_b)``, so t_ij^ab = t_ji^ba.
* ``t1 ~ N(0, 1e-3)`` (standard deviation 1e-3).

All 4-index arrays are returned with their logical index order as named and
Fortran (ColMajor) memory order, i.e. Eigen's default layout, the layout of the
reference's tensors and of the C ABI.
"""

from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class TriplesInputs:
    """Inputs of the (T) functional in the C-ABI layouts (ColMajor)."""

    o: int
    v: int
    e: np.ndarray  # (n,)
    t1: np.ndarray  # (o, v)
    t2: np.ndarray  # (o, o, v, v)
    ovvv: np.ndarray  # (o, v, v, v) = (ia|bf)
    ovoo: np.ndarray  # (o, v, o, o) = (ia|mj)
    ovov: np.ndarray  # (o, v, o, v) = (ia|jb)
    L: np.ndarray | None = None  # (naux, n, n), kept only when asked for

    @property
    def n(self) -> int:
        return self.o + self.v


def _draw(o: int, v: int, seed: int):
    rng = np.random.Generator(np.random.PCG64(seed))
    e_occ = np.sort(rng.uniform(-1.5, -0.3, o))
    e_vir = np.sort(rng.uniform(0.1, 3.0, v))
    n = o + v
    naux = 3 * v
    L = rng.standard_normal((naux, n, n)) / np.sqrt(naux)
    iu = np.triu_indices(n, 1)
    # mirror p <= q onto q > p
    L[:, iu[1], iu[0]] = L[:, iu[0], iu[1]]
    t1 = np.asfortranarray(rng.standard_normal((o, v)) * 1.0e-3)
    return np.concatenate([e_occ, e_vir]), L, t1


def blocks_from_factor(L: np.ndarray, o: int):
    """ovvv, ovoo, ovov (ColMajor, logical order as named) from L[P, p, q]."""
    naux, n, _ = L.shape
    v = n - o
    Lov = L[:, :o, o:]
    Lvv = L[:, o:, o:]
    Loo = L[:, :o, :o]
    # A2[P, (a,i)] = Lov[P,i,a]
    A2 = np.ascontiguousarray(Lov.transpose(0, 2, 1)).reshape(naux, v * o)
    # ovvv: C-order (f,b,a,i) = sum_P Lvv[P,b,f] Lov[P,i,a]
    A1 = np.ascontiguousarray(Lvv.transpose(2, 1, 0)).reshape(v * v, naux)
    ovvv = (A1 @ A2).reshape(v, v, v, o).transpose(3, 2, 1, 0)
    # ovoo: C-order (j,m,a,i) = sum_P Loo[P,m,j] Lov[P,i,a]
    B1 = np.ascontiguousarray(Loo.transpose(2, 1, 0)).reshape(o * o, naux)
    ovoo = (B1 @ A2).reshape(o, o, v, o).transpose(3, 2, 1, 0)
    # ovov: C-order (b,j,a,i) = sum_P Lov[P,j,b] Lov[P,i,a]
    C1 = np.ascontiguousarray(Lov.transpose(2, 1, 0)).reshape(v * o, naux)
    ovov = (C1 @ A2).reshape(v, o, v, o).transpose(3, 2, 1, 0)
    return ovvv, ovoo, ovov


def t2_from_ovov(ovov: np.ndarray, e: np.ndarray, o: int) -> np.ndarray:
    """t2(i,j,a,b) = (ia|jb) / (e_i + e_j - e_a - e_b), ColMajor."""
    eo, ev = e[:o], e[o:]
    D = (eo[:, None, None, None] + eo[None, :, None, None]
         - ev[None, None, :, None] - ev[None, None, None, :])
    return np.asfortranarray(ovov.transpose(0, 2, 1, 3) / D)


def make_triples_inputs(o: int, v: int, seed: int = 1, keep_factor: bool = False) -> TriplesInputs:
    """The synthetic (T) inputs at occupied/virtual sizes (o, v)."""
    e, L, t1 = _draw(o, v, seed)
    ovvv, ovoo, ovov = blocks_from_factor(L, o)
    t2 = t2_from_ovov(ovov, e, o)
    return TriplesInputs(o=o, v=v, e=e, t1=t1, t2=t2, ovvv=ovvv, ovoo=ovoo, ovov=ovov,
                         L=L if keep_factor else None)


def full_eri_chem(L: np.ndarray) -> np.ndarray:
    """(pq|rs) over all n orbitals, C-order (p,q,r,s); tiny sizes only."""
    naux, n, _ = L.shape
    M = L.reshape(naux, n * n)
    return (M.T @ M).reshape(n, n, n, n)


def full_eri_phys(L: np.ndarray) -> np.ndarray:
    """<pq|rs> = (pr|qs) over all n orbitals, ColMajor (p,q,r,s), the layout of the
    reference's TEITensor (src/Quantum/Integrals/IntegralTransforms.cpp:29-59)."""
    chem = full_eri_chem(L)
    return np.asfortranarray(chem.transpose(0, 2, 1, 3))


@dataclasses.dataclass
class CcsdInputs:
    """Integral blocks of SolveRCCSD_DIIS (src/Quantum/CC/RCCSD.cpp:31-44) in the C-ABI layouts
    (ColMajor, chemists' notation) plus the Cholesky factor of (ac|bd); no v^4 tensor."""

    o: int
    v: int
    e: np.ndarray
    blocks: dict  # ovov (k,c,l,d), oovv (k,i,a,c), ovvo (k,c,a,i), ovvv (k,d,a,c), ovoo (l,c,k,i), oooo (k,i,l,j)
    chol_vv: np.ndarray  # (a, c, P)
    naux: int


def _pair(L: np.ndarray, s1: slice, s2: slice, s3: slice, s4: slice) -> np.ndarray:
    """(pq|rs) for p in s1, q in s2, r in s3, s in s4 from L[P, p, q], ColMajor (p, q, r, s)."""
    A = L[:, s1, s2]
    B = L[:, s3, s4]
    naux = L.shape[0]
    # C-order (s, r, q, p) = sum_P B[P, r, s] A[P, p, q]; its transpose is ColMajor (p, q, r, s)
    Bm = np.ascontiguousarray(B.transpose(2, 1, 0)).reshape(-1, naux)
    Am = np.ascontiguousarray(A.transpose(0, 2, 1)).reshape(naux, -1)
    out = Bm @ Am
    return out.reshape(B.shape[2], B.shape[1], A.shape[2], A.shape[1]).transpose(3, 2, 1, 0)


def make_ccsd_inputs(o: int, v: int, seed: int = 1) -> CcsdInputs:
    """CCSD blocks of the synthetic system (same draws as make_triples_inputs) built from the factor."""
    e, L, _ = _draw(o, v, seed)
    O, V = slice(0, o), slice(o, o + v)
    blocks = {"ovov": _pair(L, O, V, O, V), "oovv": _pair(L, O, O, V, V), "ovvo": _pair(L, O, V, V, O),
              "ovvv": _pair(L, O, V, V, V), "ovoo": _pair(L, O, V, O, O), "oooo": _pair(L, O, O, O, O)}
    chol = np.asfortranarray(L[:, o:, o:].transpose(1, 2, 0))
    return CcsdInputs(o=o, v=v, e=e, blocks=blocks, chol_vv=chol, naux=L.shape[0])
