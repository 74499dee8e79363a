"""Python mirror of the reference's (T) method API over the C ABI.

``RCCSD_TriplesCorrection`` keeps the signature and meaning of
``SCESS::CC::RCCSD_TriplesCorrection(eri, orbitalEnergies, nOcc, t1, t2, thread)``
(src/Quantum/CC/CC.h:154-158): ``eri`` is <pq|rs> over all spatial orbitals,
occupied first, in Eigen's ColMajor layout (a Fortran-ordered numpy array);
``t1`` is (i,a) and ``t2`` is (i,j,a,b). Like the reference it returns E(T)
(0.0 when cancelled through ``should_stop``, the counterpart of
DH::Thread::ShouldStop(), RCCSD_T.cpp:123,161-162).

``triples_correction`` takes the three integral blocks instead of the n^4
tensor, which is what the large synthetic configurations use.
"""

from __future__ import annotations

import numpy as np

from . import gpu
from .synthetic import TriplesInputs


def unique_ijk(o: int) -> np.ndarray:
    """Occupied triples i >= j >= k in the order of the per-ijk outputs of scess_t_forward_ex
    (i outer, then j <= i, then k <= j), shape (o(o+1)(o+2)/6, 3)."""
    out = [(i, j, k) for i in range(o) for j in range(i + 1) for k in range(j + 1)]
    return np.array(out, dtype=np.int32).reshape(-1, 3)


def ijk_weights(o: int) -> np.ndarray:
    """E(T) = sum_ijk w_ijk F_ijk over unique_ijk(o): 2/deg, i.e. 2 for distinct indices, 1 when two
    are equal, 0 when all three are (r3 annihilates those)."""
    t = unique_ijk(o)
    return np.where(t[:, 0] == t[:, 2], 0.0, np.where((t[:, 0] == t[:, 1]) | (t[:, 1] == t[:, 2]), 1.0, 2.0))


def blocks_from_eri(eri: np.ndarray, nocc: int):
    """(ia|bf), (ia|mj), (ia|jb) sliced from <pq|rs> = (pr|qs), occupied first.

    The same elements as the repack of RCCSD_T.cpp:37-51:
    P(i,f,a,b) = eri(i, o+b, o+a, o+f), R(i,j,m,a) = eri(i, m, o+a, j),
    S(i,j,a,b) = eri(i, j, o+a, o+b)."""
    o = nocc
    V = slice(o, None)
    O = slice(0, o)
    # ovvv(i,a,b,f) = <ib|af> = eri[i, o+b, o+a, o+f]
    ovvv = np.asfortranarray(eri[O, V, V, V].transpose(0, 2, 1, 3))
    # ovoo(i,a,m,j) = <im|aj> = eri[i, m, o+a, j]
    ovoo = np.asfortranarray(eri[O, O, V, O].transpose(0, 2, 1, 3))
    # ovov(i,a,j,b) = <ij|ab> = eri[i, j, o+a, o+b]
    ovov = np.asfortranarray(eri[O, O, V, V].transpose(0, 2, 1, 3))
    return ovvv, ovoo, ovov


def triples_correction(inp: TriplesInputs, precision: str = "fp64", ctx: gpu.Context | None = None,
                       should_stop=None, stats: dict | None = None) -> float:
    """E(T) on the GPU from block inputs."""
    own = ctx is None
    ctx = ctx or gpu.Context(0)
    try:
        if should_stop is not None:
            ctx.set_should_stop(should_stop)
        ctx.set_triples_inputs(inp, precisions=(precision,))
        try:
            E, st = ctx.triples_forward(precision)
        except gpu.ScessError as err:
            if err.code == gpu.SCESS_CANCELLED:
                return 0.0
            raise
        if stats is not None:
            stats.update(st)
        return E
    finally:
        if own:
            ctx.close()


def RCCSD_TriplesCorrection(eri: np.ndarray, orbitalEnergies: np.ndarray, nOcc: int, t1: np.ndarray,
                            t2: np.ndarray, thread=None, precision: str = "fp64") -> float:
    """Closed-shell (T) from converged CCSD amplitudes (src/Quantum/CC/CC.h:154-158).

    thread: optional object with ShouldStop() (the DH::Thread contract)."""
    n = eri.shape[0]
    o = int(nOcc)
    ovvv, ovoo, ovov = blocks_from_eri(eri, o)
    inp = TriplesInputs(o=o, v=n - o, e=np.asarray(orbitalEnergies, dtype=np.float64), t1=t1, t2=t2,
                        ovvv=ovvv, ovoo=ovoo, ovov=ovov)
    stop = (lambda: bool(thread.ShouldStop())) if thread is not None else None
    return triples_correction(inp, precision=precision, should_stop=stop)
