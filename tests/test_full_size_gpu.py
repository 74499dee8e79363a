"""Parity at the full sizes of the synthetic (T) configurations, through the C ABI.

o=20, v=200 (FP64):
* the whole E(T) against the oracle's abc-driven restatement of RCCSD_TriplesCorrection over every
  unique virtual triple on the host cores, 1e-12 relative;
* F_ijk on a seeded sample of occupied triples against the oracle's ijk-driven restatement, 1e-12
  relative, and E(T) = sum_ijk (2/deg) F_ijk with the GPU's own F, 1e-13 relative.

o=30, v=300 (every precision mode on the same resident inputs):
* FP64 F_ijk on four sampled occupied triples (distinct indices, j == k, i == j, and one more) against
  the oracle, 1e-12 relative; E = sum w F, 1e-13 relative;
* the mixed modes against FP64 under the stated error model of their products (test docstring), and
  the int8-digit mode (five 8-bit digits) within 1 microhartree of FP64: the benchmark's bar.
"""

import sys

import numpy as np
import pytest

import oracle
from scess_amd import gpu
from scess_amd.synthetic import make_triples_inputs
from scess_amd.triples import ijk_weights, unique_ijk

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sample(o, seed, kinds):
    """One seeded occupied-triple index per kind: 'distinct' (i>j>k), 'jk' (i>j==k), 'ij' (i==j>k)."""
    t = unique_ijk(o)
    rng = np.random.default_rng(seed)
    masks = {"distinct": (t[:, 0] > t[:, 1]) & (t[:, 1] > t[:, 2]),
             "jk": (t[:, 0] > t[:, 1]) & (t[:, 1] == t[:, 2]),
             "ij": (t[:, 0] == t[:, 1]) & (t[:, 1] > t[:, 2])}
    return np.array([rng.choice(np.flatnonzero(masks[k])) for k in kinds], dtype=np.int64)


def test_o20_v200_fp64_against_oracle_sample():
    o, v = 20, 200
    _say(f"case full-size o={o} v={v}: generating")
    inp = make_triples_inputs(o, v, 1)
    ctx = gpu.Context(0)
    try:
        ctx.set_triples_inputs(inp, ("fp64",))
        E, F, st = ctx.triples_forward("fp64", with_F=True)
        _say(f"  GPU E(T)={E!r} in {st['seconds']:.2f} s")
        w = ijk_weights(o)
        assert abs(E - float(np.sum(w * F))) <= 1e-13 * abs(E)
        items = _sample(o, 20200, ("distinct", "jk"))
        t = unique_ijk(o)
        _say(f"  oracle on triples {[tuple(t[i]) for i in items]}")
        _, Fo = oracle.t_ijk(inp, items=items)
        for n, it in enumerate(items):
            assert abs(F[it] - Fo[n]) <= 1e-12 * abs(Fo[n]), (tuple(t[it]), F[it], Fo[n])
    finally:
        ctx.close()


@pytest.mark.timeout(900)
def test_o20_v200_whole_energy_against_abc_oracle():
    """E(T) of the whole o=20, v=200 job: the GPU (FP64) against the oracle's abc-driven restatement of
    RCCSD_TriplesCorrection run over every unique virtual triple on the host cores, 1e-12 relative."""
    import time
    o, v = 20, 200
    inp = make_triples_inputs(o, v, 1)
    ctx = gpu.Context(0)
    try:
        ctx.set_triples_inputs(inp, ("fp64",))
        E, st = ctx.triples_forward("fp64")
    finally:
        ctx.close()
    t = time.perf_counter()
    Eo = oracle.t_abc(inp)
    _say(f"  GPU E(T) = {E!r} ({st['seconds']:.2f} s), oracle E(T) = {Eo!r} ({time.perf_counter() - t:.0f} s on "
         f"{oracle.num_threads()} threads), relative difference {abs(E - Eo) / abs(Eo):.3e}")
    assert abs(E - Eo) <= 1e-12 * abs(Eo)


@pytest.mark.timeout(900)
def test_o30_v300_all_modes():
    """o=30, v=300, seed 1 (the benchmark workload).

    FP64: sampled F_ijk against the oracle, 1e-12 relative, and E = sum w F.
    Error model of the mixed modes (against the FP64 GPU result of the same inputs): FP32 products
    carry 2^-24 relative rounding per product and per FP32 partial sum of a 16-deep chunk, so W carries
    a few 1e-7 relative and F_ijk (quadratic in W) stays within 1e-6 relative per triple, E within
    1e-8 relative. BF16X3 drops the lo*lo product and the residual of the two-piece split (2^-17
    relative per operand), so F_ijk stays within 1e-4 per triple and E within 1e-6 relative. The int8
    mode with five 8-bit digits per element (38-39 bits of each tensor's largest entry, products exact)
    meets the benchmark's bar: |dE(T)| <= 1 microhartree, F_ijk within 1e-10 relative."""
    o, v = 30, 300
    _say(f"case o={o} v={v}: generating")
    inp = make_triples_inputs(o, v, 1)
    w = ijk_weights(o)
    ctx = gpu.Context(0)
    try:
        modes = ("fp64", "fp32", "bf16x3", "ozaki")
        ctx.set_int8_digits(5)
        ctx.set_triples_inputs(inp, modes)
        E, F, st = ctx.triples_forward("fp64", with_F=True)
        _say(f"  fp64 E(T)={E!r} in {st['seconds']:.2f} s")
        assert abs(E - float(np.sum(w * F))) <= 1e-13 * abs(E)
        res = {}
        # F tolerances relative to each triple's |F| (FP32, BF16X3), or to the largest |F| for the int8
        # mode, whose digits carry an absolute error per tensor (a fraction of its largest entry)
        for mode, tol_f, tol_e in (("fp32", 1e-6, 1e-8), ("bf16x3", 1e-4, 1e-6), ("ozaki", 1e-10, 1e-6 / abs(E))):
            Em, Fm, stm = ctx.triples_forward(mode, with_F=True)
            norm = np.abs(F).max() if mode == "ozaki" else np.abs(F[w > 0])
            rel = np.abs(Fm - F)[w > 0] / norm
            res[mode] = (Em, rel.max())
            _say(f"  {mode}: E(T)={Em!r} |dE|={abs(Em - E):.3e} max per-triple rel {rel.max():.2e} "
                 f"({stm['seconds']:.2f} s)")
            assert rel.max() <= tol_f, mode
            assert abs(Em - E) <= tol_e * abs(E), mode
    finally:
        ctx.close()
    items = _sample(o, 30300, ("distinct", "jk", "ij", "distinct"))
    t = unique_ijk(o)
    _say(f"  oracle on triples {[tuple(t[i]) for i in items]}")
    _, Fo = oracle.t_ijk(inp, items=items)
    for n, it in enumerate(items):
        assert abs(F[it] - Fo[n]) <= 1e-12 * abs(Fo[n]), (tuple(t[it]), F[it], Fo[n])
