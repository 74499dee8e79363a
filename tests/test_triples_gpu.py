"""Parity of the gfx950 (T) path (C ABI) with the CPU oracle.

Tolerances (stated per test): FP64 contractions agree with the FP64 oracle to
1e-12 relative (different but deterministic summation orders); FP32 and
split-BF16x3 contractions are compared with the FP64 GPU result.
"""

import sys

import numpy as np
import pytest

import oracle
from scess_amd import gpu
from scess_amd.synthetic import TriplesInputs, make_triples_inputs

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = gpu.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# v = 33 and 129: odd, so the last pair of fixed indices of the FP64 W-build has no second index, and
# at 129 the last 64-row tile holds one row
@pytest.mark.parametrize("o,v,seed", [(2, 3, 1), (3, 5, 2), (4, 7, 3), (5, 19, 4), (3, 17, 5), (6, 33, 6),
                                      (8, 40, 7), (7, 70, 8), (5, 129, 9)])
def test_fp64_matches_oracle(ctx, o, v, seed):
    _say(f"case fp64 o={o} v={v} seed={seed}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64",))
    E, F, st = ctx.triples_forward("fp64", with_F=True)
    Eo = oracle.t_abc(inp)
    _, Fo = oracle.t_ijk(inp)
    assert _rel(E, Eo) < 1e-12, (E, Eo)
    assert np.allclose(F, Fo, rtol=1e-11, atol=1e-13 * np.abs(Fo).max())


@pytest.mark.parametrize("digits", [5, 6, 7])
@pytest.mark.parametrize("o,v,seed", [(2, 3, 21), (5, 19, 22), (6, 33, 23), (8, 40, 24), (7, 70, 25),
                                      (3, 130, 26), (33, 9, 27), (40, 12, 28)])
def test_ozaki_matches_oracle(ctx, o, v, seed, digits):
    """Int8 digit products are exact; W differs from FP64 by the digits past S and the dropped digit
    pairs (about (S+1) 2^-8S of max|A| max|B| each): E(T) within 1e-12 relative of the oracle for S = 6
    and 7 (FP64's bar) and 1e-10 for S = 5, F_ijk within 1e-9 (S = 5) / 1e-12 of the largest F. The
    sizes cover v past one and two 64-deep chunks and o past one and two 32-wide half-chunks."""
    _say(f"case ozaki S={digits} o={o} v={v} seed={seed}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_int8_digits(digits)
    ctx.set_triples_inputs(inp, ("fp64", "ozaki"))
    E, F, _ = ctx.triples_forward("ozaki", with_F=True)
    E64, F64, _ = ctx.triples_forward("fp64", with_F=True)
    Eo = oracle.t_abc(inp)
    _say(f"  E={E!r} Eo={Eo!r} rel={_rel(E, Eo):.3e} rel(fp64)={_rel(E64, Eo):.3e}")
    assert _rel(E, Eo) < (1e-10 if digits == 5 else 1e-12), (E, Eo)
    assert np.abs(F - F64).max() <= (1e-9 if digits == 5 else 1e-12) * np.abs(F64).max()
    ctx.set_int8_digits(6)


def test_single_occupied_or_virtual_is_zero(ctx):
    for (o, v) in [(1, 9), (4, 1)]:
        _say(f"case zero o={o} v={v}")
        inp = make_triples_inputs(o, v, 9)
        ctx.set_triples_inputs(inp, ("fp64",))
        E, _ = ctx.triples_forward("fp64")
        assert abs(E) < 1e-14


def test_deterministic(ctx):
    _say("case deterministic o=6 v=40")
    inp = make_triples_inputs(6, 40, 10)
    ctx.set_triples_inputs(inp, ("fp64", "fp32", "bf16x3", "ozaki"))
    for p in ("fp64", "fp32", "bf16x3", "ozaki"):
        E1, _ = ctx.triples_forward(p)
        ctx.set_batch(3)
        E2, _ = ctx.triples_forward(p)
        ctx.set_batch(0)
        assert E1 == E2, p


@pytest.mark.parametrize("o,v,seed", [(5, 70, 15), (4, 67, 16), (6, 130, 17)])
def test_fp64_wbuild_variants_bitwise_equal(ctx, monkeypatch, o, v, seed):
    """The FP64 W-build with four fixed indices per work-group (default), two (SCESS_WBUILD_FX=2) and
    one (SCESS_WBUILD_F2=0) feeds every accumulator the same MFMAs in the same order, so E(T) and every
    F_ijk are bitwise equal; v = 67 leaves the last group of four with one live index, v = 130 with two.
    The variants are read at each launch."""
    _say(f"case fp64 W-build variants o={o} v={v}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64",))
    monkeypatch.delenv("SCESS_WBUILD_FX", raising=False)
    monkeypatch.delenv("SCESS_WBUILD_F2", raising=False)
    E4, F4, _ = ctx.triples_forward("fp64", with_F=True)
    monkeypatch.setenv("SCESS_WBUILD_FX", "2")
    E2, F2, _ = ctx.triples_forward("fp64", with_F=True)
    monkeypatch.delenv("SCESS_WBUILD_FX")
    monkeypatch.setenv("SCESS_WBUILD_F2", "0")
    E1, F1, _ = ctx.triples_forward("fp64", with_F=True)
    monkeypatch.delenv("SCESS_WBUILD_F2")
    assert E4 == E2 == E1, (E4, E2, E1)
    assert np.array_equal(F4, F2) and np.array_equal(F4, F1)
    assert _rel(E4, oracle.t_abc(inp)) < 1e-12


@pytest.mark.parametrize("o,v,seed", [(4, 20, 11), (8, 64, 12), (10, 100, 13)])
def test_mixed_precision_close_to_fp64(ctx, o, v, seed):
    _say(f"case mixed o={o} v={v}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64", "fp32", "bf16x3"))
    E64, _ = ctx.triples_forward("fp64")
    E32, _ = ctx.triples_forward("fp32")
    Eb3, _ = ctx.triples_forward("bf16x3")
    _say(f"  E64={E64!r} rel32={_rel(E32, E64):.3e} relbf16x3={_rel(Eb3, E64):.3e}")
    assert _rel(E32, E64) < 1e-5
    assert _rel(Eb3, E64) < 1e-4


def test_block_partials_sum_to_F(ctx):
    _say("case partials o=5 v=30")
    inp = make_triples_inputs(5, 30, 14)
    ctx.set_triples_inputs(inp, ("fp64",))
    E, F, P, _ = ctx.triples_forward("fp64", with_F=True, with_partials=True)
    assert np.allclose(P.sum(axis=1), F, rtol=1e-13, atol=1e-13 * np.abs(F).max())


def test_errors_and_cancellation():
    _say("case errors")
    c = gpu.Context(0)
    with pytest.raises(gpu.ScessError) as ei:
        c.triples_forward("fp64")
    assert ei.value.code == gpu.SCESS_ERR_STATE
    inp = make_triples_inputs(4, 12, 15)
    c.set_triples_inputs(inp, ("fp64",))
    with pytest.raises(gpu.ScessError) as ei:
        c.triples_forward("fp32")
    assert ei.value.code == gpu.SCESS_ERR_STATE
    c.set_should_stop(lambda: True)
    with pytest.raises(gpu.ScessError) as ei:
        c.triples_forward("fp64")
    assert ei.value.code == gpu.SCESS_CANCELLED
    c.set_should_stop(None)
    E, _ = c.triples_forward("fp64")
    assert _rel(E, oracle.t_abc(inp)) < 1e-12
    c.close()


def test_rccsd_triplescorrection_full_eri_api():
    """The n^4 <pq|rs> entry point (CC.h:154-158) slices the same blocks."""
    _say("case full-eri api o=4 v=9")
    from scess_amd.synthetic import full_eri_phys
    from scess_amd.triples import RCCSD_TriplesCorrection
    inp = make_triples_inputs(4, 9, 16, keep_factor=True)
    eri = full_eri_phys(inp.L)
    E = RCCSD_TriplesCorrection(eri, inp.e, 4, inp.t1, inp.t2)
    assert _rel(E, oracle.t_abc(inp)) < 1e-12


def test_shards_partition_the_work(ctx):
    """scess_t_set_shard: every occupied triple is evaluated by exactly one shard (F bitwise equal
    to the unsharded run), E(T) and the adjoint add up over shards."""
    _say("case shards o=6 v=40")
    inp = make_triples_inputs(6, 40, 21)
    ctx.set_triples_inputs(inp, ("fp64",))
    E, F, _ = ctx.triples_forward("fp64", with_F=True)
    g = ctx.triples_backward("fp64")
    n = 3
    Fs = np.zeros_like(F)
    Es = 0.0
    gs = {k: np.zeros_like(x) for k, x in g.items()}
    try:
        for s in range(n):
            ctx.set_shard(s, n)
            Ei, Fi, _ = ctx.triples_forward("fp64", with_F=True)
            assert np.all((Fi == 0.0) | (Fs == 0.0))  # disjoint
            Fs += Fi
            Es += Ei
            gi = ctx.triples_backward("fp64")
            for k in gs:
                gs[k] += gi[k]
    finally:
        ctx.set_shard(0, 1)
    assert np.array_equal(Fs, F)
    assert _rel(Es, E) < 1e-13
    for k in g:
        assert np.abs(gs[k] - g[k]).max() <= 1e-12 * np.abs(g[k]).max(), k
    with pytest.raises(gpu.ScessError):
        ctx.set_shard(3, 3)
