"""Adjoint of the gfx950 (T) path (scess_t_backward) against the FP64 oracle
backward and against central finite differences.

Central differences use h = 1e-5 and are taken as local differences: the
forward returns per-(ijk, orbit-run) partials of E, and the partials an
entry does not touch are bitwise identical in E(x+h) and E(x-h) (the kernels
are deterministic), so they cancel exactly and the rounding error of |E| does
not swamp the derivative. E is quadratic in every single input entry, so the
central difference is exact up to rounding. Tolerance: 1e-6 relative
(|fd - g| <= 1e-6 * max(|g|, 1e-3 * max|grad|)).
"""

import sys

import numpy as np
import pytest

import oracle
from scess_amd import gpu
from scess_amd.synthetic import TriplesInputs, make_triples_inputs

pytestmark = pytest.mark.gpu
NAMES = ("ovvv", "ovoo", "ovov", "t1", "t2")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _replace(inp, **kw):
    d = dict(o=inp.o, v=inp.v, e=inp.e, t1=inp.t1, t2=inp.t2, ovvv=inp.ovvv, ovoo=inp.ovoo, ovov=inp.ovov)
    d.update(kw)
    return TriplesInputs(**d)


@pytest.fixture(scope="module")
def ctx():
    c = gpu.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("o,v,seed", [(2, 3, 1), (3, 5, 2), (4, 9, 3), (5, 17, 4), (3, 70, 5), (6, 3, 6),
                                      (34, 7, 7)])
def test_backward_matches_oracle(ctx, o, v, seed):
    """(3, 70): v past one 64-wide tile of the back-contraction GEMMs; v even, so the batched ones fold
    the batch into their rows with tiles that span two batch entries (odd v keep one tile set per
    entry). (6, 3) and (34, 7): o > v, where
    the (ooov) gradient (o^3 v entries) is the largest array the backward stages for download; (34, 7)
    also has o past two 16-wide occupied chunks and N = o past 32 in the dRn and dTM back-contractions."""
    _say(f"case backward o={o} v={v}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64",))
    g = ctx.triples_backward("fp64", dE=1.0)
    go = oracle.t_grad(inp)
    for k in NAMES:
        scale = max(np.abs(go[k]).max(), 1e-300)
        err = np.abs(g[k] - go[k]).max() / scale
        assert err < 1e-11, (k, err)


@pytest.mark.parametrize("o,v,seed", [(4, 9, 31), (34, 7, 32)])
def test_backward_with_ozaki_w_matches_oracle(ctx, o, v, seed):
    """W recomputed from six 8-bit digit products per FP64 product (~2^-44 product error) in the
    backward: the gradients keep 1e-11 relative to the oracle; o = 34 spans two 32-wide occupied
    half-chunks and has o > v (the (ooov) gradient, o^3 v entries, is then the largest one)."""
    _say(f"case backward ozaki o={o} v={v}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_int8_digits(6)
    ctx.set_triples_inputs(inp, ("fp64", "ozaki"))
    g = ctx.triples_backward("ozaki", dE=1.0)
    go = oracle.t_grad(inp)
    for k in NAMES:
        scale = max(np.abs(go[k]).max(), 1e-300)
        err = np.abs(g[k] - go[k]).max() / scale
        assert err < 1e-11, (k, err)


def test_backward_scales_with_dE_and_is_deterministic(ctx):
    _say("case backward dE scaling")
    inp = make_triples_inputs(4, 20, 5)
    ctx.set_triples_inputs(inp, ("fp64",))
    g1 = ctx.triples_backward("fp64", dE=1.0)
    g2 = ctx.triples_backward("fp64", dE=1.0)
    g3 = ctx.triples_backward("fp64", dE=-2.5)
    for k in NAMES:
        assert np.array_equal(g1[k], g2[k]), k
        # dE enters the per-ijk weights before the back-contractions, so the rounding differs
        assert np.abs(g3[k] + 2.5 * g1[k]).max() <= 1e-13 * np.abs(g1[k]).max(), k


@pytest.mark.parametrize("o,v,seed", [(3, 70, 8), (4, 40, 9)])
def test_backward_folded_batches_match_unfolded(ctx, monkeypatch, o, v, seed):
    """The batched back-contractions (dX, dTM) with the batch folded into the rows (default for even v)
    against one tile set per batch entry (SCESS_KN_FOLD=0, read at each launch): the same MFMAs reach
    every output element in the same order, only the element's row within its 16 x 16 block differs."""
    _say(f"case backward fold o={o} v={v}")
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64",))
    monkeypatch.delenv("SCESS_KN_FOLD", raising=False)
    g1 = ctx.triples_backward("fp64")
    monkeypatch.setenv("SCESS_KN_FOLD", "0")
    g0 = ctx.triples_backward("fp64")
    monkeypatch.delenv("SCESS_KN_FOLD")
    for k in NAMES:
        scale = max(np.abs(g0[k]).max(), 1e-300)
        assert np.abs(g1[k] - g0[k]).max() <= 1e-14 * scale, k


def local_fd(ctx, inp, name, flat_index, h, precision="fp64"):
    """Central difference of E(T) in one entry from local partial differences."""
    arr = getattr(inp, name)
    ix = np.unravel_index(flat_index, arr.shape, order="F")
    parts = []
    for sgn in (+1.0, -1.0):
        a = np.array(arr, order="F", copy=True)
        a[ix] += sgn * h
        ctx.set_triples_inputs(_replace(inp, **{name: a}), (precision,))
        _, F, P, _ = ctx.triples_forward(precision, with_F=True, with_partials=True)
        parts.append(P)
    ijk = oracle.unique_ijk(inp.o)
    w = np.where(ijk[:, 0] == ijk[:, 2], 0.0, np.where((ijk[:, 0] == ijk[:, 1]) | (ijk[:, 1] == ijk[:, 2]), 1.0, 2.0))
    diff = (parts[0] - parts[1]) * w[:, None]
    return float(np.sum(diff[diff != 0.0])) / (2 * h)


# h per input: 1e-5 for T2, (vvvo) and (ooov). E is linear in (ovov) and quadratic in T1, and
# their derivatives are ~1e-4 of the others (V carries a factor t1 ~ 1e-3), so at h = 1e-5 the
# rounding of the ~|E|/n partials (~1e-16 |E| / h) is ~1e-6 of those derivatives; the central
# difference of a quadratic is exact for every h, so h = 1e-3 is used there.
FD_H = {"ovvv": 1e-5, "ovoo": 1e-5, "t2": 1e-5, "ovov": 1e-3, "t1": 1e-3}


@pytest.mark.parametrize("o,v,seed", [(5, 24, 6), (6, 40, 7)])
def test_backward_against_local_central_differences(ctx, o, v, seed):
    inp = make_triples_inputs(o, v, seed)
    ctx.set_triples_inputs(inp, ("fp64",))
    g = ctx.triples_backward("fp64")
    rng = np.random.default_rng(seed)
    for name in NAMES:
        gmax = np.abs(g[name]).max()
        for f in rng.choice(getattr(inp, name).size, size=4, replace=False):
            _say(f"case fd o={o} v={v} {name}[{f}]")
            fd = local_fd(ctx, inp, name, int(f), FD_H[name])
            ix = np.unravel_index(int(f), getattr(inp, name).shape, order="F")
            ref = g[name][ix]
            assert abs(fd - ref) <= 1e-6 * max(abs(ref), 1e-3 * gmax), (name, ix, fd, ref)
