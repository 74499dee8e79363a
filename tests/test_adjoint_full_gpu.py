"""The (T) adjoint at o=20, v=200 (FP64 W recompute), through the C ABI.

* 64 seeded entries of dE/dT2, dE/d(vvvo) and dE/d(ooov) against central differences (h = 1e-5),
  1e-6 relative, in two halves of 32 entries. Each difference is local: the forward returns per-(ijk,
  orbit-run) partials, the partials an entry does not reach are bitwise equal in E(x+h) and E(x-h), so
  only the touched ones are summed (otherwise eps |E| / h swamps the bar). E is quadratic in every
  single entry, so the central difference is exact up to rounding.
* The backward of single occupied triples (a one-item shard of the work list: one with distinct
  indices, one with two equal indices, which runs the halved W-build and the twin back-contractions)
  against the oracle's FP64 backward of the same triple, 1e-11 of the largest gradient entry.
"""

import sys

import numpy as np
import pytest

import oracle
from scess_amd import gpu
from scess_amd.synthetic import TriplesInputs, make_triples_inputs
from scess_amd.triples import ijk_weights, unique_ijk

pytestmark = pytest.mark.gpu

O, V, SEED = 20, 200, 1


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c4():
    _say(f"case adjoint o={O} v={V}: generating")
    inp = make_triples_inputs(O, V, SEED)
    ctx = gpu.Context(0)
    ctx.set_triples_inputs(inp, ("fp64",))
    g = ctx.triples_backward("fp64")
    yield inp, ctx, g
    ctx.close()


def _replace(inp, **kw):
    d = dict(o=inp.o, v=inp.v, e=inp.e, t1=inp.t1, t2=inp.t2, ovvv=inp.ovvv, ovoo=inp.ovoo, ovov=inp.ovov)
    d.update(kw)
    return TriplesInputs(**d)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("half", [0, 1])
def test_backward_central_differences(c4, half):
    inp, ctx, g = c4
    wgt = ijk_weights(O)
    rng = np.random.default_rng(SEED + 1000 + half)
    h = 1e-5
    worst = 0.0
    for n in range(32):
        name = ("t2", "ovvv", "ovoo")[int(rng.integers(3))]
        arr = getattr(inp, name)
        ix = np.unravel_index(int(rng.integers(arr.size)), arr.shape, order="F")
        parts = []
        for sgn in (1.0, -1.0):
            a = np.array(arr, order="F", copy=True)
            a[ix] += sgn * h
            ctx.set_triples_inputs(_replace(inp, **{name: a}), ("fp64",))
            _, _, P, _ = ctx.triples_forward("fp64", with_F=True, with_partials=True)
            parts.append(P)
        diff = (parts[0] - parts[1]) * wgt[:, None]
        fd = float(np.sum(diff[diff != 0.0])) / (2 * h)
        an = float(g[name][ix])
        rel = abs(fd - an) / max(abs(an), 1e-300)
        worst = max(worst, rel)
        assert rel <= 1e-6, (name, ix, an, fd, rel)
    ctx.set_triples_inputs(inp, ("fp64",))
    _say(f"  half {half}: 32 entries, max relative error {worst:.2e}")


@pytest.mark.timeout(600)
def test_backward_single_triples_against_oracle(c4):
    inp, ctx, g_all = c4
    t = unique_ijk(O)
    work = np.flatnonzero(t[:, 0] != t[:, 2])  # the library's work list: unique triples with i != k
    rng = np.random.default_rng(4242)
    distinct = np.flatnonzero((t[work, 0] > t[work, 1]) & (t[work, 1] > t[work, 2]))
    twoeq = np.flatnonzero((t[work, 0] == t[work, 1]) | (t[work, 1] == t[work, 2]))
    try:
        for s in (int(rng.choice(distinct)), int(rng.choice(twoeq))):
            tri = tuple(int(x) for x in t[work[s]])
            ctx.set_shard(s, len(work))  # shard s of len(work) is work item s alone
            gg = ctx.triples_backward("fp64")
            go = oracle.t_grad(inp, items=[work[s]])
            for k in go:
                scale = np.abs(go[k]).max()
                err = np.abs(gg[k] - go[k]).max()
                _say(f"  triple {tri} d{k}: max |gpu - oracle| {err:.2e} of max {scale:.2e}")
                assert err <= 1e-11 * scale, (tri, k, err, scale)
                assert scale > 0.0 or k == "t1"
    finally:
        ctx.set_shard(0, 1)
