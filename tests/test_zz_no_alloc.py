"""scess_t_forward_noalloc (the (T) forward that allocates no device memory) through the C ABI against
scess_t_forward on the same resident inputs: bitwise equal E(T) in every precision, on two calls in a
row, the second at a smaller size (the shared scratch keeps the last call's contents, so whatever must
start at zero has to be cleared by the call). Sizes past its scratch are rejected with an error."""

import ctypes
import sys

import pytest

from scess_amd import gpu
from scess_amd.synthetic import make_triples_inputs

pytestmark = pytest.mark.gpu
MODES = ("fp64", "fp32", "bf16x3", "ozaki")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(ctx, mode):
    E = ctypes.c_double(0.0)
    rc = ctx.lib.scess_t_forward_noalloc(ctx.h, gpu.PRECISIONS[mode], ctypes.byref(E))
    return rc, E.value


def test_noalloc_forward_matches_forward():
    with gpu.Context(0) as ctx:
        for (o, v, seed) in ((9, 150, 31), (4, 37, 32)):  # the second call is smaller
            _say(f"case noalloc o={o} v={v}")
            ctx.set_triples_inputs(make_triples_inputs(o, v, seed), MODES)
            for mode in MODES:
                rc, E = _noalloc(ctx, mode)
                assert rc == gpu.SCESS_OK, ctx.lib.scess_last_error(ctx.h)
                Eref, _ = ctx.triples_forward(mode)
                _say(f"  {mode}: noalloc {E!r} forward {Eref!r}")
                assert E == Eref, (mode, E, Eref)


def test_noalloc_rejects_sizes_past_its_scratch():
    with gpu.Context(0) as ctx:
        _say("case noalloc too large: o=2 v=170")
        ctx.set_triples_inputs(make_triples_inputs(2, 170, 33), ("fp64",))
        rc, E = _noalloc(ctx, "fp64")
        assert rc == gpu.SCESS_ERR_ARG and E == 0.0
        assert b"scratch" in ctx.lib.scess_last_error(ctx.h)
