"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module to check the GPU results and as the reported
CPU baseline. The product path never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build(arch: str | None = None, out: str | None = None) -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    cmd = ["make", "-s", "-C", _HERE]
    if arch:
        cmd.append(f"ARCH={arch}")
    if out:
        cmd += [f"OUT={out}", "-B"]
    subprocess.run(cmd, check=True)
    return out or os.path.join(_HERE, "liboracle.so")


def _bind(lib):
    lib.oracle_t_abc.restype = ctypes.c_double
    lib.oracle_t_abc.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                 _i64p, ctypes.c_int64, ctypes.c_int, _dp]
    lib.oracle_t_abc_prepare.restype = ctypes.c_void_p
    lib.oracle_t_abc_prepare.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_t_abc_eval.restype = ctypes.c_double
    lib.oracle_t_abc_eval.argtypes = [ctypes.c_void_p, _i64p, ctypes.c_int64, ctypes.c_int, _dp]
    lib.oracle_t_abc_free.restype = None
    lib.oracle_t_abc_free.argtypes = [ctypes.c_void_p]
    lib.oracle_t_ijk.restype = ctypes.c_double
    lib.oracle_t_ijk.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                 _i64p, ctypes.c_int64, ctypes.c_int, _dp]
    lib.oracle_t_brute.restype = ctypes.c_double
    lib.oracle_t_brute.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_t_grad.restype = None
    lib.oracle_t_grad.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                  _dp, _dp, _dp, _dp, _dp]
    lib.oracle_t_grad_items.restype = None
    lib.oracle_t_grad_items.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _i64p, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_set_dgemm.restype = None
    lib.oracle_set_dgemm.argtypes = [ctypes.c_void_p]
    lib.oracle_num_unique_triples.restype = ctypes.c_int64
    lib.oracle_num_unique_triples.argtypes = [ctypes.c_int64]
    lib.oracle_num_unique_ijk.restype = ctypes.c_int64
    lib.oracle_num_unique_ijk.argtypes = [ctypes.c_int64]
    lib.oracle_unique_ijk.restype = None
    lib.oracle_unique_ijk.argtypes = [ctypes.c_int64, _i32p]
    return lib


def load(path: str | None = None):
    """Load (building first if needed) the oracle library."""
    global _LIB
    if path is not None:
        return _bind(ctypes.CDLL(path))
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = _bind(ctypes.CDLL(so))
    return _LIB


def _f(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    if a.ndim > 1 and not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    return a


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _args(inp):
    arrs = [_f(inp.e), _f(inp.t1), _f(inp.t2), _f(inp.ovvv), _f(inp.ovoo), _f(inp.ovov)]
    return arrs


def blas_dgemm():
    """Address of the CBLAS dgemm with 64-bit integers of numpy's bundled OpenBLAS (already loaded by
    numpy), with its own threading set to one thread: for oracle_set_dgemm, which runs GetW's products in
    BLAS as the reference does (EIGEN_USE_BLAS) while the oracle spreads the triples over threads."""
    import glob
    libs = glob.glob(os.path.join(os.path.dirname(np.__file__), "..", "numpy.libs", "*openblas*"))
    if not libs:
        raise RuntimeError("numpy's OpenBLAS not found")
    blas = ctypes.CDLL(libs[0])
    blas.scipy_openblas_set_num_threads64_.argtypes = [ctypes.c_int]
    blas.scipy_openblas_set_num_threads64_(1)
    return ctypes.cast(blas.scipy_cblas_dgemm64_, ctypes.c_void_p).value


def num_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)))


def t_abc(inp, items=None, nthreads=None, item_energy=False, lib=None):
    """E(T) by the abc-driven restatement of RCCSD_TriplesCorrection.

    items: optional int64 indices of unique triples a>=b>=c to evaluate (then the
    return value is 2 * the sum over those). item_energy=True also returns the
    per-triple energies."""
    lib = lib or load()
    arrs = _args(inp)
    it = None if items is None else np.ascontiguousarray(items, dtype=np.int64)
    n = lib.oracle_num_unique_triples(inp.v) if it is None else len(it)
    out = np.zeros(max(n, 1)) if item_energy else None
    E = lib.oracle_t_abc(inp.o, inp.v, *[_p(a) for a in arrs],
                         it.ctypes.data_as(_i64p) if it is not None else None,
                         0 if it is None else len(it), nthreads or num_threads(), _p(out))
    return (E, out[:n]) if item_energy else E


class AbcPrepared:
    """Repacked oracle inputs, for timing the per-triple work apart from the one-time repack."""

    def __init__(self, inp, lib=None):
        self.lib = lib or load()
        self.arrs = _args(inp)
        self.v = inp.v
        self.h = self.lib.oracle_t_abc_prepare(inp.o, inp.v, *[_p(a) for a in self.arrs])

    def eval(self, items=None, nthreads=None):
        it = None if items is None else np.ascontiguousarray(items, dtype=np.int64)
        return self.lib.oracle_t_abc_eval(self.h, it.ctypes.data_as(_i64p) if it is not None else None,
                                          0 if it is None else len(it), nthreads or num_threads(), None)

    def close(self):
        if self.h:
            self.lib.oracle_t_abc_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def t_ijk(inp, items=None, nthreads=None, lib=None):
    """(E, F) by the ijk-driven form: F[t] = sum_abc W r3(Y)/D for the listed unique
    occupied triples (all when items is None), E = sum 2/deg F."""
    lib = lib or load()
    arrs = _args(inp)
    it = None if items is None else np.ascontiguousarray(items, dtype=np.int64)
    n = lib.oracle_num_unique_ijk(inp.o) if it is None else len(it)
    F = np.zeros(max(n, 1))
    E = lib.oracle_t_ijk(inp.o, inp.v, *[_p(a) for a in arrs],
                         it.ctypes.data_as(_i64p) if it is not None else None,
                         0 if it is None else len(it), nthreads or num_threads(), _p(F))
    return E, F[:n]


def t_brute(inp, lib=None):
    lib = lib or load()
    arrs = _args(inp)
    return lib.oracle_t_brute(inp.o, inp.v, *[_p(a) for a in arrs])


def t_grad(inp, items=None, lib=None):
    """dict of FP64 gradients of E(T) in the input layouts; with items (indices into unique_ijk's
    order), the gradient of sum over those occupied triples of 2/deg F_ijk only."""
    lib = lib or load()
    arrs = _args(inp)
    g = {k: np.zeros(getattr(inp, k).shape, order="F") for k in ("ovvv", "ovoo", "ovov", "t1", "t2")}
    outs = (_p(g["ovvv"]), _p(g["ovoo"]), _p(g["ovov"]), _p(g["t1"]), _p(g["t2"]))
    if items is None:
        lib.oracle_t_grad(inp.o, inp.v, *[_p(a) for a in arrs], *outs)
    else:
        it = np.ascontiguousarray(items, dtype=np.int64)
        lib.oracle_t_grad_items(inp.o, inp.v, *[_p(a) for a in arrs], it.ctypes.data_as(_i64p), len(it), *outs)
    return g


def unique_ijk(o: int, lib=None) -> np.ndarray:
    lib = lib or load()
    n = lib.oracle_num_unique_ijk(o)
    out = np.zeros((max(n, 1), 3), dtype=np.int32)
    lib.oracle_unique_ijk(o, out.ctypes.data_as(_i32p))
    return out[:n]
