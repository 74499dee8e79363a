"""ctypes binding of the C ABI of libscess_gpu.so (include/scess_gpu.h).

There is no CPU fallback: if the library is missing, or no GPU is present,
every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscess_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "scess_gpu.h")

SCESS_OK = 0
SCESS_ERR_ARG = -1
SCESS_ERR_HIP = -2
SCESS_ERR_OOM = -3
SCESS_ERR_STATE = -4
SCESS_CANCELLED = -5

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "ozaki": 3}

_dp = ctypes.POINTER(ctypes.c_double)


class TIn(ctypes.Structure):
    _fields_ = [("o", ctypes.c_int64), ("v", ctypes.c_int64), ("e", _dp), ("t1", _dp), ("t2", _dp),
                ("ovvv", _dp), ("ovoo", _dp), ("ovov", _dp)]


class TStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("w_build_ms", ctypes.c_double), ("epilogue_ms", ctypes.c_double),
                ("w_build_launches", ctypes.c_int64), ("epilogue_launches", ctypes.c_int64),
                ("n_ijk", ctypes.c_int64), ("flops_algorithmic", ctypes.c_double),
                ("flops_executed", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TGrad(ctypes.Structure):
    _fields_ = [("ovvv", _dp), ("ovoo", _dp), ("ovov", _dp), ("t1", _dp), ("t2", _dp)]


class CcsdIn(ctypes.Structure):
    _fields_ = [("o", ctypes.c_int64), ("v", ctypes.c_int64), ("e", _dp), ("ovov", _dp), ("oovv", _dp),
                ("ovvo", _dp), ("ovvv", _dp), ("ovoo", _dp), ("oooo", _dp), ("vvvv", _dp), ("chol_vv", _dp),
                ("naux", ctypes.c_int64), ("t2_driver", _dp), ("f12_energy", _dp)]


class CcsdResult(ctypes.Structure):
    _fields_ = [("E_CCSD", ctypes.c_double), ("E_F12_Coupling", ctypes.c_double), ("converged", ctypes.c_int),
                ("iterations", ctypes.c_int), ("last_dE", ctypes.c_double), ("last_rms", ctypes.c_double)]


STOP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double)

_LIB = None


class ScessError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scess error {code}: {msg}")
        self.code = code


def header_functions() -> list[str]:
    """Names of the functions declared in include/scess_gpu.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"\b(scess_[a-z0-9_]+)\s*\(", txt)


def load(path: str = LIB_PATH):
    """Load libscess_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m scess_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    c = ctypes
    lib.scess_create.argtypes = [c.POINTER(c.c_void_p), c.c_int]
    lib.scess_create.restype = c.c_int
    lib.scess_destroy.argtypes = [c.c_void_p]
    lib.scess_destroy.restype = None
    lib.scess_last_error.argtypes = [c.c_void_p]
    lib.scess_last_error.restype = c.c_char_p
    lib.scess_set_should_stop.argtypes = [c.c_void_p, STOP_FN, c.c_void_p]
    lib.scess_set_should_stop.restype = c.c_int
    lib.scess_set_batch.argtypes = [c.c_void_p, c.c_int64]
    lib.scess_set_batch.restype = c.c_int
    lib.scess_t_set_inputs.argtypes = [c.c_void_p, c.POINTER(TIn), c.c_int]
    lib.scess_t_set_inputs.restype = c.c_int
    lib.scess_t_set_int8_digits.argtypes = [c.c_void_p, c.c_int]
    lib.scess_t_set_int8_digits.restype = c.c_int
    lib.scess_t_forward.argtypes = [c.c_void_p, c.c_int, _dp, c.POINTER(TStats)]
    lib.scess_t_forward.restype = c.c_int
    lib.scess_t_forward_ex.argtypes = [c.c_void_p, c.c_int, _dp, _dp, _dp, c.POINTER(TStats)]
    lib.scess_t_forward_ex.restype = c.c_int
    lib.scess_t_set_shard.argtypes = [c.c_void_p, c.c_int64, c.c_int64]
    lib.scess_t_set_shard.restype = c.c_int
    lib.scess_t_num_ijk.argtypes = [c.c_void_p]
    lib.scess_t_num_ijk.restype = c.c_int64
    lib.scess_t_num_partials.argtypes = [c.c_void_p]
    lib.scess_t_num_partials.restype = c.c_int64
    lib.scess_t_forward_noalloc.argtypes = [c.c_void_p, c.c_int, _dp]
    lib.scess_t_forward_noalloc.restype = c.c_int
    lib.scess_t_backward.argtypes = [c.c_void_p, c.c_int, c.c_double, c.POINTER(TGrad)]
    lib.scess_t_backward.restype = c.c_int
    lib.scess_ccsd_set_inputs.argtypes = [c.c_void_p, c.POINTER(CcsdIn)]
    lib.scess_ccsd_set_inputs.restype = c.c_int
    lib.scess_ccsd_update.argtypes = [c.c_void_p, _dp, _dp, _dp, _dp]
    lib.scess_ccsd_update.restype = c.c_int
    lib.scess_ccsd_energy.argtypes = [c.c_void_p, _dp, _dp, _dp]
    lib.scess_ccsd_energy.restype = c.c_int
    lib.scess_ccsd_solve.argtypes = [c.c_void_p, _dp, _dp, c.POINTER(CcsdResult), LOG_FN, c.c_void_p]
    lib.scess_ccsd_solve.restype = c.c_int
    lib.scess_ccsd_set_precision.argtypes = [c.c_void_p, c.c_int]
    lib.scess_ccsd_set_precision.restype = c.c_int
    lib.scess_transform_tei.argtypes = [c.c_void_p, c.c_int64, c.c_int64, _dp, _dp, c.c_int, _dp]
    lib.scess_transform_tei.restype = c.c_int
    _LIB = lib
    return lib


def _f64(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    if a.ndim > 1 and not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    elif a.ndim <= 1:
        a = np.ascontiguousarray(a)
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Context:
    """One scess_ctx (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.scess_create(ctypes.byref(h), device)
        if rc != SCESS_OK:
            raise ScessError(rc, f"scess_create(device={device}) failed")
        self.h = h
        self._keep = None
        self._stop = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.scess_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, rc):
        if rc != SCESS_OK:
            msg = self.lib.scess_last_error(self.h).decode(errors="replace")
            raise ScessError(rc, msg)
        return rc

    def set_batch(self, n: int):
        self.check(self.lib.scess_set_batch(self.h, int(n)))

    def set_should_stop(self, fn):
        """fn() -> bool, polled between batches of occupied triples."""
        self._stop = STOP_FN(lambda _u: 1 if fn() else 0) if fn is not None else STOP_FN(0)
        self.check(self.lib.scess_set_should_stop(self.h, self._stop, None))

    def set_shard(self, shard: int, nshards: int):
        """Evaluate only shard `shard` of `nshards` equal ranges of occupied triples."""
        self.check(self.lib.scess_t_set_shard(self.h, int(shard), int(nshards)))

    def set_int8_digits(self, digits: int):
        """Digits per element of the "ozaki" mode (5, 6 or 7), used by the next set_triples_inputs."""
        self.check(self.lib.scess_t_set_int8_digits(self.h, int(digits)))

    def set_triples_inputs(self, inp, precisions=("fp64",)):
        arrs = dict(e=_f64(inp.e), t1=_f64(inp.t1), t2=_f64(inp.t2), ovvv=_f64(inp.ovvv),
                    ovoo=_f64(inp.ovoo), ovov=_f64(inp.ovov))
        o, v = int(inp.o), int(inp.v)
        expect = dict(e=(o + v,), t1=(o, v), t2=(o, o, v, v), ovvv=(o, v, v, v), ovoo=(o, v, o, o),
                      ovov=(o, v, o, v))
        for k, shp in expect.items():
            if arrs[k].shape != shp:
                raise ValueError(f"{k} has shape {arrs[k].shape}, expected {shp}")
        tin = TIn(o, v, *[_ptr(arrs[k]) for k in ("e", "t1", "t2", "ovvv", "ovoo", "ovov")])
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISIONS[p]
        self.check(self.lib.scess_t_set_inputs(self.h, ctypes.byref(tin), mask))
        self.o, self.v = o, v

    def triples_forward(self, precision="fp64", with_F=False, with_partials=False):
        """E(T) [, F_ijk][, partials] and the stats dict."""
        E = ctypes.c_double(0.0)
        st = TStats()
        F = P = None
        if with_F:
            F = np.zeros(self.lib.scess_t_num_ijk(self.h))
        if with_partials:
            P = np.zeros((self.lib.scess_t_num_ijk(self.h), self.lib.scess_t_num_partials(self.h)))
        rc = self.lib.scess_t_forward_ex(self.h, PRECISIONS[precision], ctypes.byref(E),
                                         _ptr(F) if F is not None else None,
                                         _ptr(P) if P is not None else None, ctypes.byref(st))
        self.check(rc)
        out = [E.value]
        if with_F:
            out.append(F)
        if with_partials:
            out.append(P)
        out.append(st.as_dict())
        return tuple(out)

    def triples_backward(self, precision="fp64", dE=1.0):
        """Gradients of dE * E(T) with respect to ovvv, ovoo, ovov, t1, t2 (input layouts)."""
        o, v = self.o, self.v
        g = {"ovvv": np.zeros((o, v, v, v), order="F"), "ovoo": np.zeros((o, v, o, o), order="F"),
             "ovov": np.zeros((o, v, o, v), order="F"), "t1": np.zeros((o, v), order="F"),
             "t2": np.zeros((o, o, v, v), order="F")}
        tg = TGrad(*[_ptr(g[k]) for k in ("ovvv", "ovoo", "ovov", "t1", "t2")])
        self.check(self.lib.scess_t_backward(self.h, PRECISIONS[precision], float(dE), ctypes.byref(tg)))
        return g

    # ------------------------------------------------------------------ CCSD
    def set_ccsd_inputs(self, o, v, e, blocks, chol_vv=None, t2_driver=None, f12_energy=None):
        """blocks: dict of chemists' blocks (ovov, oovv, ovvo, ovvv, ovoo, oooo and optionally vvvv),
        logical index order as in RCCSD.cpp:31-44; chol_vv: (a, c, P) factor of (ac|bd)."""
        keep = {k: _f64(blocks[k]) for k in ("ovov", "oovv", "ovvo", "ovvv", "ovoo", "oooo")}
        keep["e"] = _f64(e)
        vvvv = _f64(blocks["vvvv"]) if blocks.get("vvvv") is not None else None
        ch = _f64(chol_vv) if chol_vv is not None else None
        drv = _f64(t2_driver) if t2_driver is not None else None
        fen = _f64(f12_energy) if f12_energy is not None else None
        cin = CcsdIn(int(o), int(v), _ptr(keep["e"]), _ptr(keep["ovov"]), _ptr(keep["oovv"]), _ptr(keep["ovvo"]),
                     _ptr(keep["ovvv"]), _ptr(keep["ovoo"]), _ptr(keep["oooo"]),
                     _ptr(vvvv) if vvvv is not None else None, _ptr(ch) if ch is not None else None,
                     int(ch.shape[2]) if ch is not None else 0,
                     _ptr(drv) if drv is not None else None, _ptr(fen) if fen is not None else None)
        self.check(self.lib.scess_ccsd_set_inputs(self.h, ctypes.byref(cin)))
        self.cc_o, self.cc_v = int(o), int(v)

    def set_ccsd_precision(self, precision="fp64"):
        """Product precision of the CCSD residual contractions ("fp64", "fp32", "bf16x3")."""
        self.check(self.lib.scess_ccsd_set_precision(self.h, PRECISIONS[precision]))

    def ccsd_update(self, t1, t2):
        t1 = _f64(t1)
        t2 = _f64(t2)
        n1 = np.zeros(t1.shape, order="F")
        n2 = np.zeros(t2.shape, order="F")
        self.check(self.lib.scess_ccsd_update(self.h, _ptr(t1), _ptr(t2), _ptr(n1), _ptr(n2)))
        return n1, n2

    def ccsd_energy(self, t1, t2):
        t1 = _f64(t1)
        t2 = _f64(t2)
        e = ctypes.c_double(0.0)
        self.check(self.lib.scess_ccsd_energy(self.h, _ptr(t1), _ptr(t2), ctypes.byref(e)))
        return e.value

    def ccsd_solve(self, log=None):
        o, v = self.cc_o, self.cc_v
        t1 = np.zeros((o, v), order="F")
        t2 = np.zeros((o, o, v, v), order="F")
        res = CcsdResult()
        cb = LOG_FN(lambda _u, it, E, dE, rms: log(it, E, dE, rms)) if log else LOG_FN(0)
        self.check(self.lib.scess_ccsd_solve(self.h, _ptr(t1), _ptr(t2), ctypes.byref(res), cb, None))
        return {k: getattr(res, k) for k, _ in res._fields_} | {"t1": t1, "t2": t2}

    def transform_tei(self, tei, C, chem_to_phys=False):
        """TransformTEI (IntegralTransforms.cpp:29-59) on the GPU: AO (mu nu|kap lam) -> MO, optionally <pq|rs>."""
        tei = _f64(tei)
        C = _f64(C)
        nao, nmo = C.shape
        if tei.shape != (nao,) * 4:
            raise ValueError("tei must be nao^4")
        out = np.zeros((nmo,) * 4, order="F")
        self.check(self.lib.scess_transform_tei(self.h, nao, nmo, _ptr(tei), _ptr(C), int(bool(chem_to_phys)),
                                                _ptr(out)))
        return out
