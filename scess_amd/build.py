"""Builds libscess_gpu.so (gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m scess_amd.build   (or __graft_entry__.build()).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libscess_gpu.so")
CLI = os.path.join(HERE, "scess_cc")
CPP_CHECK = os.path.join(ROOT, "tests", "cpp", "triples_function_check")
SO_CHECK = os.path.join(ROOT, "tests", "cpp", "spin_orbital_check")
SOURCES = ["triples_kernels.hip", "triples_adjoint.hip", "ccsd_engine.hip", "scess_capi.hip", "cc_api.cpp",
           "scf_host.cpp", "spin_orbital_cc.cpp", "regression.cpp"]
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        out = _run([hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
                    os.path.join(CSRC, src), "-o", obj])
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # command-line runner over the C++ method API (host code only)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", CLI + ".tmp", os.path.join(CSRC, "scess_cc_main.cpp"),
          "-L", HERE, "-lscess_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    # C++ check of the SCESS::CC method API (TriplesFunction), run on the GPU by tests/test_cpp_api_gpu.py
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", CPP_CHECK + ".tmp",
          CPP_CHECK + ".cpp", "-L", HERE, "-lscess_gpu", "-Wl,-rpath,$ORIGIN/../../scess_amd"])
    os.replace(CPP_CHECK + ".tmp", CPP_CHECK)
    # host-only check of the spin-orbital CCSD(T) (tests/test_spin_orbital_cpu.py)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", SO_CHECK + ".tmp",
          SO_CHECK + ".cpp", "-L", HERE, "-lscess_gpu", "-Wl,-rpath,$ORIGIN/../../scess_amd"])
    os.replace(SO_CHECK + ".tmp", SO_CHECK)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
