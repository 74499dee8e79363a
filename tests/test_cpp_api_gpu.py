"""The C++ method API on the GPU: SCESS::CC::TriplesFunction (include/scess_cc.h), the autograd-Function
style forward/backward pair of (T), checked by the C++ program tests/cpp/triples_function_check.cpp
(built with the library): forward equals RCCSD_TriplesCorrection bitwise, backward is linear in dE and
agrees with central differences of RCCSD_TriplesCorrection for dE/d<pq|rs>, dE/dt1 and dE/dt2, the
gradient is zero outside the index patterns (T) reads, and cancellation returns 0.0."""
import os
import subprocess

import pytest

CHECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "triples_function_check")


@pytest.mark.gpu
def test_triples_function_forward_backward():
    assert os.path.exists(CHECK), f"{CHECK} is missing: build with __graft_entry__.build()"
    print("case TriplesFunction check o=4 v=9", flush=True)
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
