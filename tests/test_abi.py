"""CPU-side checks of the C ABI: libscess_gpu.so loads and exports every symbol
include/scess_gpu.h declares; without a GPU, scess_create fails with an error
code instead of crashing. No compute call is made here."""

import ctypes
import os

import pytest

from scess_amd import gpu


def test_header_declares_expected_entry_points():
    names = set(gpu.header_functions())
    for n in ("scess_create", "scess_destroy", "scess_last_error", "scess_t_set_inputs", "scess_t_forward",
              "scess_t_forward_ex", "scess_t_num_ijk", "scess_t_num_partials", "scess_set_should_stop"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libscess_gpu.so not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in gpu.header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_returns_error_code():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libscess_gpu.so not built")
    import torch  # noqa: F401  (device count probe only)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = gpu.load()
    h = ctypes.c_void_p()
    assert lib.scess_create(ctypes.byref(h), 0) != 0
    assert lib.scess_last_error(None) == b"null context"


def test_library_exports_the_cxx_method_api():
    """include/scess_cc.h: the C++ mirror of SCESS::CC (CC.h:136-164) is defined in the same library."""
    import shutil
    import subprocess
    if not os.path.exists(gpu.LIB_PATH) or not shutil.which("nm"):
        pytest.skip("libscess_gpu.so not built or nm missing")
    out = subprocess.run(["nm", "-DC", "--defined-only", gpu.LIB_PATH], capture_output=True, text=True).stdout
    for n in ("SCESS::CC::SolveRCCSD_DIIS(", "SCESS::CC::RCCSD_TriplesCorrection(", "SCESS::CC::RCCSD_T(",
              "SCESS::CC::RCCSD_T_FromMO(", "SCESS::Integrals::TransformTEI(", "SCESS::HF::RHF(",
              "SCESS::Integrals::EvaluateMolecularIntegrals(", "SCESS::Basis::LoadBasisSet(",
              "SCESS::Molecule::GenerateWaterMolecule(", "SCESS::CC::TriplesFunction::TriplesFunction(",
              "SCESS::CC::TriplesFunction::forward(", "SCESS::CC::TriplesFunction::backward(",
              "SCESS::Driver::RHF(", "SCESS::Driver::CCSD(", "SCESS::Driver::CCSD_T(", "SCESS::CC::RCCSD_DIIS("):
        assert n in out, n
