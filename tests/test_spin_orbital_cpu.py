"""The host spin-orbital CCSD(T) of the regression runner (SCESS::CC::CCSD_DIIS_FromMO, written from the
Stanton-Gauss equations and the spin-orbital (T) expressions) on water / STO-3G and H2 / STO-3G, without
a GPU: host AO integrals and RHF, a plain-loop MO transform, then the spin-orbital solver. Pinned by the
reference's published energies (src/SCESS/Tests.cpp:36-44): E(RHF), E(CCSD) and E(T) of water, E(CCSD)
of H2 (= FCI - RHF, exact for two electrons) with E(T) = 0."""

import os
import subprocess

import pytest

from scess_amd import gpu

CHECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "spin_orbital_check")


def _run(mol, tmp_path):
    from test_cli import make_basis_dir
    env = dict(os.environ, SCESS_BASIS_DIR=make_basis_dir(tmp_path))
    r = subprocess.run([CHECK, mol], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    return [float(x) for x in r.stdout.split()]


@pytest.mark.skipif(not os.path.exists(CHECK) or not os.path.exists(gpu.LIB_PATH), reason="check not built")
def test_spin_orbital_ccsd_t_water_sto3g(tmp_path):
    E_HF, E_CCSD, E_T, conv = _run("H2O", tmp_path)
    assert conv == 1.0
    assert abs(E_HF - (-74.9629832288)) < 1e-6    # Tests.cpp:41 (its tolerance, :567)
    assert abs(E_CCSD - (-0.0494026324)) < 1e-7   # Tests.cpp:43
    assert abs(E_T - (-0.0000673499)) < 1e-8      # Tests.cpp:44


@pytest.mark.skipif(not os.path.exists(CHECK) or not os.path.exists(gpu.LIB_PATH), reason="check not built")
def test_spin_orbital_ccsd_h2_sto3g(tmp_path):
    E_HF, E_CCSD, E_T, conv = _run("H2", tmp_path)
    assert conv == 1.0
    assert abs(E_HF - (-1.1175058851)) < 1e-8     # Tests.cpp:36
    assert abs(E_CCSD - (-0.0193419373)) < 1e-8   # Tests.cpp:38
    assert abs(E_T) < 1e-14                        # one occupied spatial orbital: no triples
