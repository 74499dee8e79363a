"""The scess_cc command-line runner's host side (no GPU): FCIDUMP parsing, the Fock diagonal
and E_HF, against the RHF of the water STO-3G oracle (src/SCESS/Tests.cpp:41-44 pins that RHF)."""

import os
import re
import subprocess

import numpy as np
import pytest

from fcidump import write_fcidump
from scess_amd import gpu

CLI = os.path.join(os.path.dirname(gpu.LIB_PATH), "scess_cc")


@pytest.mark.skipif(not os.path.exists(CLI), reason="scess_cc not built")
def test_cli_hf_from_fcidump(tmp_path):
    from oracle import scf
    atoms = scf.water_builtin()
    S, H, G = scf.ao_integrals("STO-3G", atoms)
    E_HF, e, C = scf.rhf(S, H, G, 5, scf.nuclear_repulsion(atoms))
    dump = tmp_path / "FCIDUMP"
    write_fcidump(dump, C.T @ H @ C, scf.mo_eri(G, C), 10, scf.nuclear_repulsion(atoms))
    r = subprocess.run([CLI, "hf", str(dump)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    ehf = float(re.search(r"E\(Hartree Fock\) = (-?[0-9.]+)", r.stdout).group(1))
    eps = np.array([float(x) for x in re.findall(r"e\[\d+\] = (-?[0-9.eE+-]+)", r.stdout)])
    assert abs(ehf - E_HF) < 1e-9
    assert abs(ehf - (-74.9629832288)) < 1e-6  # Tests.cpp:41
    assert eps.shape == e.shape and np.abs(eps - e).max() < 1e-8
    assert "warning" not in r.stdout


@pytest.mark.skipif(not os.path.exists(CLI), reason="scess_cc not built")
def test_cli_rejects_bad_usage(tmp_path):
    assert subprocess.run([CLI], capture_output=True).returncode == 1
    assert subprocess.run([CLI, "fci", "x"], capture_output=True).returncode == 1
    r = subprocess.run([CLI, "hf", str(tmp_path / "missing")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
