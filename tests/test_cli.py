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
def test_cli_hf_fcidump_with_orbital_energy_lines(tmp_path):
    """`e_p p 0 0 0` lines (written by some programs) are skipped: the run equals the one without them.
    A line with any other zero pattern (here `x 1 0 2 0`) is an error, not a silent write."""
    from oracle import scf
    atoms = scf.water_builtin()
    S, H, G = scf.ao_integrals("STO-3G", atoms)
    E_HF, e, C = scf.rhf(S, H, G, 5, scf.nuclear_repulsion(atoms))
    plain, with_e = tmp_path / "plain", tmp_path / "with_e"
    write_fcidump(plain, C.T @ H @ C, scf.mo_eri(G, C), 10, scf.nuclear_repulsion(atoms))
    write_fcidump(with_e, C.T @ H @ C, scf.mo_eri(G, C), 10, scf.nuclear_repulsion(atoms), orbital_energies=e + 7.0)
    r0 = subprocess.run([CLI, "hf", str(plain)], capture_output=True, text=True, timeout=60)
    r1 = subprocess.run([CLI, "hf", str(with_e)], capture_output=True, text=True, timeout=60)
    assert r0.returncode == 0 and r1.returncode == 0, r1.stderr
    assert r1.stdout == r0.stdout
    bad = tmp_path / "bad"
    bad.write_text(with_e.read_text() + "1.0 1 0 2 0\n")
    r2 = subprocess.run([CLI, "hf", str(bad)], capture_output=True, text=True, timeout=60)
    assert r2.returncode == 1 and "invalid index pattern" in r2.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="scess_cc not built")
def test_cli_rejects_bad_usage(tmp_path):
    assert subprocess.run([CLI], capture_output=True).returncode == 1
    assert subprocess.run([CLI, "fci", "x"], capture_output=True).returncode == 1
    r = subprocess.run([CLI, "hf", str(tmp_path / "missing")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


GOLDEN_BASIS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "basis_h_o.json")


def make_basis_dir(root):
    """The reference's Assets/Basis layout (family folders of BSE JSON files) with the H and O entries
    of cc-pVDZ and STO-3G from the golden fixture."""
    import json
    data = json.load(open(GOLDEN_BASIS))
    for fam, name in (("Dunning", "cc-pVDZ"), ("Pople", "STO-3G"), ("Pople", "6-31G")):
        os.makedirs(os.path.join(root, fam), exist_ok=True)
        with open(os.path.join(root, fam, name + ".json"), "w") as f:
            json.dump({"name": name, "elements": data[name]}, f)
    return str(root)


def _rhf(args, basis_dir):
    env = dict(os.environ, SCESS_BASIS_DIR=basis_dir)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.skipif(not os.path.exists(CLI), reason="scess_cc not built")
@pytest.mark.parametrize("mol,basis,ref,tol", [("H2O", "cc-pVDZ", -76.0267859418, 1e-8),
                                                ("H2O", "STO-3G", -74.9629832288, 1e-6),
                                                ("H2O", "6-31G", -75.9839769556, 1e-8),
                                                ("H2", "STO-3G", -1.1175058851, 1e-8)])
def test_cli_rhf_builtin_molecule(tmp_path, mol, basis, ref, tol):
    """`scess_cc rhf <molecule> <basis>`: the host AO integrals and RHF against the published RHF energies
    (src/SCESS/Tests.cpp:36-53, same tolerances as the oracle chain) and, for water, the oracle's own
    integrals + RHF (1e-9)."""
    r = _rhf(["rhf", mol, basis], make_basis_dir(tmp_path))
    assert r.returncode == 0, r.stderr
    e = float(re.search(r"E\(RHF\) = (-?[0-9.]+)", r.stdout).group(1))
    assert abs(e - ref) < tol
    if mol == "H2O" and basis != "6-31G":  # the oracle chain is pinned on the cc-pVDZ and STO-3G entries
        from oracle import scf
        atoms = scf.water_builtin()
        S, H, G = scf.ao_integrals(basis, atoms)
        E_HF = scf.rhf(S, H, G, 5, scf.nuclear_repulsion(atoms))[0]
        assert abs(e - E_HF) < 1e-9


@pytest.mark.skipif(not os.path.exists(CLI), reason="scess_cc not built")
def test_cli_basis_lookup_and_errors(tmp_path):
    d = make_basis_dir(tmp_path)
    ok = _rhf(["RHF", "H2O", "cc-pvdz"], d)  # method and basis names are case-insensitive
    assert ok.returncode == 0, ok.stderr
    path = os.path.join(d, "Pople", "STO-3G.json")
    by_path = _rhf(["rhf", "H2", path], d)  # a file path instead of a name
    assert by_path.returncode == 0 and "E(RHF) = -1.11750588" in by_path.stdout
    r = _rhf(["rhf", "H2O", "def2-TZVP"], d)
    assert r.returncode == 1 and "not an available basis set" in r.stderr
    r = _rhf(["rhf", "CH4", "STO-3G"], d)
    assert r.returncode == 1 and "not an available molecule" in r.stderr
    r = _rhf(["rhf", "LiH", "STO-3G"], d)  # the fixture holds H and O only
    assert r.returncode == 1 and "no entry for Li" in r.stderr
    r = _rhf(["mp2", "H2O", "STO-3G"], d)
    assert r.returncode == 1 and "not a valid method" in r.stderr
    env = {k: v for k, v in os.environ.items() if k != "SCESS_BASIS_DIR"}
    r = subprocess.run([CLI, "rhf", "H2O", "STO-3G"], capture_output=True, text=True, env=env)
    assert r.returncode == 1 and "SCESS_BASIS_DIR" in r.stderr
