"""Parity of the gfx950 CCSD residual engine (scess_ccsd_*) with the CCSD oracle
(oracle/ccsd_oracle.py, restating src/Quantum/CC/RCCSD.cpp).

Tolerances: one UpdateAmplitudes step agrees to 1e-11 relative to the largest
amplitude (FP64, different summation order); converged correlation energies to
1e-9 Eh (the solver's own convergence threshold is 1e-10)."""

import sys

import numpy as np
import pytest

from oracle import ccsd_oracle as co
from scess_amd import gpu
from fcidump import write_fcidump
from scess_amd.synthetic import _draw, full_eri_chem

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(o, v, seed, scale):
    e, L, _ = _draw(o, v, seed)
    L = L * scale
    eri = full_eri_chem(L)
    O, V = slice(0, o), slice(o, None)
    blocks = {"ovov": eri[O, V, O, V], "oovv": eri[O, O, V, V], "ovvo": eri[O, V, V, O],
              "ovvv": eri[O, V, V, V], "ovoo": eri[O, V, O, O], "oooo": eri[O, O, O, O],
              "vvvv": eri[V, V, V, V]}
    chol = np.asfortranarray(L[:, o:, o:].transpose(1, 2, 0))
    return eri, e, blocks, chol


@pytest.fixture(scope="module")
def ctx():
    c = gpu.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("o,v,seed,use_chol", [(2, 3, 1, False), (3, 7, 2, False), (3, 7, 2, True),
                                               (5, 20, 3, True), (4, 37, 4, True)])
def test_update_matches_oracle(ctx, o, v, seed, use_chol):
    _say(f"case ccsd update o={o} v={v} chol={use_chol}")
    eri, e, blocks, chol = _system(o, v, seed, 0.3)
    if use_chol:
        blocks = dict(blocks, vvvv=None)
    ctx.set_ccsd_inputs(o, v, e, blocks, chol_vv=chol if use_chol else None)
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal((o, v)) * 1e-2
    t2 = rng.standard_normal((o, o, v, v)) * 1e-2
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    g1, g2 = ctx.ccsd_update(t1, t2)
    ints = co.make_blocks(eri, o)
    eia, eijab = co.denominators(e, o)
    r1, r2 = co.update_amplitudes(ints, eia, eijab, None, t1, t2)
    assert np.abs(g1 - r1).max() <= 1e-11 * np.abs(r1).max()
    assert np.abs(g2 - r2).max() <= 1e-11 * np.abs(r2).max()
    assert abs(ctx.ccsd_energy(t1, t2) - co.correlation_energy(ints, t1, t2)) < 1e-12 * max(1.0, abs(
        co.correlation_energy(ints, t1, t2)))


@pytest.mark.parametrize("o,v,seed,use_chol", [(2, 4, 2, False), (3, 8, 3, True)])
def test_solve_matches_oracle(ctx, o, v, seed, use_chol):
    _say(f"case ccsd solve o={o} v={v} chol={use_chol}")
    eri, e, blocks, chol = _system(o, v, seed, 0.25)
    if use_chol:
        blocks = dict(blocks, vvvv=None)
    ctx.set_ccsd_inputs(o, v, e, blocks, chol_vv=chol if use_chol else None)
    r = ctx.ccsd_solve()
    ro = co.solve_rccsd_diis(eri, e, o)
    assert r["converged"] and ro["converged"]
    assert abs(r["E_CCSD"] - ro["E_CCSD"]) < 1e-9
    assert np.abs(r["t2"] - ro["t2"]).max() < 1e-7


def test_f12_driver_and_coupling(ctx):
    _say("case ccsd f12 driver")
    o, v = 2, 3
    eri, e, blocks, chol = _system(o, v, 6, 0.3)
    rng = np.random.default_rng(1)
    res = rng.standard_normal((o, o, v, v)) * 1e-3
    res = res + res.transpose(1, 0, 3, 2)
    en = rng.standard_normal((o, o, v, v)) * 1e-3
    ctx.set_ccsd_inputs(o, v, e, blocks, t2_driver=res, f12_energy=en)
    r = ctx.ccsd_solve()
    ro = co.solve_rccsd_diis(eri, e, o, f12={"residual": res, "energy": en})
    assert r["converged"] and abs(r["E_CCSD"] - ro["E_CCSD"]) < 1e-9
    assert abs(r["E_F12_Coupling"] - co.coupling_energy({"energy": en}, ro["t2"])) < 1e-9


@pytest.mark.parametrize("basis", ["STO-3G", "cc-pVDZ"])
def test_water_end_to_end_on_gpu(ctx, basis):
    """Water at the reference's built-in geometry: CCSD (scess_ccsd_solve) and (T)
    (scess_t_forward, FP64) on the GPU against the published energies
    (Tests.cpp:43-44, 52-53) and against the CPU oracles (1e-10 Eh)."""
    _say(f"case water {basis}")
    import oracle
    from oracle import scf
    from scess_amd.synthetic import TriplesInputs
    ref = {"STO-3G": (-0.0494026324, -0.0000673499, 1e-7, 1e-8),
           "cc-pVDZ": (-0.2133077024, -0.0030574132, 1e-7, 1e-7)}[basis]
    r = scf.molecule_inputs(basis)
    o, eri, e = r["nocc"], r["eri"], r["e"]
    v = e.size - o
    O, V = slice(0, o), slice(o, None)
    blocks = {"ovov": eri[O, V, O, V], "oovv": eri[O, O, V, V], "ovvo": eri[O, V, V, O], "ovvv": eri[O, V, V, V],
              "ovoo": eri[O, V, O, O], "oooo": eri[O, O, O, O], "vvvv": eri[V, V, V, V]}
    ctx.set_ccsd_inputs(o, v, e, blocks)
    cc = ctx.ccsd_solve()
    assert cc["converged"]
    assert abs(cc["E_CCSD"] - ref[0]) <= ref[2]
    cco = co.solve_rccsd_diis(eri, e, o)
    assert abs(cc["E_CCSD"] - cco["E_CCSD"]) < 1e-10
    inp = TriplesInputs(o=o, v=v, e=e, t1=cc["t1"], t2=cc["t2"], ovvv=np.asfortranarray(eri[O, V, V, V]),
                        ovoo=np.asfortranarray(eri[O, V, O, O]), ovov=np.asfortranarray(eri[O, V, O, V]))
    ctx.set_triples_inputs(inp, ("fp64", "fp32", "bf16x3"))
    ET, _ = ctx.triples_forward("fp64")
    assert abs(ET - ref[1]) <= ref[3]
    assert abs(ET - oracle.t_abc(inp)) < 1e-10
    # at chemical magnitudes (|E(T)| ~ 1e-3 Eh) the mixed-precision contractions stay within 1 microhartree
    for mode in ("fp32", "bf16x3"):
        Em, _ = ctx.triples_forward(mode)
        _say(f"  {basis} {mode}: |dE(T)| = {abs(Em - ET):.3e} Eh")
        assert abs(Em - ET) <= 1e-6


@pytest.mark.parametrize("nao,nmo,phys", [(1, 1, False), (9, 7, False), (9, 9, True), (24, 17, True)])
def test_transform_tei(ctx, nao, nmo, phys):
    """TransformTEI on the GPU against numpy's four-index transform (FP64, 1e-12 relative).
    A transposed C or a missing chemists'->physicists' swap fails it: the tensor has no symmetry."""
    _say(f"case transform nao={nao} nmo={nmo} phys={phys}")
    rng = np.random.default_rng(nao * 100 + nmo)
    tei = rng.standard_normal((nao,) * 4)
    C = rng.standard_normal((nao, nmo))
    ref = np.einsum("abcd,ap,bq,cr,ds->pqrs", tei, C, C, C, C, optimize=True)
    if phys:
        ref = ref.transpose(0, 2, 1, 3)
    out = ctx.transform_tei(tei, C, chem_to_phys=phys)
    assert np.abs(out - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("basis", ["STO-3G", "cc-pVDZ"])
def test_cli_ccsd_t_water(tmp_path, basis):
    """scess_cc ccsd_t on a water FCIDUMP: the C++ method API (SolveRCCSD_DIIS, RCCSD_TriplesCorrection)
    end to end on the GPU, against the published energies (Tests.cpp:43-44, 52-53)."""
    import os
    import re
    import subprocess
    from oracle import scf
    _say(f"case cli water {basis}")
    cli = os.path.join(os.path.dirname(gpu.LIB_PATH), "scess_cc")
    ref = {"STO-3G": (-74.9629832288, -0.0494026324, -0.0000673499, 1e-6),
           "cc-pVDZ": (-76.0267859418, -0.2133077024, -0.0030574132, 1e-8)}[basis]
    atoms = scf.water_builtin()
    S, H, G = scf.ao_integrals(basis, atoms)
    E_HF, e, C = scf.rhf(S, H, G, 5, scf.nuclear_repulsion(atoms))
    eri = scf.mo_eri(G, C)
    dump = tmp_path / "FCIDUMP"
    write_fcidump(dump, C.T @ H @ C, eri, 10, scf.nuclear_repulsion(atoms))
    r = subprocess.run([cli, "ccsd_t", str(dump)], capture_output=True, text=True, timeout=120)
    _say(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    num = r"(-?[0-9.]+(?:e[-+]?\d+)?)"
    ehf = float(re.search(r"E\(Hartree Fock\) = " + num, r.stdout).group(1))
    ecc = float(re.search(r"E\(CCSD correlation\) = " + num, r.stdout).group(1))
    et = float(re.search(r"E\(T\) = " + num, r.stdout).group(1))
    assert abs(ehf - E_HF) < 1e-9
    assert abs(ehf - ref[0]) < ref[3]
    assert abs(ecc - ref[1]) < 1e-7
    assert abs(et - ref[2]) < 1e-7
    assert "Turning on DIIS" in r.stdout and "CCSD procedure converged" in r.stdout


@pytest.mark.parametrize("method,basis", [("ccsd_t", "cc-pVDZ"), ("ccsd", "STO-3G")])
def test_cli_builtin_water(tmp_path, method, basis):
    """`scess_cc ccsd_t H2O cc-pVDZ` as the reference's command line: host AO integrals and RHF, then the MO
    transform, CCSD and (T) on the GPU, against the published energies (Tests.cpp:41-53)."""
    import os
    import re
    import subprocess
    from test_cli import make_basis_dir
    _say(f"case cli builtin water {method} {basis}")
    cli = os.path.join(os.path.dirname(gpu.LIB_PATH), "scess_cc")
    ref = {"STO-3G": (-74.9629832288, -0.0494026324, -0.0000673499, 1e-6),
           "cc-pVDZ": (-76.0267859418, -0.2133077024, -0.0030574132, 1e-8)}[basis]
    env = dict(os.environ, SCESS_BASIS_DIR=make_basis_dir(tmp_path))
    r = subprocess.run([cli, method, "H2O", basis], capture_output=True, text=True, timeout=120, env=env)
    _say(r.stdout[-700:], r.stderr[-600:])
    assert r.returncode == 0
    num = r"(-?[0-9.]+(?:e[-+]?\d+)?)"
    ehf = float(re.search(r"E\(RHF\) = " + num, r.stdout).group(1))
    assert abs(ehf - ref[0]) < ref[3]
    if method == "ccsd_t":
        corr = float(re.search(r"E\(CCSD\(T\) correlation\) = " + num, r.stdout).group(1))
        assert abs(corr - (ref[1] + ref[2])) < 2e-7
        et = float(re.search(r"E\(T\) = " + num, r.stdout).group(1))
        assert abs(et - ref[2]) < 1e-7
    else:
        corr = float(re.search(r"E\(CCSD correlation\) = " + num, r.stdout).group(1))
        assert abs(corr - ref[1]) < 1e-7
    total = float(re.search(r"E\(Total\) = " + num + " hartrees\n*$", r.stdout).group(1))
    assert abs(total - (ehf + corr)) < 1e-9


def test_cli_builtin_hydrogen(tmp_path):
    """`scess_cc ccsd_t H2 STO-3G`: one occupied and one virtual orbital. CCSD is exact for two electrons,
    so E(RHF) + E(CCSD) is the FCI energy (Tests.cpp:34-38), and (T) vanishes with one occupied orbital."""
    import os
    import re
    import subprocess
    from test_cli import make_basis_dir
    _say("case cli builtin H2 ccsd_t STO-3G (o=1, v=1)")
    cli = os.path.join(os.path.dirname(gpu.LIB_PATH), "scess_cc")
    env = dict(os.environ, SCESS_BASIS_DIR=make_basis_dir(tmp_path))
    r = subprocess.run([cli, "ccsd_t", "H2", "STO-3G"], capture_output=True, text=True, timeout=120, env=env)
    _say(r.stdout[-700:], r.stderr[-600:])
    assert r.returncode == 0
    num = r"(-?[0-9.]+(?:e[-+]?\d+)?)"
    ehf = float(re.search(r"E\(RHF\) = " + num, r.stdout).group(1))
    et = float(re.search(r"E\(T\) = " + num, r.stdout).group(1))
    corr = float(re.search(r"E\(CCSD\(T\) correlation\) = " + num, r.stdout).group(1))
    total = float(re.search(r"E\(Total\) = " + num + " hartrees\n*$", r.stdout).group(1))
    assert abs(ehf - (-1.1175058851)) < 1e-8
    assert et == 0.0
    assert abs(corr - (-0.0193419373)) < 1e-8
    assert abs(total - (-1.1368478224)) < 1e-8


def test_water_baseline_geometry_gpu_vs_oracle(ctx):
    """Water at O-H 0.9572 A, H-O-H 104.52 deg, cc-pVDZ: CCSD and (T) on the GPU against the CPU
    oracles (1e-10 Eh); no published energy exists for this geometry."""
    import oracle
    from oracle import scf
    from scess_amd.synthetic import TriplesInputs
    _say("case water baseline geometry cc-pVDZ")
    r = scf.molecule_inputs("cc-pVDZ", scf.water_baseline_geometry())
    o, eri, e = r["nocc"], r["eri"], r["e"]
    v = e.size - o
    O, V = slice(0, o), slice(o, None)
    blocks = {"ovov": eri[O, V, O, V], "oovv": eri[O, O, V, V], "ovvo": eri[O, V, V, O], "ovvv": eri[O, V, V, V],
              "ovoo": eri[O, V, O, O], "oooo": eri[O, O, O, O], "vvvv": eri[V, V, V, V]}
    ctx.set_ccsd_inputs(o, v, e, blocks)
    cc = ctx.ccsd_solve()
    cco = co.solve_rccsd_diis(eri, e, o)
    assert cc["converged"] and abs(cc["E_CCSD"] - cco["E_CCSD"]) < 1e-10
    inp = TriplesInputs(o=o, v=v, e=e, t1=cc["t1"], t2=cc["t2"], ovvv=np.asfortranarray(eri[O, V, V, V]),
                        ovoo=np.asfortranarray(eri[O, V, O, O]), ovov=np.asfortranarray(eri[O, V, O, V]))
    ctx.set_triples_inputs(inp, ("fp64",))
    ET, _ = ctx.triples_forward("fp64")
    assert abs(ET - oracle.t_abc(inp)) < 1e-10
    _say(f"  E(CCSD)={cc['E_CCSD']!r} E(T)={ET!r}")


@pytest.mark.parametrize("mode,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_mixed_precision_contractions(ctx, mode, tol):
    """scess_ccsd_set_precision: one update with FP32 / split-BF16x3 products against the FP64 one
    (relative to the largest amplitude), and the water cc-pVDZ CCSD energy within 1 microhartree."""
    _say(f"case ccsd mixed {mode}")
    o, v = 5, 24
    eri, e, blocks, chol = _system(o, v, 9, 0.3)
    ctx.set_ccsd_inputs(o, v, e, blocks, chol_vv=chol)
    rng = np.random.default_rng(3)
    t1 = np.asfortranarray(rng.standard_normal((o, v)) * 1e-2)
    t2 = rng.standard_normal((o, o, v, v)) * 1e-2
    t2 = np.asfortranarray(t2 + t2.transpose(1, 0, 3, 2))
    try:
        r1, r2 = ctx.ccsd_update(t1, t2)
        ctx.set_ccsd_precision(mode)
        m1, m2 = ctx.ccsd_update(t1, t2)
        scale = np.abs(r2).max()
        err = max(np.abs(m2 - r2).max(), np.abs(m1 - r1).max()) / scale
        _say(f"  update rel err {err:.2e}")
        assert 0.0 < err <= tol
        from oracle import scf
        r = scf.molecule_inputs("cc-pVDZ")
        oo, eri_w, ew = r["nocc"], r["eri"], r["e"]
        O, V = slice(0, oo), slice(oo, None)
        bw = {"ovov": eri_w[O, V, O, V], "oovv": eri_w[O, O, V, V], "ovvo": eri_w[O, V, V, O],
              "ovvv": eri_w[O, V, V, V], "ovoo": eri_w[O, V, O, O], "oooo": eri_w[O, O, O, O],
              "vvvv": eri_w[V, V, V, V]}
        ctx.set_ccsd_inputs(oo, ew.size - oo, ew, bw)
        cm = ctx.ccsd_solve()
        ctx.set_ccsd_precision("fp64")
        c64 = ctx.ccsd_solve()
        _say(f"  water E(CCSD) {mode} - fp64 = {cm['E_CCSD'] - c64['E_CCSD']:.3e} Eh")
        assert cm["converged"] and abs(cm["E_CCSD"] - c64["E_CCSD"]) < 1e-6
    finally:
        ctx.set_ccsd_precision("fp64")


def test_cli_regression_runner(tmp_path):
    """`scess_cc test` as the reference's `SCESS test` (Application.cpp:49-58, Tests.cpp:494-737): the CC
    rows print [PASS] lines with the reference's tolerances (CheckEnergy, Tests.cpp:77-91), including the
    closed-shell vs spin-orbital (T) and CCSD cross-checks, and the exit code is 0 only when all pass.
    With one basis set missing, its sections fail and the exit code is 1."""
    import os
    import subprocess
    from test_cli import make_basis_dir
    _say("case cli regression runner")
    cli = os.path.join(os.path.dirname(gpu.LIB_PATH), "scess_cc")
    basis_dir = make_basis_dir(tmp_path)
    r = subprocess.run([cli, "test"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, SCESS_BASIS_DIR=basis_dir))
    _say(r.stdout[-2500:], r.stderr[-600:])
    lines = [l for l in r.stdout.splitlines() if l.strip().startswith("[")]
    assert r.returncode == 0
    assert len(lines) == 16 and all(l.strip().startswith("[PASS]") for l in lines)
    assert "closed-shell vs spin orbitals" in r.stdout and "spatial vs spin orbitals" in r.stdout
    assert "All regression tests passed" in r.stdout
    os.remove(os.path.join(basis_dir, "Pople", "6-31G.json"))
    r = subprocess.run([cli, "test"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, SCESS_BASIS_DIR=basis_dir))
    assert r.returncode == 1 and "[FAIL] H2O / 6-31G" in r.stdout and "1 regression test(s) failed" in r.stdout
