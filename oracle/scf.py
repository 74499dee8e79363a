"""Water inputs for the end-to-end checks: basis sets, AO integrals, RHF, MO
integrals (numpy + oracle/integrals.c).

TEST INFRASTRUCTURE ONLY. The reference computes these with libint2, its own
RHF (src/Quantum/HF/RHF.cpp:34-211, Pulay DIIS on the Fock matrix) and the
n^5 transform (src/Quantum/Integrals/IntegralTransforms.cpp:29-59). Only the
published energies (src/SCESS/Tests.cpp:31-53) pin them; this module exists
to reproduce those energies so the CCSD and (T) oracles, and the GPU path, can
be checked against the reference's own numbers.

Geometry: the reference's built-in water (src/Quantum/Molecule.cpp:194-205),
H at (+-0.7569520, -0.586349, 0) Angstrom, O at the origin, converted with
ANGSTROM_2_BOHR = 1.8897259886 (src/Quantum/Molecule.h:36). Basis data:
tests/golden/basis_h_o.json (the H and O entries of the reference's
Assets/Basis files). Spherical d shells (cc-pVDZ) span xy, xz, yz, x^2-y^2,
2z^2-x^2-y^2; energies depend only on that span.
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import load as _load_oracle

ANGSTROM_2_BOHR = 1.8897259886
HERE = os.path.dirname(os.path.abspath(__file__))
BASIS_FILE = os.path.join(os.path.dirname(HERE), "tests", "golden", "basis_h_o.json")
Z_OF = {"H": 1, "O": 8}


def water_builtin():
    """[(symbol, xyz in bohr)] of the reference's built-in water geometry."""
    a = ANGSTROM_2_BOHR
    return [("H", np.array([0.7569520, -0.586349, 0.0]) * a), ("O", np.zeros(3)),
            ("H", np.array([-0.7569520, -0.586349, 0.0]) * a)]


def water_baseline_geometry():
    """Water with O-H 0.9572 Angstrom and H-O-H 104.52 degrees (H in the xy plane, like the built-in one)."""
    a = ANGSTROM_2_BOHR
    r, th = 0.9572, np.deg2rad(104.52) / 2
    return [("H", np.array([r * np.sin(th), -r * np.cos(th), 0.0]) * a), ("O", np.zeros(3)),
            ("H", np.array([-r * np.sin(th), -r * np.cos(th), 0.0]) * a)]


def hydrogen_molecule_builtin():
    """H2 of the reference (src/Quantum/Molecule.cpp:207-217): bond 0.7122300374 Angstrom along x."""
    return [("H", np.zeros(3)), ("H", np.array([0.7122300374 * ANGSTROM_2_BOHR, 0.0, 0.0]))]


def _double_factorial(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def build_shells(basis: str, atoms):
    """Flat shell arrays (one shell per contraction column) and each shell's spherical flag."""
    data = json.load(open(BASIS_FILE))[basis]
    l_, npr, off, ctr, ex, co, sph = [], [], [], [], [], [], []
    for sym, xyz in atoms:
        for sh in data[str(Z_OF[sym])]["electron_shells"]:
            exps = np.array([float(x) for x in sh["exponents"]])
            am = sh["angular_momentum"]
            cols = sh["coefficients"]
            spherical = sh["function_type"] == "gto_spherical"
            for ci, col in enumerate(cols):
                l = am[ci] if len(am) > 1 else am[0]
                c = np.array([float(x) for x in col])
                keep = c != 0.0
                e_k, c_k = exps[keep], c[keep]
                # normalised primitives (x^l component), then the contracted function normalised
                nrm = (2 * e_k / np.pi) ** 0.75 * (4 * e_k) ** (l / 2) / np.sqrt(_double_factorial(2 * l - 1))
                cc = c_k * nrm
                ss = 0.0
                for a_, ca in zip(e_k, cc):
                    for b_, cb in zip(e_k, cc):
                        ss += ca * cb * (np.pi / (a_ + b_)) ** 1.5 * _double_factorial(2 * l - 1) / (2 * (a_ + b_)) ** l
                cc = cc / np.sqrt(ss)
                l_.append(l)
                npr.append(len(e_k))
                off.append(len(ex))
                ctr.extend(xyz.tolist())
                ex.extend(e_k.tolist())
                co.extend(cc.tolist())
                sph.append(spherical and l >= 2)
    return (np.array(l_, dtype=np.int32), np.array(npr, dtype=np.int32), np.array(off, dtype=np.int32),
            np.array(ctr, dtype=np.float64), np.array(ex, dtype=np.float64), np.array(co, dtype=np.float64), sph)


def _cart_to_sph(l_, sph):
    """Columns of the Cartesian -> basis transformation (span of the spherical d set)."""
    cols = []
    ncart_tot = sum((l + 1) * (l + 2) // 2 for l in l_)
    f = 0
    for l, s in zip(l_, sph):
        nc = (l + 1) * (l + 2) // 2
        if l == 2 and s:
            # Cartesian order xx, xy, xz, yy, yz, zz
            for vec in ([0, 1, 0, 0, 0, 0], [0, 0, 0, 0, 1, 0], [-1, 0, 0, -1, 0, 2], [0, 0, 1, 0, 0, 0],
                        [1, 0, 0, -1, 0, 0]):
                col = np.zeros(ncart_tot)
                col[f:f + 6] = vec
                cols.append(col)
        else:
            for k in range(nc):
                col = np.zeros(ncart_tot)
                col[f + k] = 1.0
                cols.append(col)
        f += nc
    return np.array(cols).T


def ao_integrals(basis: str, atoms):
    """S, H = T + V and the (pq|rs) AO tensor (chemists', C order) in the basis of the shells."""
    lib = _load_oracle()
    i32 = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.oracle_ncart_total.restype = ctypes.c_int64
    lib.oracle_ncart_total.argtypes = [ctypes.c_int32, i32]
    lib.oracle_one_electron.argtypes = [ctypes.c_int32, i32, i32, i32, dp, dp, dp, ctypes.c_int32, dp, dp, dp, dp, dp]
    lib.oracle_two_electron.argtypes = [ctypes.c_int32, i32, i32, i32, dp, dp, dp, dp]
    l_, npr, off, ctr, ex, co, sph = build_shells(basis, atoms)
    ns = len(l_)
    nb = int(lib.oracle_ncart_total(ns, l_.ctypes.data_as(i32)))
    S = np.zeros((nb, nb))
    T = np.zeros((nb, nb))
    V = np.zeros((nb, nb))
    G = np.zeros((nb, nb, nb, nb))
    axyz = np.array([x for _, x in atoms]).ravel()
    az = np.array([float(Z_OF[s]) for s, _ in atoms])
    args = (ns, l_.ctypes.data_as(i32), npr.ctypes.data_as(i32), off.ctypes.data_as(i32), ctr.ctypes.data_as(dp),
            ex.ctypes.data_as(dp), co.ctypes.data_as(dp))
    lib.oracle_one_electron(*args, len(atoms), axyz.ctypes.data_as(dp), az.ctypes.data_as(dp),
                            S.ctypes.data_as(dp), T.ctypes.data_as(dp), V.ctypes.data_as(dp))
    lib.oracle_two_electron(*args, G.ctypes.data_as(dp))
    X = _cart_to_sph(l_, sph)
    S = X.T @ S @ X
    H = X.T @ (T + V) @ X
    G = np.einsum("pqrs,pa,qb,rc,sd->abcd", G, X, X, X, X, optimize=True)
    return S, H, G


def nuclear_repulsion(atoms):
    e = 0.0
    for a in range(len(atoms)):
        for b in range(a):
            e += Z_OF[atoms[a][0]] * Z_OF[atoms[b][0]] / np.linalg.norm(atoms[a][1] - atoms[b][1])
    return e


def rhf(S, H, G, nocc, enuc, tol=1e-12, maxit=200):
    """Closed-shell RHF with Pulay DIIS on the Fock matrix. Returns (E_HF, eps, C)."""
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    Fs, Es = [], []
    Eold = 0.0
    for it in range(maxit):
        J = np.einsum("pqrs,rs->pq", G, D)
        K = np.einsum("prqs,rs->pq", G, D)
        F = H + J - 0.5 * K
        Eel = 0.5 * np.sum(D * (H + F))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        Fs.append(F)
        Es.append(err)
        if len(Fs) > 8:
            Fs.pop(0)
            Es.pop(0)
        n = len(Fs)
        if n > 1 and np.abs(err).max() > 1e-14:
            B = -np.ones((n + 1, n + 1))
            B[n, n] = 0.0
            for i in range(n):
                for j in range(n):
                    B[i, j] = np.sum(Es[i] * Es[j])
            rhs = np.zeros(n + 1)
            rhs[n] = -1.0
            c = np.linalg.lstsq(B, rhs, rcond=None)[0]
            F = sum(c[i] * Fs[i] for i in range(n))
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        Dn = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        if abs(Eel - Eold) < tol and np.abs(Dn - D).max() < 1e-9:
            D = Dn
            break
        D = Dn
        Eold = Eel
    J = np.einsum("pqrs,rs->pq", G, D)
    K = np.einsum("prqs,rs->pq", G, D)
    F = H + J - 0.5 * K
    e, Cp = np.linalg.eigh(X.T @ F @ X)
    C = X @ Cp
    Eel = 0.5 * np.sum(D * (H + F))
    return Eel + enuc, e, C


def mo_eri(G, C):
    """(pq|rs) over MOs, C order."""
    return np.einsum("pqrs,pa,qb,rc,sd->abcd", G, C, C, C, C, optimize=True)


def mp2(eri_mo, e, o):
    g = eri_mo[:o, o:, :o, o:]  # (ia|jb)
    eo, ev = e[:o], e[o:]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(g * (2 * g - g.transpose(0, 3, 2, 1)) / D))


def molecule_inputs(basis="cc-pVDZ", atoms=None, nelec=10):
    """RHF + MO integrals: dict(E_HF, e, eri (MO chemists', C order), nocc, E_MP2)."""
    atoms = atoms if atoms is not None else water_builtin()
    S, H, G = ao_integrals(basis, atoms)
    nocc = nelec // 2
    E, e, C = rhf(S, H, G, nocc, nuclear_repulsion(atoms))
    eri = mo_eri(G, C)
    return dict(E_HF=E, e=e, eri=eri, nocc=nocc, E_MP2=mp2(eri, e, nocc), C=C)
