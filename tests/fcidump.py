"""FCIDUMP writer for the command-line runner tests (test helper)."""


def write_fcidump(path, h, eri, nelec, core, orbital_energies=None):
    """FCIDUMP of MO integrals: h (n x n), chemists' (pq|rs) (C order), each unique element once; with
    orbital_energies, also the `e_p p 0 0 0` lines some programs write after the one-electron block."""
    n = h.shape[0]
    with open(path, "w") as f:
        f.write(f" &FCI NORB={n},NELEC={nelec},MS2=0,\n  ORBSYM={','.join(['1'] * n)},\n  ISYM=1,\n &END\n")
        for p in range(n):
            for q in range(p + 1):
                for r in range(n):
                    for s in range(r + 1):
                        if p * (p + 1) // 2 + q >= r * (r + 1) // 2 + s and eri[p, q, r, s] != 0.0:
                            f.write(f"{eri[p, q, r, s]:.17e} {p + 1} {q + 1} {r + 1} {s + 1}\n")
        for p in range(n):
            for q in range(p + 1):
                f.write(f"{h[p, q]:.17e} {p + 1} {q + 1} 0 0\n")
        if orbital_energies is not None:
            for p in range(n):
                f.write(f"{orbital_energies[p]:.17e} {p + 1} 0 0 0\n")
        f.write(f"{core:.17e} 0 0 0 0\n")
