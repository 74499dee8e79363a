"""Pins the CPU oracle (oracle/triples_oracle.c) before it is trusted as the checker.

Checks, each chosen so that a plausible mistake in the restatement of
src/Quantum/CC/RCCSD_T.cpp:17-168 fails one of them:
  * abc-driven (the reference's loop structure) == ijk-driven == brute-force
    6-index sum (a wrong symmetrisation index, r3 coefficient or degeneracy
    weight breaks the agreement of differently organised sums);
  * == an independent spin-orbital (T) written from the textbook formula
    (pins the closed-shell formula itself: the 1/2 in V, r3, E = 2 sum);
  * E = 0 for o = 1 and for v = 1 (r3 annihilates i=j=k and a=b=c);
  * invariance under a uniform shift of all orbital energies and under
    relabelling occupied / virtual orbitals (exercises a=b, b=c weights);
  * E(s * integrals, t1 = 0) = s^2 E;
  * the FP64 backward against central differences (E is quadratic in every
    single input entry, so central differences are exact up to rounding);
  * the thread count does not change the result.
"""

import numpy as np
import pytest

import oracle
from scess_amd.synthetic import TriplesInputs, full_eri_chem, make_triples_inputs
from spin_orbital_t import spin_orbital_triples


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("o,v,seed", [(2, 3, 1), (3, 4, 2), (3, 5, 3), (4, 4, 4), (4, 6, 5)])
def test_three_forms_agree(o, v, seed):
    inp = make_triples_inputs(o, v, seed)
    Ea = oracle.t_abc(inp)
    Ei, _ = oracle.t_ijk(inp)
    Eb = oracle.t_brute(inp)
    assert _rel(Ei, Ea) < 1e-13
    assert _rel(Eb, Ea) < 1e-13


@pytest.mark.parametrize("o,v,seed", [(2, 3, 5), (3, 4, 5), (2, 5, 6), (3, 3, 7)])
def test_spin_orbital_textbook(o, v, seed):
    inp = make_triples_inputs(o, v, seed, keep_factor=True)
    E = oracle.t_abc(inp)
    Es = spin_orbital_triples(inp.e, np.asarray(inp.t1), np.asarray(inp.t2), full_eri_chem(inp.L), o)
    assert _rel(E, Es) < 1e-12


def test_zero_for_single_occupied_or_virtual():
    inp = make_triples_inputs(1, 6, 1)
    assert abs(oracle.t_abc(inp)) < 1e-14
    assert abs(oracle.t_ijk(inp)[0]) < 1e-14
    inp = make_triples_inputs(5, 1, 1)
    assert abs(oracle.t_abc(inp)) < 1e-14
    assert abs(oracle.t_ijk(inp)[0]) < 1e-14


def _replace(inp, **kw):
    d = dict(o=inp.o, v=inp.v, e=inp.e, t1=inp.t1, t2=inp.t2, ovvv=inp.ovvv, ovoo=inp.ovoo, ovov=inp.ovov)
    d.update(kw)
    return TriplesInputs(**{k: (np.asfortranarray(x) if isinstance(x, np.ndarray) and x.ndim > 1 else x)
                            for k, x in d.items()})


def test_energy_shift_invariance():
    inp = make_triples_inputs(4, 7, 11)
    E = oracle.t_abc(inp)
    E2 = oracle.t_abc(_replace(inp, e=inp.e + 0.37))
    assert _rel(E2, E) < 1e-12


def test_relabelling_invariance():
    inp = make_triples_inputs(4, 7, 12)
    o, v = inp.o, inp.v
    E = oracle.t_abc(inp)
    rng = np.random.default_rng(0)
    po = rng.permutation(o)
    pv = rng.permutation(v)
    e = np.concatenate([inp.e[:o][po], inp.e[o:][pv]])
    t1 = inp.t1[np.ix_(po, pv)]
    t2 = inp.t2[np.ix_(po, po, pv, pv)]
    ovvv = inp.ovvv[np.ix_(po, pv, pv, pv)]
    ovoo = inp.ovoo[np.ix_(po, pv, po, po)]
    ovov = inp.ovov[np.ix_(po, pv, po, pv)]
    E2 = oracle.t_abc(_replace(inp, e=e, t1=t1, t2=t2, ovvv=ovvv, ovoo=ovoo, ovov=ovov))
    assert _rel(E2, E) < 1e-12


def test_degenerate_virtual_energies():
    # equal virtual energies exercise nothing special in the formula but the
    # relabelling of equal-energy orbitals must stay exact
    inp = make_triples_inputs(3, 6, 13)
    e = inp.e.copy()
    e[inp.o + 2] = e[inp.o + 3]
    inp2 = _replace(inp, e=e)
    assert _rel(oracle.t_abc(inp2), oracle.t_ijk(inp2)[0]) < 1e-13


def test_quadratic_scaling():
    inp = make_triples_inputs(3, 6, 14)
    inp0 = _replace(inp, t1=np.zeros_like(inp.t1))
    E0 = oracle.t_abc(inp0)
    s = 1.7
    Es = oracle.t_abc(_replace(inp0, ovvv=inp.ovvv * s, ovoo=inp.ovoo * s))
    assert _rel(Es, s * s * E0) < 1e-12


def test_thread_count_independent():
    inp = make_triples_inputs(5, 12, 15)
    assert oracle.t_abc(inp, nthreads=1) == oracle.t_abc(inp, nthreads=7)
    assert oracle.t_ijk(inp, nthreads=1)[0] == oracle.t_ijk(inp, nthreads=5)[0]


def test_sampled_items_sum_to_total():
    inp = make_triples_inputs(4, 9, 16)
    E, per = oracle.t_abc(inp, item_energy=True)
    items = np.arange(0, len(per), 3)
    Es = oracle.t_abc(inp, items=items)
    assert _rel(Es, 2.0 * per[items].sum()) < 1e-13
    Ei, F = oracle.t_ijk(inp)
    ijk = oracle.unique_ijk(inp.o)
    sel = np.array([1, 4, 7, 9])
    Es2, F2 = oracle.t_ijk(inp, items=sel)
    assert np.array_equal(F2, F[sel])


@pytest.mark.parametrize("name", ["ovvv", "ovoo", "ovov", "t1", "t2"])
def test_backward_against_central_differences(name):
    inp = make_triples_inputs(3, 5, 17)
    g = oracle.t_grad(inp)[name]
    rng = np.random.default_rng(1)
    arr = getattr(inp, name)
    flat = rng.choice(arr.size, size=12, replace=False)
    h = 1e-3 if name != "t1" else 1e-4
    for f in flat:
        ix = np.unravel_index(f, arr.shape, order="F")
        plus = arr.copy(order="F")
        minus = arr.copy(order="F")
        plus[ix] += h
        minus[ix] -= h
        Ep = oracle.t_ijk(_replace(inp, **{name: plus}))[0]
        Em = oracle.t_ijk(_replace(inp, **{name: minus}))[0]
        fd = (Ep - Em) / (2 * h)
        assert abs(fd - g[ix]) <= 1e-7 * max(1.0, abs(g[ix])), (name, ix, fd, g[ix])


def test_product_ijk_enumeration_matches_oracle():
    """scess_amd.triples.unique_ijk / ijk_weights (used to weight per-ijk outputs) follow the oracle's order."""
    from scess_amd.triples import ijk_weights, unique_ijk
    for o in (1, 2, 5, 9):
        t = oracle.unique_ijk(o)
        assert np.array_equal(unique_ijk(o), t)
        w = ijk_weights(o)
        assert w.sum() == 2 * (o * (o - 1) * (o - 2) // 6) + 2 * o * (o - 1) // 2  # 2/deg per unique ijk


def test_grad_items_partition_and_local_differences():
    """oracle.t_grad(items=...) is the gradient of sum over the listed triples of 2/deg F_ijk: over a
    partition of the triples the parts add up to the full gradient (1e-12 relative), and one part
    matches central differences of the same restricted sum (E is quadratic per entry)."""
    from scess_amd.triples import ijk_weights
    inp = make_triples_inputs(4, 6, 17)
    full = oracle.t_grad(inp)
    n = len(oracle.unique_ijk(4))
    part_a = np.arange(0, n, 2)
    part_b = np.arange(1, n, 2)
    ga, gb = oracle.t_grad(inp, items=part_a), oracle.t_grad(inp, items=part_b)
    for k in full:
        scale = np.abs(full[k]).max()
        assert np.abs(ga[k] + gb[k] - full[k]).max() <= 1e-12 * scale, k
        assert np.abs(ga[k]).max() > 0.0
    w = ijk_weights(4)[part_a]
    rng = np.random.default_rng(3)
    h = 1e-3
    for name in ("ovvv", "ovoo", "t2", "ovov", "t1"):
        arr = getattr(inp, name)
        ix = np.unravel_index(int(rng.integers(arr.size)), arr.shape, order="F")
        vals = []
        for sgn in (1.0, -1.0):
            a = arr.copy(order="F")
            a[ix] += sgn * h
            vals.append(float(np.dot(w, oracle.t_ijk(_replace(inp, **{name: a}), items=part_a)[1])))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - ga[name][ix]) <= 1e-7 * max(1.0, abs(ga[name][ix])), (name, ix, fd, ga[name][ix])


def test_blas_ijk_restatement_matches_oracle():
    """oracle.triples_blas.t_ijk_blas (dense BLAS form used for single triples at large v) equals the C
    oracle's ijk-driven F_ijk, 1e-12 relative, on triples with distinct and with two equal indices."""
    from oracle.triples_blas import t_ijk_blas
    from scess_amd.triples import unique_ijk
    for (o, v, seed) in ((3, 5, 1), (5, 13, 2), (4, 24, 3)):
        inp = make_triples_inputs(o, v, seed)
        t = unique_ijk(o)
        _, F = oracle.t_ijk(inp)
        for n, (i, j, k) in enumerate(t):
            if i == k:
                continue
            Fb = t_ijk_blas(inp, int(i), int(j), int(k))
            assert abs(Fb - F[n]) <= 1e-12 * max(abs(F[n]), 1e-300), (o, v, (i, j, k), Fb, F[n])


def test_abc_oracle_with_blas_dgemm_equals_loops():
    """GetW's two products through BLAS dgemm (the reference's EIGEN_USE_BLAS path, used for the CPU
    baseline) give the same E(T) as the oracle's loop GEMMs, 1e-14 relative."""
    lib = oracle.load()
    fn = oracle.blas_dgemm()
    try:
        for (o, v, seed) in ((3, 5, 1), (6, 13, 2), (7, 20, 3)):
            inp = make_triples_inputs(o, v, seed)
            lib.oracle_set_dgemm(None)
            E0 = oracle.t_abc(inp)
            lib.oracle_set_dgemm(fn)
            E1 = oracle.t_abc(inp)
            assert abs(E1 - E0) <= 1e-14 * abs(E0), (o, v, E0, E1)
    finally:
        lib.oracle_set_dgemm(None)
