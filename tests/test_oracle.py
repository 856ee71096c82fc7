"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (dropped term,
wrong sign or index, transposed operand) fails: the reference's own test
cases of test/core/test_neighborlist.py, closed forms, finite differences,
Newton's third law, brute force over an explicit image box, the NaCl
Madelung constant and exact one-step integrator formulas.
"""

import itertools
import math

import numpy as np
import pytest

import oracle
from kups_amd import workloads


def nl_set(pos, lattice, cutoff, excl=None, sizes=None):
    pos = np.asarray(pos, float)
    lat = np.asarray(lattice, float).reshape(-1, 3, 3)
    if sizes is None:
        sizes = [len(pos)]
    off = np.concatenate([[0], np.cumsum(sizes)])
    tril = lat[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]
    if excl is None:
        excl = np.arange(len(pos))
    cut = np.broadcast_to(np.asarray(cutoff, float), (len(sizes),))
    ij, sh = oracle.neighborlist(pos, excl, off, tril, cut)
    return {(int(i), int(j), tuple(int(x) for x in s)) for (i, j), s in zip(ij, sh)}


def pairs(es):
    return {(i, j) for i, j, _ in es}


# ---- neighbor list: the reference's own cases (test_neighborlist.py) ----

def test_nl_basic_chain():
    # 298-316: three particles in a row, only nearest pairs
    es = nl_set([[0, 0, 0], [1, 0, 0], [2.1, 0, 0]], np.eye(3) * 10, 1.5)
    assert pairs(es) == {(0, 1), (1, 0), (1, 2), (2, 1)}


def test_nl_exact_cutoff_excluded():
    # 335-354: distance == cutoff is not a neighbor (strict <)
    assert nl_set([[0, 0, 0], [1, 0, 0]], np.eye(3) * 10, 1.0) == set()
    assert len(nl_set([[0, 0, 0], [1, 0, 0]], np.eye(3) * 10, 1.0 + 1e-12)) == 2


def test_nl_periodic_boundary():
    # 356-370: 0.1 and 2.9 in a 3.0 box are 0.2 apart through the boundary
    es = nl_set([[0.1, 0, 0], [2.9, 0, 0]], np.eye(3) * 3, 1.0)
    assert es == {(0, 1, (-1, 0, 0)), (1, 0, (1, 0, 0))}


def test_nl_multi_system_isolation():
    es = nl_set([[0, 0, 0], [1, 0, 0], [0, 0, 0], [1, 0, 0]],
                [np.eye(3) * 10, np.eye(3) * 10], 1.5, sizes=[2, 2])
    assert pairs(es) == {(0, 1), (1, 0), (2, 3), (3, 2)}


def test_nl_self_images():
    # 1055-1116: one particle, 0.5 cell, cutoff 0.55 -> six self images
    es = nl_set([[0, 0, 0]], np.eye(3) * 0.5, 0.55, excl=[0])
    exp = {(0, 0, s) for s in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]}
    assert es == exp


def test_nl_infinite_cutoff_no_images():
    # 1118-1163
    es = nl_set([[0, 0, 0], [0.3, 0, 0]], np.eye(3), np.inf)
    assert es == {(0, 1, (0, 0, 0)), (1, 0, (0, 0, 0))}


def test_nl_exclusion_only_minimum_image():
    # 1165-1236
    es = nl_set([[0.1, 0.1, 0.1], [0.4, 0.1, 0.1]], np.eye(3) * 0.5, 0.4, excl=[0, 0])
    assert (0, 1, (-1, 0, 0)) not in es and (1, 0, (1, 0, 0)) not in es
    assert (0, 1, (0, 0, 0)) in es and (1, 0, (0, 0, 0)) in es


def test_nl_batched_small_and_large_cells():
    # 895-949
    es = nl_set([[0, 0, 0], [0.4, 0, 0], [0, 0, 0], [1, 0, 0]],
                [np.diag([1.0, 10, 10]), np.eye(3) * 10], [0.8, 1.5], sizes=[2, 2])
    p = pairs(es)
    assert {(0, 1), (1, 0), (2, 3), (3, 2)} <= p
    assert not any((i < 2) != (j < 2) for i, j in p)


def _brute_images(pos, lat, rc, nimg=3):
    """Every (i, j, s) with |r_j - r_i + s@L| < rc over an explicit image box."""
    out = set()
    n = len(pos)
    for i in range(n):
        for j in range(n):
            for s in itertools.product(range(-nimg, nimg + 1), repeat=3):
                if i == j and s == (0, 0, 0):
                    continue
                d = pos[j] - pos[i] + np.asarray(s) @ lat
                if d @ d < rc * rc:
                    out.add((i, j, s))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nl_matches_brute_force_large_box(seed):
    # 678-722 (compare to naive): no images needed, min image is the only image
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-5, 5, (15, 3))
    lat = np.eye(3) * 10.0
    assert nl_set(pos, lat, 3.0) == _brute_images(pos, lat, 3.0, nimg=1)


def test_nl_matches_brute_force_small_triclinic():
    # 1007-1053: triclinic cell narrower than 2 rc, wrapped positions: the image
    # grid of the reference covers every image within rc
    rng = np.random.default_rng(3)
    lat = np.array([[2.0, 0, 0], [0.4, 1.8, 0], [0.3, -0.2, 2.1]])
    f = rng.uniform(-0.5, 0.5, (5, 3))
    pos = f @ lat
    rc = 1.5
    got = nl_set(pos, lat, rc)
    assert got == _brute_images(pos, lat, rc, nimg=3)


# ---- Lennard-Jones ----

def _lj_two(r, sigma=3.405, eps=0.0103):
    pos = np.array([[0.0, 0, 0], [r, 0, 0]])
    e, f, w = oracle.lj_cells(pos, [0, 0], [0, 2], [40, 0, 40, 0, 0, 40], [10.0],
                              [[sigma]], [[eps]])
    return e[0], f, w[0]


def test_lj_closed_form_pair():
    # test_lennard_jones.py:172-289: u = 4 eps ((s/r)^12 - (s/r)^6)
    for r in [3.0, 3.405, 3.82, 5.0, 9.9]:
        e, f, w = _lj_two(r)
        x = 3.405 / r
        assert e == pytest.approx(4 * 0.0103 * (x**12 - x**6), rel=1e-12)
        dudr = 4 * 0.0103 * (-12 * x**12 + 6 * x**6) / r
        assert f[1, 0] == pytest.approx(-dudr, rel=1e-10)   # F_j = -dU/dx_j
        assert f[0, 0] == pytest.approx(dudr, rel=1e-10)
        assert w[0, 0] == pytest.approx(-dudr * r, rel=1e-10)  # W = r . F
    # beyond the hard cutoff: nothing
    assert _lj_two(10.5)[0] == 0.0


def test_lj_forces_finite_difference_and_third_law():
    b = workloads.batched_lj_boxes(sizes=[70, 200], seed=5)
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = [workloads.AR_CUTOFF] * 2
    e, f, w = oracle.lj_cells(b.positions, b.species, b.sys_off, b.tril, cut, sig, eps)
    np.testing.assert_allclose(f[:70].sum(0), 0, atol=1e-10)
    np.testing.assert_allclose(f[70:].sum(0), 0, atol=1e-10)
    h = 1e-5
    for atom in [3, 100, 150]:
        for k in range(3):
            p = b.positions.copy(); p[atom, k] += h
            ep = oracle.lj_cells(p, b.species, b.sys_off, b.tril, cut, sig, eps)[0].sum()
            p[atom, k] -= 2 * h
            em = oracle.lj_cells(p, b.species, b.sys_off, b.tril, cut, sig, eps)[0].sum()
            assert -(ep - em) / (2 * h) == pytest.approx(f[atom, k], rel=2e-5, abs=1e-8)


def test_lj_cells_equals_edge_list():
    b = workloads.batched_lj_boxes(sizes=[64, 90, 500], seed=7)  # 64, 90 need images
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = np.full(3, workloads.AR_CUTOFF)
    ij, sh = oracle.neighborlist(b.positions, np.arange(b.n), b.sys_off, b.tril, cut)
    e1, f1, w1 = oracle.lj_edges(b.positions, b.species, b.sys_off, b.tril, cut, sig, eps, ij, sh)
    e2, f2, w2 = oracle.lj_cells(b.positions, b.species, b.sys_off, b.tril, cut, sig, eps)
    np.testing.assert_allclose(e1, e2, rtol=1e-12)
    np.testing.assert_allclose(f1, f2, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(w1, w2, rtol=1e-9, atol=1e-12)


def test_lj_virial_is_strain_derivative():
    # W = -dU/d(strain): scale the box and positions by (1 + e) along x
    b = workloads.batched_lj_boxes(sizes=[300], seed=9)
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = [workloads.AR_CUTOFF]
    _, _, w = oracle.lj_cells(b.positions, b.species, b.sys_off, b.tril, cut, sig, eps)
    h = 1e-6
    es = []
    for sgn in (1, -1):
        m = np.diag([1 + sgn * h, 1, 1])
        es.append(oracle.lj_cells(b.positions @ m, b.species, b.sys_off,
                                  (b.tril * [1 + sgn * h, 1, 1, 1, 1, 1]), cut, sig, eps)[0][0])
    assert w[0, 0, 0] == pytest.approx(-(es[0] - es[1]) / (2 * h), rel=1e-5)


def test_lj_switch_polynomial_properties():
    # S = 1 and S' = 0 at r_tr, S = S' = 0 at rc, dS/d(r^2) = FD of S, S decreasing between
    rc, rtr = 8.5, 6.0
    S, dS = oracle.lj_switch_factor(np.array([rtr**2, rc**2]), rc, rtr)
    assert S[0] == 1.0 and S[1] == 0.0
    r2 = np.linspace(rtr**2 + 1e-3, rc**2 - 1e-3, 57)
    S, dS = oracle.lj_switch_factor(r2, rc, rtr)
    h = 1e-6
    fd = (oracle.lj_switch_factor(r2 + h, rc, rtr)[0] - oracle.lj_switch_factor(r2 - h, rc, rtr)[0]) / (2 * h)
    np.testing.assert_allclose(dS, fd, rtol=1e-6, atol=1e-9)
    assert np.all(np.diff(S) < 0) and abs(dS[0]) < 1e-3 and abs(dS[-1]) < 1e-3
    assert oracle.lj_switch_factor(rtr**2 + 1e-12, rc, rtr)[0] == pytest.approx(1.0, abs=1e-12)


def test_lj_switched_reference_regimes():
    # test_lennard_jones.py:316-387: particles A, B one apart, a single directed edge
    sig = np.array([[1.0, 1.2], [1.2, 1.0]])
    eps = np.array([[1.0, 0.8], [0.8, 1.0]])
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    args = (pos, [0, 1], [0, 2], [10, 0, 10, 0, 0, 10])
    ij, sh = np.array([[0, 1]]), np.zeros((1, 3))
    target = oracle.lj_edges(*args, [12.0], sig, eps, ij, sh)[0][0]
    def sw(rc, rtr):
        return oracle.lj_switched_edges(*args, [rc], [rtr], sig, eps, ij, sh)[0][0]
    assert sw(2.0, 1.5) == pytest.approx(target, rel=1e-15)   # r < r_tr
    assert sw(1.2, 1.0) == pytest.approx(target, rel=1e-15)   # r = r_tr
    assert 0.0 < sw(1.2, 0.8) < target                         # switched
    assert sw(1.0, 0.8) == 0.0                                 # r = rc


def test_lj_switched_forces_and_virial_are_derivatives():
    b = workloads.batched_lj_boxes(sizes=[90, 300], seed=11)
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = np.full(2, workloads.AR_CUTOFF)
    rtr = np.array([0.7, 0.85]) * workloads.AR_CUTOFF
    def run(pos, tril):
        ij, sh = oracle.neighborlist(pos, np.arange(b.n), b.sys_off, tril, cut)
        return oracle.lj_switched_edges(pos, b.species, b.sys_off, tril, cut, rtr, sig, eps, ij, sh)
    e, f, w = run(b.positions, b.tril)
    np.testing.assert_allclose(f[:90].sum(0), 0, atol=1e-10)
    h = 1e-5
    for atom in [3, 100, 250]:
        for k in range(3):
            p = b.positions.copy(); p[atom, k] += h
            ep = run(p, b.tril)[0].sum()
            p[atom, k] -= 2 * h
            em = run(p, b.tril)[0].sum()
            assert -(ep - em) / (2 * h) == pytest.approx(f[atom, k], rel=2e-5, abs=1e-8)
    hs = 1e-6
    es = []
    for sgn in (1, -1):
        m = np.diag([1 + sgn * hs, 1, 1])
        es.append(run(b.positions @ m, b.tril * [1 + sgn * hs, 1, 1, 1, 1, 1])[0][1])
    assert w[1, 0, 0] == pytest.approx(-(es[0] - es[1]) / (2 * hs), rel=1e-5)
    # no switch when r_tr >= rc
    ij, sh = oracle.neighborlist(b.positions, np.arange(b.n), b.sys_off, b.tril, cut)
    e0 = oracle.lj_edges(b.positions, b.species, b.sys_off, b.tril, cut, sig, eps, ij, sh)[0]
    e1 = oracle.lj_switched_edges(b.positions, b.species, b.sys_off, b.tril, cut, cut, sig, eps, ij, sh)[0]
    np.testing.assert_allclose(e1, e0, rtol=1e-12)
    assert np.all(e > e0)  # the switch removes attractive tail energy


# ---- Morse ----

def test_morse_closed_form():
    # morse.py:142-168 / test_morse.py: D (1 - exp(-a (r - r0)))^2, zero at r0,
    # D at dissociation
    pos = np.array([[0.0, 0, 0], [1.3, 0, 0]])
    tril = [30, 0, 30, 0, 0, 30]
    ij = np.array([[0, 1]]); sh = np.zeros((1, 3), int)
    for r, exp_e in [(1.3, None)]:
        e, f, w = oracle.morse_edges(pos, [0, 0], [0, 2], tril, [[4.5]], [[1.9]], [[1.1]], ij, sh)
        assert e[0] == pytest.approx(4.5 * (1 - math.exp(-1.9 * 0.2)) ** 2, rel=1e-12)
        dudr = 2 * 4.5 * (1 - math.exp(-1.9 * 0.2)) * 1.9 * math.exp(-1.9 * 0.2)
        assert f[1, 0] == pytest.approx(-dudr, rel=1e-12) and f[0, 0] == pytest.approx(dudr)
    e0 = oracle.morse_edges(np.array([[0, 0, 0], [1.1, 0, 0]]), [0, 0], [0, 2], tril,
                            [[4.5]], [[1.9]], [[1.1]], ij, sh)[0][0]
    assert e0 == 0.0
    ef = oracle.morse_edges(np.array([[0, 0, 0], [14.0, 0, 0]]), [0, 0], [0, 2], tril,
                            [[4.5]], [[1.9]], [[1.1]], ij, sh)[0][0]
    assert ef == pytest.approx(4.5, rel=1e-9)


# ---- Ewald ----

def nacl_supercell(rep=5):
    base = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1],
                     [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    q = np.array([-1, -1, -1, -1, 1, 1, 1, 1], float)
    g = np.stack(np.meshgrid(*[np.arange(rep)] * 3, indexing="ij"), -1).reshape(-1, 1, 3) * 2.0
    pos = (g + base[None]).reshape(-1, 3)
    return pos, np.tile(q, rep**3), 2.0 * rep


def ewald_total(pos, q, L, alpha, rc, nmax):
    tril = [L, 0, L, 0, 0, L]
    off = [0, len(pos)]
    ij, sh = oracle.neighborlist(pos, np.arange(len(pos)), off, tril, [rc])
    esr, fsr, wsr = oracle.ewald_real_edges(pos, q, off, tril, [alpha], [rc], ij, sh)
    kv = oracle.kvecs(tril, nmax)[None]
    elr, flr, wlr, _ = oracle.ewald_recip(pos, q, off, tril, [alpha], kv)
    es = oracle.ewald_self(q, off, [alpha])
    return esr[0] + elr[0] + es[0], fsr + flr, wsr[0] + wlr[0]


def test_ewald_madelung_nacl():
    # test_ewald.py:272-333: E / K = M * rep^3 * 4 with M = -1.7475646
    pos, q, L = nacl_supercell(5)
    e, f, _ = ewald_total(pos, q, L, alpha=0.7, rc=4.9, nmax=8)
    assert e / oracle.coulomb_k() == pytest.approx(-1.7475646 * 125 * 4, rel=5e-5)
    assert np.abs(f).max() < 1e-4  # perfect lattice: no net force


def test_ewald_alpha_independence_and_forces():
    rng = np.random.default_rng(11)
    L = 12.0
    pos = rng.uniform(0, L, (24, 3))
    q = np.tile([1.0, -1.0], 12)
    e1, f1, w1 = ewald_total(pos, q, L, 0.62, 5.9, 14)
    e2, f2, w2 = ewald_total(pos, q, L, 0.72, 5.9, 16)
    assert e1 == pytest.approx(e2, rel=1e-6)
    np.testing.assert_allclose(f1, f2, atol=2e-4 * np.abs(f1).max())
    np.testing.assert_allclose(w1, w2, atol=1e-5 * np.abs(w1).max())
    h = 1e-6
    for atom, k in [(0, 0), (5, 1), (17, 2)]:
        p = pos.copy(); p[atom, k] += h
        ep = ewald_total(p, q, L, 0.62, 5.9, 14)[0]
        p[atom, k] -= 2 * h
        em = ewald_total(p, q, L, 0.62, 5.9, 14)[0]
        assert -(ep - em) / (2 * h) == pytest.approx(f1[atom, k], rel=1e-5, abs=1e-6)


def test_coulomb_vacuum_identities():
    # test_ewald.py:137-205: bonded pairs at distance 1 and 2 -> -2K, -0.5K
    K = oracle.coulomb_k()
    assert K == pytest.approx(14.399645, rel=1e-7)
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [10, 0, 0], [11, 0, 0]])
    q = np.array([1.0, -1, 1, -1])
    ij = np.array([[0, 1], [1, 0], [2, 3], [3, 2]]); sh = np.zeros((4, 3), int)
    tril = [100, 0, 100, 0, 0, 100]
    # mode 1 with alpha -> 0 and a zero screened part: screened - vacuum
    e_corr = oracle.ewald_real_edges(pos, q, [0, 4], tril, [1e-9], [50.0], ij, sh, True)[0][0]
    e_scr = oracle.ewald_real_edges(pos, q, [0, 4], tril, [1e-9], [50.0], ij, sh)[0][0]
    assert e_scr - e_corr == pytest.approx(-2.0 * K, rel=1e-9)
    pos2 = np.array([[1.0, 0, 0], [19.0, 0, 0]])
    e2 = oracle.ewald_real_edges(pos2, [1.0, -1.0], [0, 2], [20, 0, 20, 0, 0, 20], [1e-12], [50.0],
                                 np.array([[0, 1], [1, 0]]), np.array([[-1, 0, 0], [1, 0, 0]]))[0][0]
    assert e2 == pytest.approx(-0.5 * K, rel=1e-9)


def test_kvec_grid_counts():
    tril = [56.4, 0, 56.4, 0, 0, 56.4]
    assert len(oracle.kvecs(tril, 8)) == 9 * 17 * 17
    kv = oracle.kvecs(tril, 3, kmax=2 * math.pi / 56.4 * 1.01)
    # n_x in [0..n], n_y, n_z in [-n..n]: 0, (1,0,0), (0,+-1,0), (0,0,+-1)
    assert {tuple(v) for v in kv} == {(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                                      (0, 0, -1)}


# ---- integrators ----

def _md_setup(sizes=(100,), seed=0):
    b = workloads.batched_lj_boxes(sizes=list(sizes), seed=seed)
    rng = np.random.default_rng(seed)
    mom = rng.normal(0, 1, (b.n, 3)) * 0.5
    return b, mom


def test_vv_single_step_formula():
    # test_integrators.py:292-327: p' = p + F dt/2, r' = wrap(r + p' dt / m), ...
    b, mom = _md_setup()
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = [workloads.AR_CUTOFF]
    pos = b.positions.copy(); p = mom.copy()
    _, f0, _ = oracle.lj_cells(pos, b.species, b.sys_off, b.tril, cut, sig, eps)
    f = f0.copy()
    dt = 0.2
    oracle.md_step(0, pos, p, b.masses, b.species, b.sys_off, b.tril, cut, sig, eps, [dt], [0],
                   [0], np.zeros((b.n, 3)), f)
    ph = mom + f0 * dt / 2
    r1 = workloads.wrap_cubic(b.positions + ph / b.masses[:, None] * dt, b.tril[0, 0])
    np.testing.assert_allclose(pos, r1, atol=1e-12)
    _, f1, _ = oracle.lj_cells(r1, b.species, b.sys_off, b.tril, cut, sig, eps)
    np.testing.assert_allclose(p, ph + f1 * dt / 2, atol=1e-12)


def test_vv_energy_conservation_and_reversibility():
    b, _ = _md_setup((256,), seed=3)
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    cut = [workloads.AR_CUTOFF]
    rng = np.random.default_rng(1)
    p = rng.normal(0, np.sqrt(b.masses[:, None] * 94.4 * 8.617e-5), (b.n, 3))
    pos = b.positions.copy()
    _, f, _ = oracle.lj_cells(pos, b.species, b.sys_off, b.tril, cut, sig, eps)
    dt = 0.02
    eta = np.zeros((b.n, 3))
    tot = []
    for _ in range(200):
        e, _ = oracle.md_step(0, pos, p, b.masses, b.species, b.sys_off, b.tril, cut, sig, eps,
                              [dt], [0], [0], eta, f)
        tot.append(e[0] + (0.5 * p**2 / b.masses[:, None]).sum())
    tot = np.array(tot)
    assert np.abs(tot - tot[0]).max() < 1e-3 * abs(tot[0])
    # reverse momenta, run back, reverse: initial state
    p = -p
    for _ in range(200):
        oracle.md_step(0, pos, p, b.masses, b.species, b.sys_off, b.tril, cut, sig, eps,
                       [dt], [0], [0], eta, f)
    d = workloads.wrap_cubic(pos - b.positions, b.tril[0, 0])
    assert np.abs(d).max() < 1e-6


def test_baoab_o_step_formula():
    # md/integrators.py:330-402: p <- c1 p + sqrt(kT (1 - c1^2)) sqrt(m) eta;
    # with F = 0 (no neighbors) the step is A(dt/2) O A(dt/2) on free particles
    pos = np.array([[0.0, 0, 0], [50.0, 0, 0]])
    p = np.array([[1.0, 2, 3], [-1, 0.5, 0.2]])
    m = np.array([10.0, 20.0])
    eta = np.array([[0.3, -0.2, 1.1], [0.5, 0.1, -0.7]])
    f = np.zeros((2, 3))
    tril = [200, 0, 200, 0, 0, 200]
    dt, kT, g = 0.5, 0.025, 0.1
    pos0 = pos.copy()
    oracle.md_step(1, pos, p, m, [0, 0], [0, 2], tril, [5.0], [[3.0]], [[0.01]], [dt], [kT], [g],
                   eta, f)
    c1 = math.exp(-g * dt)
    pe = c1 * np.array([[1.0, 2, 3], [-1, 0.5, 0.2]]) + math.sqrt(kT * (1 - c1**2)) * np.sqrt(m)[:, None] * eta
    np.testing.assert_allclose(p, pe, rtol=1e-12)
    x = pos0 + np.array([[1.0, 2, 3], [-1, 0.5, 0.2]]) / m[:, None] * dt / 2 + pe / m[:, None] * dt / 2
    np.testing.assert_allclose(pos, x, rtol=1e-12, atol=1e-12)


def test_morse_parameter_constructors():
    # test_morse.py:126-133 (from_harmonic) and UFF equations 2 and 4
    from kups_amd.potential.morse import MorseBondParameters, uff_bond_length

    r0 = np.array([[1.0, 1.2], [1.2, 1.5]])
    k = np.array([[100.0, 80.0], [80.0, 120.0]])
    D = np.array([[2.0, 1.5], [1.5, 2.5]])
    p = MorseBondParameters.from_harmonic(("A", "B"), r0, k, D)
    np.testing.assert_allclose(p.alpha, np.sqrt(k / D))
    # equal electronegativity: no correction, r_ij = r_i + r_j
    r = uff_bond_length([0.7, 0.9], [2.5, 2.5])
    np.testing.assert_allclose(r, [[1.4, 1.6], [1.6, 1.8]])
    # C-O (UFF radii 0.757, 0.658; GMP chi 5.343, 8.741): r_EN by hand
    r = uff_bond_length([0.757, 0.658], [5.343, 8.741])
    ren = 0.757 * 0.658 * (np.sqrt(5.343) - np.sqrt(8.741)) ** 2 / (5.343 * 0.757 + 8.741 * 0.658)
    assert r[0, 1] == pytest.approx(0.757 + 0.658 - ren, rel=1e-14)
    assert r[0, 1] == r[1, 0]
    # near equilibrium Morse ~ harmonic with k = 2 D alpha^2 (U = k/2 x^2)
    pu = MorseBondParameters.from_uff(("C", "O"), [0.757, 0.658], [5.343, 8.741], [1.912, 2.3],
                                      np.full((2, 2), 3.0))
    kk = 664.12 * 1.912 * 2.3 / pu.r0[0, 1] ** 3
    assert 2 * pu.D[0, 1] * pu.alpha[0, 1] ** 2 == pytest.approx(kk, rel=1e-12)


def test_pressure_observables():
    """Ideal gas: P = N kT / V from K = 3/2 N kT with zero stress; LJ: the
    configurational pressure tr(W / V) / 3 equals -dU/dV under isotropic
    scaling of cell and positions (central differences of the oracle energy)."""
    import torch

    from kups_amd.core.constants import BOLTZMANN_CONSTANT
    from kups_amd.md.observables import instantaneous_pressure, stress_via_virial, volume

    kT = 300.0 * BOLTZMANN_CONSTANT
    tril = torch.tensor([[20.0, 0.0, 21.0, 0.0, 0.0, 22.0]], dtype=torch.float64)
    V = volume(tril)
    P = instantaneous_pressure(torch.tensor([1.5 * 100 * kT], dtype=torch.float64),
                               torch.zeros((1, 3, 3), dtype=torch.float64), V)
    assert P.item() == pytest.approx(100 * kT / (20 * 21 * 22), rel=1e-14)

    b = workloads.argon_fcc(4)
    pos = b.positions.astype(np.float64)
    rc = [workloads.AR_CUTOFF]
    sig, eps = np.array([[workloads.AR_SIGMA]]), np.array([[workloads.AR_EPS]])
    e, f, w = oracle.lj_cells(pos, np.zeros(len(pos), np.int32), b.sys_off, b.tril, rc, sig, eps)
    V0 = b.tril[0, 0] * b.tril[0, 2] * b.tril[0, 5]
    p_conf = np.trace(stress_via_virial(torch.as_tensor(w.reshape(1, 9)),
                                        torch.as_tensor(b.tril)).numpy()[0]) / 3
    h = 1e-5
    es = []
    for s in (1 + h, 1 - h):
        es.append(oracle.lj_cells(pos * s, np.zeros(len(pos), np.int32), b.sys_off, b.tril * s, rc,
                                  sig, eps)[0][0])
    dUdV = (es[0] - es[1]) / (V0 * ((1 + h) ** 3 - (1 - h) ** 3))
    assert p_conf == pytest.approx(-dUdV, rel=1e-6)
