"""GPU parity of the pair sets the force kernels actually evaluate.

The fused MD kernel (LDS Verlet lists or per-step cell lists), the Verlet
pair path of the LJ + Ewald MD step and the cell-list pair path each record
the directed pairs (i, j, shift) their force loops kept (PairSetRecorder,
the recording instantiation of the same kernels).  Bar: bit-exact set
equality with the oracle's neighbor list (strict r < rc in FP64 on the same
FP32 positions, core/neighborlist.py:542-564, 589; minimum-image-only
exclusion, 731-735): same pairs, same shifts, each directed pair once.  Cases:
lists reused after tens of steps of motion, per-step cell lists, image-grid
systems, triclinic cells, cells rescaled by the barostat, exclusion ids, and
pairs placed inside the FP32 band around the cutoff where the kernels
re-decide in FP64.  At the full batched sizes the recorded sets are compared
with the cell-list export (itself bit-exact against the oracle below) and
with the oracle on a sample of systems.
"""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import _lib, workloads
from kups_amd.core.batch import ParticleBatch, SystemBatch
from kups_amd.core.neighborlist import CellListNeighborList, DenseNearestNeighborList
from kups_amd.core.pairset import PairSetRecorder
from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
from kups_amd.md.integrators import (BAOAB_LANGEVIN, FusedLJPropagator, EwaldLJPropagator,
                                     make_baoab_langevin_step, make_csvr_npt_step,
                                     make_velocity_verlet_step)
from kups_amd.potential.ewald import EwaldParameters
from kups_amd.potential.lennard_jones import LennardJonesParameters

pytestmark = pytest.mark.gpu

AR = LennardJonesParameters.from_dict(workloads.AR_CUTOFF,
                                      {"Ar": (workloads.AR_SIGMA, workloads.AR_EPS)})
NACL_LJ = LennardJonesParameters.from_dict(12.0, {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)})


def _f64(t):
    return t.detach().cpu().numpy().astype(np.float64)


def keys(ij, sh, n):
    """Canonical sorted int64 keys of directed edges (i, j, s), |s| <= 7."""
    ij = np.asarray(ij, np.int64)
    s = np.asarray(sh, np.int64) + 8
    assert s.size == 0 or (s.min() >= 0 and s.max() < 16)
    k = ((ij[:, 0] * n + ij[:, 1]) << 12) | (s[:, 0] << 8) | (s[:, 1] << 4) | s[:, 2]
    return np.sort(k)


def _decode(k, n):
    s = np.stack([(k >> 8) & 15, (k >> 4) & 15, k & 15], 1) - 8
    ij = k >> 12
    return np.stack([ij // n, ij % n], 1), s


def assert_same_set(got_ij, got_sh, ref_ij, ref_sh, n):
    kg, kr = keys(got_ij, got_sh, n), keys(ref_ij, ref_sh, n)
    assert len(np.unique(kg)) == len(kg), "a directed pair was evaluated twice"
    if len(kg) != len(kr) or not np.array_equal(kg, kr):
        miss, extra = np.setdiff1d(kr, kg), np.setdiff1d(kg, kr)
        mi, ms = _decode(miss[:8], n)
        ei, es = _decode(extra[:8], n)
        raise AssertionError(
            f"{len(kg)} pairs evaluated, the oracle has {len(kr)}: {len(miss)} missing, e.g. "
            f"{list(zip(mi.tolist(), ms.tolist()))}; {len(extra)} extra, e.g. "
            f"{list(zip(ei.tolist(), es.tolist()))}")


def oracle_set(pos, excl, sys_off, tril, cutoff):
    cut = np.broadcast_to(np.asarray(cutoff, np.float64), (len(sys_off) - 1,))
    return oracle.neighborlist(pos, excl, sys_off, tril, cut)


def _md_state(b, temperature=94.4, dt_fs=2.0, seed=0):
    p = maxwell_boltzmann_momenta(b.masses, b.sizes, temperature, np.random.default_rng(seed))
    return make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril,
                         workloads.AR_CUTOFF, time_step_fs=dt_fs, temperature=temperature,
                         friction_per_fs=0.001, momenta=p.astype(np.float32))


def _record_md(prop, st, n_steps, capacity):
    rec = PairSetRecorder(capacity).arm()
    prop(st, n_steps)
    return rec.result()


def _check_md(st, ij, sh, terms):
    pos = _f64(st.positions)
    tril = _f64(st.tril)
    off = st.sys_off.cpu().numpy()
    assert (terms == 1).all()
    ri, rs = oracle_set(pos, np.arange(st.n), off, tril, workloads.AR_CUTOFF)
    assert_same_set(ij, sh, ri, rs, st.n)
    return len(ri)


# ------------------------------------------------------------- fused MD --

def test_md_lj_verlet_list_pairset_after_motion(gpu):
    """VV with Verlet lists (skin 0.3 sigma) after 40 steps; the 64/90-atom
    boxes are narrower than 2 rc and take the image-grid cell list."""
    b = workloads.batched_lj_boxes(sizes=[64, 90, 333, 1000, 2100, 4096], seed=31)
    st = _md_state(b)
    prop = make_velocity_verlet_step(AR)
    prop.evaluate_forces(st)
    prop(st, 38)
    # the record covers the last step of a launch (lists are used from the
    # second step of a launch on)
    ij, sh, terms = _record_md(prop, st, 2, 8_000_000)
    assert _check_md(st, ij, sh, terms) > 300_000


def test_md_lj_cell_list_pairset_baoab(gpu):
    b = workloads.batched_lj_boxes(sizes=[77, 500, 1500, 3000], seed=32)
    st = _md_state(b)
    prop = make_baoab_langevin_step(AR, seed=5, skin=0.0)
    prop.evaluate_forces(st)
    ij, sh, terms = _record_md(prop, st, 12, 8_000_000)
    _check_md(st, ij, sh, terms)


def test_md_lj_force_evaluation_pairset(gpu):
    """kupsmi_lj_forces (zero steps): positions inside the cell are used as given."""
    b = workloads.batched_lj_boxes(sizes=[70, 800, 2500], seed=33)
    st = _md_state(b)
    prop = make_velocity_verlet_step(AR)
    rec = PairSetRecorder(4_000_000).arm()
    prop.evaluate_forces(st)
    _check_md(st, *rec.result())


def test_md_lj_npt_pairset_rescaled_cells(gpu):
    b = workloads.batched_lj_boxes(sizes=[300, 1200], seed=34)
    st = _md_state(b, temperature=120.0)
    prop = make_csvr_npt_step(AR, seed=3)
    prop.evaluate_forces(st)
    tril0 = _f64(st.tril)
    ij, sh, terms = _record_md(prop, st, 6, 4_000_000)
    assert not np.array_equal(_f64(st.tril), tril0)  # the cell moved
    _check_md(st, ij, sh, terms)


def band_positions(lat, n_anchor, rc, deltas, seed):
    """Anchors spread over the cell plus, for each delta, a partner at
    r^2 = rc^2 (1 + delta) from an anchor (all wrapped into the cell, FP32)."""
    rng = np.random.default_rng(seed)
    lat = np.asarray(lat, float)
    f = rng.uniform(-0.45, 0.45, (n_anchor, 3))
    anchors = f @ lat
    parts = []
    for k, d in enumerate(deltas):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        p = anchors[k % n_anchor] + np.sqrt(rc * rc * (1.0 + d)) * u
        fp = p @ np.linalg.inv(lat)
        fp = (fp + 0.5) % 1.0 - 0.5
        parts.append(fp @ lat)
    return np.concatenate([anchors, np.array(parts)]).astype(np.float32)


DELTAS = [s * d for d in (2e-7, 6e-7, 1.5e-6, 2.5e-6, 3.5e-6, 3.9e-6, 4.5e-6, 8e-6)
          for s in (-1.0, 1.0)] * 3


def _band_count(pos, off, tril, rc, rel=4e-6):
    """Oracle pairs and near misses within rel of rc^2 (in FP64)."""
    ri, rs = oracle_set(pos, np.arange(len(pos)), off, tril, rc * (1 + rel))
    lat = np.zeros((3, 3))
    lat[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]] = tril[0]
    d = pos[ri[:, 1]] - pos[ri[:, 0]] + rs @ lat
    r2 = (d * d).sum(1) / (rc * rc)
    return int(np.sum(np.abs(r2 - 1.0) < rel))


@pytest.mark.parametrize("tri", [False, True])
@pytest.mark.parametrize("mode", ["list", "cells", "forces"])
def test_md_lj_cutoff_band_pairs(gpu, mode, tri):
    """Pairs placed inside the FP32 band around rc (re-decided in FP64) and
    just outside it; nothing moves (dt = 0), so the lists built at step 0 are
    the ones used at step 1."""
    lat = np.array([[31.0, 0, 0], [4.0, 30.0, 0], [-3.0, 5.0, 29.0]]) if tri else np.eye(3) * 30.0
    rc = workloads.AR_CUTOFF
    pos = band_positions(lat, 150, rc, DELTAS, seed=7 + tri)
    n = len(pos)
    tril = lat[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]][None]
    st = make_md_state(pos, np.full(n, workloads.AR_MASS), np.zeros(n, np.int32), [n], tril, rc,
                       time_step_fs=0.0, temperature=0.0)
    assert _band_count(pos.astype(np.float64), [0, n], tril, rc) >= 20
    rec = PairSetRecorder(1_000_000).arm()
    if mode == "forces":
        make_velocity_verlet_step(AR).evaluate_forces(st)
    else:
        make_velocity_verlet_step(AR, skin=0.0 if mode == "cells" else None)(st, 2)
    ij, sh, terms = rec.result()
    np.testing.assert_array_equal(_f64(st.positions), pos.astype(np.float64))
    ri, rs = oracle_set(pos.astype(np.float64), np.arange(n), [0, n], tril, rc)
    assert_same_set(ij, sh, ri, rs, n)


# ---------------------------------------------------------- Verlet path --

def _melt_state(ms, seed, temperature=1200.0):
    b = workloads.nacl_melts(ms=ms, seed=seed)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, temperature, np.random.default_rng(seed))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=temperature, friction_per_fs=0.001,
                       momenta=p0.astype(np.float32), charges=b.charges)
    return b, st


def test_lj_ewald_bench_path_pairset_after_motion(gpu):
    """The bench's LJ + Ewald step (BAOAB, Verlet skin 1.5 A): after 25 steps
    the pair sets of both terms equal the oracle's; the 64-ion melt is
    narrower than 2 (rc + skin) and its list holds periodic images."""
    b, st = _melt_state([2, 5, 6], seed=41)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    prop = EwaldLJPropagator(NACL_LJ, ep, BAOAB_LANGEVIN, seed=2, skin=1.5)
    prop.evaluate_forces(st)
    prop(st, 25)
    rebuilds, overflow = prop.verlet.counters()
    assert overflow == 0
    rec = PairSetRecorder(20_000_000).arm(prop.verlet)
    prop.evaluate_forces(st)
    ij, sh, terms = rec.result()
    pos, off, tril = _f64(st.positions), st.sys_off.cpu().numpy(), _f64(st.tril)
    for bit, rc in ((1, 12.0), (2, 12.0)):
        m = (terms & bit) != 0
        ri, rs = oracle_set(pos, np.arange(st.n), off, tril, rc)
        assert_same_set(ij[m], sh[m], ri, rs, st.n)


def _verlet_call(v, st, ep, lj_excl=None, coul_excl=None, rec=None):
    dev = st.positions.device
    keep = []

    def t(a, dt):
        keep.append(torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev))
        return _lib.ptr(keep[-1])

    f = torch.empty((st.n, 3), dtype=torch.float32, device=dev)
    e = torch.empty((st.n_sys, 2), dtype=torch.float64, device=dev)
    if rec is not None:
        rec.arm(v)
    _lib.check(_lib.load().kupsmi_pair_forces_verlet(
        v.handle, _lib.stream_ptr(), st.n_sys, _lib.ptr(st.sys_off), st.n, int(st.sizes.max()),
        _lib.ptr(st.positions), _lib.ptr(st.charges), _lib.ptr(st.species), NACL_LJ.n_species,
        t(NACL_LJ.sigma, torch.float64), t(NACL_LJ.epsilon, torch.float64), _lib.ptr(st.cutoff),
        None if lj_excl is None else t(lj_excl, torch.int32), t(ep.alpha, torch.float64),
        t(ep.cutoff, torch.float64), None if coul_excl is None else t(coul_excl, torch.int32),
        _lib.ptr(st.tril), v.skin, v.max_neighbors, 0, _lib.ptr(f), _lib.ptr(e), None))
    torch.cuda.synchronize()
    return f, e


@pytest.mark.parametrize("tri", [False, True])
def test_verlet_path_cutoff_band_pairs(gpu, tri):
    from kups_amd.core.verlet import VerletList

    lat = np.array([[36.0, 0, 0], [5.0, 35.0, 0], [-4.0, 3.0, 34.0]]) if tri else np.eye(3) * 36.0
    rc = 12.0
    pos = band_positions(lat, 220, rc, DELTAS, seed=17 + tri)
    n = len(pos)
    tril = lat[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]][None]
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    st = make_md_state(pos, np.ones(n), (np.arange(n) % 2).astype(np.int32), [n], tril, rc,
                       time_step_fs=1.0, temperature=0.0, charges=q)
    assert _band_count(pos.astype(np.float64), [0, n], tril, rc) >= 20
    ep = EwaldParameters.box(1, alpha=0.3, cutoff=rc, nmax=4)
    v = VerletList(1.0, 600)
    _verlet_call(v, st, ep)  # builds the list
    rec = PairSetRecorder(2_000_000)
    _verlet_call(v, st, ep, rec=rec)  # uses it
    ij, sh, terms = rec.result()
    ri, rs = oracle_set(pos.astype(np.float64), np.arange(n), [0, n], tril, rc)
    for bit in (1, 2):
        m = (terms & bit) != 0
        assert_same_set(ij[m], sh[m], ri, rs, n)


def test_pair_paths_pairset_with_exclusion_ids(gpu):
    """GCMC exclusion semantics on the cell-list path and the Verlet path:
    LJ dropped within a group id (host atoms share one), Coulomb by particle;
    an image-grid system checks exclusion on the minimum image only."""
    from kups_amd.core.verlet import VerletList

    rng = np.random.default_rng(5)
    sizes = [300, 40]
    Ls = [25.0, 14.0]
    pos = np.concatenate([rng.uniform(-L / 2, L / 2, (n, 3)) for n, L in zip(sizes, Ls)])
    pos = pos.astype(np.float32)
    n = len(pos)
    grp = np.where(np.arange(n) % 100 < 60, 10**6, np.arange(n) // 3).astype(np.int32)
    q = rng.normal(0, 0.3, n)
    tril = np.concatenate([workloads.cubic_tril(L) for L in Ls])
    off = np.array([0, sizes[0], n])
    st = make_md_state(pos, np.ones(n), (np.arange(n) % 2).astype(np.int32), sizes, tril, 11.0,
                       time_step_fs=1.0, temperature=0.0, charges=q)
    ep = EwaldParameters.box(2, alpha=0.26, cutoff=12.0, nmax=4)
    p64 = pos.astype(np.float64)
    ri_lj, rs_lj = oracle_set(p64, grp, off, tril, 11.0)
    ri_cl, rs_cl = oracle_set(p64, np.arange(n), off, tril, 12.0)
    # cell-list path
    dev = st.positions.device
    keep = []

    def t(a, dt):
        keep.append(torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev))
        return _lib.ptr(keep[-1])

    f = torch.empty((n, 3), dtype=torch.float32, device=dev)
    e = torch.empty((2, 2), dtype=torch.float64, device=dev)
    rec = PairSetRecorder(2_000_000).arm()
    _lib.check(_lib.load().kupsmi_pair_forces(
        _lib.ctx(), _lib.stream_ptr(), 2, _lib.ptr(st.sys_off), n, max(sizes),
        _lib.ptr(st.positions), _lib.ptr(st.charges), _lib.ptr(st.species), 2,
        t(NACL_LJ.sigma, torch.float64), t(NACL_LJ.epsilon, torch.float64),
        t([11.0, 11.0], torch.float64), t(grp, torch.int32), t(ep.alpha, torch.float64),
        t(ep.cutoff, torch.float64), None, _lib.ptr(st.tril), 0, _lib.ptr(f), _lib.ptr(e), None,
        None))
    ij, sh, terms = rec.result()
    assert_same_set(ij[(terms & 1) != 0], sh[(terms & 1) != 0], ri_lj, rs_lj, n)
    assert_same_set(ij[(terms & 2) != 0], sh[(terms & 2) != 0], ri_cl, rs_cl, n)
    # Verlet path (second call uses the list)
    st.cutoff.fill_(11.0)
    v = VerletList(1.0, 800)
    _verlet_call(v, st, ep, lj_excl=grp)
    rec = PairSetRecorder(2_000_000)
    _verlet_call(v, st, ep, lj_excl=grp, rec=rec)
    ij, sh, terms = rec.result()
    assert_same_set(ij[(terms & 1) != 0], sh[(terms & 1) != 0], ri_lj, rs_lj, n)
    assert_same_set(ij[(terms & 2) != 0], sh[(terms & 2) != 0], ri_cl, rs_cl, n)


# ------------------------------------------------------ cell-list export --

def _batch(pos, sizes, tril, excl=None, dev="cuda"):
    system = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    pb = ParticleBatch(torch.as_tensor(np.asarray(pos, np.float32), device=dev),
                       torch.as_tensor(system, device=dev),
                       exclusion=None if excl is None else
                       torch.as_tensor(np.asarray(excl, np.int32), device=dev))
    sb = SystemBatch(torch.as_tensor(np.asarray(tril, np.float64).reshape(-1, 6), device=dev))
    return pb, sb


def test_cell_list_export_equals_oracle_set(gpu):
    b = workloads.batched_lj_boxes(sizes=[64, 90, 333, 1000, 2100, 4096, 9000], seed=35)
    pb, sb = _batch(b.positions, b.sizes, b.tril)
    cut = torch.full((b.n_sys,), workloads.AR_CUTOFF, dtype=torch.float64, device=gpu)
    e = CellListNeighborList(avg_edges=1)(pb, None, sb, cut)  # avg_edges=1 forces a regrow
    ri, rs = oracle_set(b.positions.astype(np.float32).astype(np.float64), np.arange(b.n),
                        b.sys_off, b.tril, workloads.AR_CUTOFF)
    assert_same_set(e.indices.cpu().numpy(), e.shifts[:, 0].cpu().numpy(), ri, rs, b.n)
    # i-major output
    assert np.all(np.diff(e.indices[:, 0].cpu().numpy()) >= 0)


def test_cell_list_export_triclinic_and_exclusions(gpu):
    rng = np.random.default_rng(8)
    lat = np.array([[19.0, 0, 0], [3.0, 18.5, 0], [2.5, -2.0, 19.5]])
    pos = rng.uniform(-0.5, 0.5, (700, 3)) @ lat
    tril = lat[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]][None]
    excl = np.arange(700) // 3
    for rc in (4.2, 6.0, 9.9):
        pb, sb = _batch(pos, [700], tril, excl)
        e = CellListNeighborList()(pb, None, sb, torch.tensor([rc], dtype=torch.float64,
                                                              device=gpu))
        ri, rs = oracle_set(pos.astype(np.float32).astype(np.float64), excl, [0, 700], tril, rc)
        assert_same_set(e.indices.cpu().numpy(), e.shifts[:, 0].cpu().numpy(), ri, rs, 700)


def test_dense_export_bit_exact_order(gpu):
    b = workloads.batched_lj_boxes(sizes=[64, 500], seed=36)
    pb, sb = _batch(b.positions, b.sizes, b.tril)
    cut = torch.full((b.n_sys,), workloads.AR_CUTOFF, dtype=torch.float64, device=gpu)
    e = DenseNearestNeighborList()(pb, None, sb, cut)
    ri, rs = oracle_set(b.positions.astype(np.float32).astype(np.float64), np.arange(b.n),
                        b.sys_off, b.tril, workloads.AR_CUTOFF)
    np.testing.assert_array_equal(e.indices.cpu().numpy(), ri)
    np.testing.assert_array_equal(e.shifts[:, 0].cpu().numpy(), rs)


# ------------------------------------------------------------ rh queries --

def _query(lh_pos, sizes, tril, rh_pos, rh_sys, remap, cutoff, rh_excl=None, nl=None):
    pb, sb = _batch(lh_pos, sizes, tril)
    rpb = ParticleBatch(torch.as_tensor(np.asarray(rh_pos, np.float32), device="cuda"),
                        torch.as_tensor(np.asarray(rh_sys, np.int32), device="cuda"),
                        exclusion=None if rh_excl is None else
                        torch.as_tensor(np.asarray(rh_excl, np.int32), device="cuda"))
    cut = torch.as_tensor(np.broadcast_to(np.asarray(cutoff, np.float64), (len(sizes),)).copy(),
                          device="cuda")
    e = (nl or CellListNeighborList(avg_query_edges=1))(
        pb, rpb, sb, cut, None if remap is None else torch.as_tensor(remap, device="cuda"))
    return e.indices.cpu().numpy(), e.shifts[:, 0].cpu().numpy()


def _oracle_query(lh_pos, sizes, tril, rh_pos, rh_sys, remap, cutoff, rh_excl=None):
    off = np.concatenate([[0], np.cumsum(sizes)])
    n = off[-1]
    if rh_excl is None:
        rh_excl = remap if remap is not None else n + np.arange(len(rh_sys))
    cut = np.broadcast_to(np.asarray(cutoff, np.float64), (len(sizes),))
    return oracle.neighborlist_query(np.asarray(lh_pos, np.float32).astype(np.float64),
                                     np.arange(n), off, np.asarray(tril).reshape(-1, 6), cut,
                                     np.asarray(rh_pos, np.float32).astype(np.float64), rh_sys,
                                     rh_excl, remap)


QUERY_CASES = [
    # the reference's rh cases (test_neighborlist.py:402-676), lh sorted by system
    ([[0, 0, 0], [2, 0, 0], [5, 0, 0]], [3], [[1, 0, 0], [3, 0, 0]], [0, 0], [2, 0], [1.5]),
    ([[0, 0, 0], [5, 0, 0], [0, 0, 0], [5, 0, 0]], [2, 2], [[1, 0, 0], [1, 0, 0]], [0, 1],
     [1, 3], [1.5, 1.5]),
    ([[0, 0, 0], [5, 0, 0], [0, 0, 0], [5, 0, 0]], [2, 2], [[0.5, 0, 0], [0.5, 0, 0]], [1, 0],
     [3, 1], [1.5, 1.5]),
    ([[0, 0, 0], [2, 0, 0], [4, 0, 0]], [3], [[1, 0, 0], [3, 0, 0]], [0, 0], [0, 2], [1.5]),
    ([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]], [4],
     [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]], [0, 0, 0, 0], [0, 2, 1, 4], [2.5]),
    ([[0, 0, 0], [1, 0, 0]], [2], [[0, 0, 0], [1, 0, 0]], [0, 0], [0, 1], [1.5]),
    ([[0, 0, 0]] * 3, [2, 1], [[1, 0, 0]] * 3, [0, 1, 0], [1, 2, 0], [1.5, 1.5]),
    # no remap: rh indexes itself
    ([[0, 0, 0], [4, 0, 0]], [2], [[1, 0, 0], [9.5, 0, 0]], [0, 0], None, [1.6]),
]


@pytest.mark.parametrize("case", range(len(QUERY_CASES)))
def test_rh_query_reference_cases_bit_exact(gpu, case):
    lh, sizes, rh, rh_sys, remap, cut = QUERY_CASES[case]
    tril = np.tile([10.0, 0, 10.0, 0, 0, 10.0], (len(sizes), 1))
    remap = None if remap is None else np.asarray(remap)
    gi, gs = _query(lh, sizes, tril, rh, rh_sys, remap, cut)
    ri, rs = _oracle_query(lh, sizes, tril, rh, rh_sys, remap, cut)
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gs, rs)


def test_rh_query_moved_particles_bit_exact(gpu):
    """A patch of displaced particles (the RadiusGraphConstructor form of an
    MC move) in batched boxes, image-grid systems included, both list
    classes; plus the subset identity against the full list."""
    b = workloads.batched_lj_boxes(sizes=[8, 64, 500, 1500], seed=37)
    rng = np.random.default_rng(9)
    idx = np.sort(rng.choice(b.n, 60, replace=False))
    sys_of = np.searchsorted(b.sys_off, np.arange(b.n), side="right") - 1
    moved = b.positions[idx] + rng.normal(0, 0.7, (60, 3))
    for nl in (CellListNeighborList(), DenseNearestNeighborList()):
        for remap in (idx, None):
            gi, gs = _query(b.positions, b.sizes, b.tril, moved, sys_of[idx], remap,
                            workloads.AR_CUTOFF, nl=nl)
            ri, rs = _oracle_query(b.positions, b.sizes, b.tril, moved, sys_of[idx], remap,
                                   workloads.AR_CUTOFF)
            np.testing.assert_array_equal(gi, ri)
            np.testing.assert_array_equal(gs, rs)


# ------------------------------------------------------ full batch sizes --

def _gpu_keys(ij, sh, n):
    ij = ij.to(torch.int64)
    s = sh.to(torch.int64) + 8
    assert int(s.min()) >= 0 and int(s.max()) < 16
    k = ((ij[:, 0] * n + ij[:, 1]) << 12) | (s[:, 0] << 8) | (s[:, 1] << 4) | s[:, 2]
    return torch.sort(k).values


def _export_keys(st, cutoff):
    pb = ParticleBatch(st.positions, torch.repeat_interleave(
        torch.arange(st.n_sys, device=st.positions.device, dtype=torch.int32),
        torch.as_tensor(st.sizes, device=st.positions.device)))
    e = CellListNeighborList(avg_edges=64)(pb, None, SystemBatch(st.tril),
                                           torch.full((st.n_sys,), cutoff, dtype=torch.float64,
                                                      device=st.positions.device))
    return _gpu_keys(e.indices, e.shifts[:, 0], st.n)


def _sample_oracle(st, ij, sh, terms_mask, every, cutoff):
    """Oracle sets of every `every`-th system against the recorded pairs."""
    pos, off, tril = _f64(st.positions), st.sys_off.cpu().numpy(), _f64(st.tril)
    for s in range(0, st.n_sys, every):
        a, b = off[s], off[s + 1]
        m = terms_mask & (ij[:, 0] >= a) & (ij[:, 0] < b)
        ri, rs = oracle_set(pos[a:b], np.arange(b - a), [0, b - a], tril[s:s + 1], cutoff)
        assert_same_set(ij[m] - a, sh[m], ri, rs, b - a)


def test_md_lj_full_batch_pairset(gpu):
    """2048 boxes of 64-4096 atoms (4.36 M atoms) after 20 VV steps with lists:
    the recorded set equals the cell-list export; every 128th system against
    the oracle."""
    b = workloads.batched_lj_boxes(n_sys=2048, seed=0)
    st = _md_state(b)
    prop = make_velocity_verlet_step(AR)
    prop.evaluate_forces(st)
    prop(st, 18)
    rec = PairSetRecorder(int(b.n * 60)).arm()
    prop(st, 2)
    torch.cuda.synchronize()
    n = int(rec.count.item())
    assert 0 < n <= rec.capacity
    kg = _gpu_keys(rec.ij[:n], rec.shift[:n], st.n)
    assert bool((kg[1:] != kg[:-1]).all())
    ke = _export_keys(st, workloads.AR_CUTOFF)
    assert kg.numel() == ke.numel() and bool((kg == ke).all())
    ij, sh, terms = rec.result()
    _sample_oracle(st, ij, sh, np.ones(len(ij), bool), 128, workloads.AR_CUTOFF)


def test_lj_ewald_full_batch_pairset(gpu):
    """512 melts of 1000-8000 ions (2.1 M ions), the bench's step (BAOAB,
    skin 1.5 A) after 10 steps: both terms' recorded sets equal the cell-list
    export at 12 A; every 64th melt against the oracle."""
    b = workloads.nacl_melts(n_sys=512, seed=0)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(0))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=1200.0, friction_per_fs=0.001,
                       momenta=p0.astype(np.float32), charges=b.charges)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    prop = EwaldLJPropagator(NACL_LJ, ep, BAOAB_LANGEVIN, seed=1, skin=1.5)
    prop.evaluate_forces(st)
    prop(st, 10)
    rec = PairSetRecorder(int(b.n * 340)).arm(prop.verlet)
    prop.evaluate_forces(st)
    torch.cuda.synchronize()
    n = int(rec.count.item())
    assert 0 < n <= rec.capacity
    terms = rec.terms[:n]
    assert bool((terms == 3).all())  # equal cutoffs, no exclusion ids: both terms
    kg = _gpu_keys(rec.ij[:n], rec.shift[:n], st.n)
    assert bool((kg[1:] != kg[:-1]).all())
    ke = _export_keys(st, 12.0)
    assert kg.numel() == ke.numel() and bool((kg == ke).all())
    del kg, ke
    small = np.nonzero(b.sizes <= 1728)[0][:4]
    ij, sh, tm = rec.result()
    pos, off, tril = _f64(st.positions), st.sys_off.cpu().numpy(), _f64(st.tril)
    for s in small:
        a, c = off[s], off[s + 1]
        m = (ij[:, 0] >= a) & (ij[:, 0] < c)
        ri, rs = oracle_set(pos[a:c], np.arange(c - a), [0, c - a], tril[s:s + 1], 12.0)
        assert_same_set(ij[m] - a, sh[m], ri, rs, c - a)
