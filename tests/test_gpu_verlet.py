"""Verlet list with a skin (kupsmi_pair_forces_verlet) against the cell-list
pair path (kupsmi_pair_forces), which is itself pinned to the FP64 oracle:
same pair set and FP64 pair vectors, so energies agree to FP32 summation order
(1e-6 relative, 1e-5 of the largest force) while positions move below and
beyond skin/2, atoms cross the cell faces and are wrapped back, lists
overflow, narrow cells need image shifts, and exclusion groups apply."""

import numpy as np
import pytest
import torch

from kups_amd import _lib, workloads
from kups_amd.core.verlet import VerletList, neighbor_capacity
from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
from kups_amd.md.integrators import VELOCITY_VERLET, EwaldLJPropagator
from kups_amd.potential.ewald import EwaldParameters
from kups_amd.potential.lennard_jones import LennardJonesParameters

pytestmark = pytest.mark.gpu

NACL_LJ = LennardJonesParameters.from_dict(12.0, {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)})


class Case:
    def __init__(self, b, dev, lj, rc_lj, alpha=None, rc_cl=None, lj_excl=None):
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        self.n_sys, self.n = b.n_sys, int(b.sizes.sum())
        self.maxn = int(b.sizes.max())
        self.sizes = b.sizes
        self.off = t(np.concatenate([[0], np.cumsum(b.sizes)]).astype(np.int32), torch.int32)
        self.pos = t(b.positions.astype(np.float32), torch.float32)
        self.q = None if alpha is None else t(b.charges.astype(np.float32), torch.float32)
        self.sp = t(b.species, torch.int32)
        self.tril = t(b.tril, torch.float64)
        self.lj = lj
        self.sig, self.eps = t(lj.sigma, torch.float64), t(lj.epsilon, torch.float64)
        self.rc_lj = t(np.full(b.n_sys, rc_lj), torch.float64)
        self.alpha = None if alpha is None else t(np.full(b.n_sys, alpha), torch.float64)
        self.rc_cl = None if rc_cl is None else t(np.full(b.n_sys, rc_cl), torch.float64)
        self.lj_excl = None if lj_excl is None else t(lj_excl.astype(np.int32), torch.int32)
        self.dev = dev

    def _out(self):
        return (torch.empty((self.n, 3), dtype=torch.float32, device=self.dev),
                torch.empty((self.n_sys, 2), dtype=torch.float64, device=self.dev),
                torch.empty((self.n_sys, 9), dtype=torch.float64, device=self.dev))

    def cells(self):
        f, e, w = self._out()
        _lib.check(_lib.load().kupsmi_pair_forces(
            _lib.ctx(), _lib.stream_ptr(), self.n_sys, _lib.ptr(self.off), self.n, self.maxn,
            _lib.ptr(self.pos), _lib.ptr(self.q), _lib.ptr(self.sp), self.lj.n_species,
            _lib.ptr(self.sig), _lib.ptr(self.eps), _lib.ptr(self.rc_lj), _lib.ptr(self.lj_excl),
            _lib.ptr(self.alpha), _lib.ptr(self.rc_cl), None, _lib.ptr(self.tril), 0, _lib.ptr(f),
            _lib.ptr(e), _lib.ptr(w), None))
        return f.cpu().numpy(), e.cpu().numpy(), w.cpu().numpy()

    def verlet(self, vl: VerletList):
        f, e, w = self._out()
        _lib.check(_lib.load().kupsmi_pair_forces_verlet(
            vl.handle, _lib.stream_ptr(), self.n_sys, _lib.ptr(self.off), self.n, self.maxn,
            _lib.ptr(self.pos), _lib.ptr(self.q), _lib.ptr(self.sp), self.lj.n_species,
            _lib.ptr(self.sig), _lib.ptr(self.eps), _lib.ptr(self.rc_lj), _lib.ptr(self.lj_excl),
            _lib.ptr(self.alpha), _lib.ptr(self.rc_cl), None, _lib.ptr(self.tril), vl.skin,
            vl.max_neighbors, 0, _lib.ptr(f), _lib.ptr(e), _lib.ptr(w)))
        return f.cpu().numpy(), e.cpu().numpy(), w.cpu().numpy()

    def wrap(self):
        """Wrap positions into the cell (cubic cells of the workloads here)."""
        L = self.tril[:, 0].repeat_interleave(torch.as_tensor(self.sizes, device=self.dev))
        p = self.pos.double()
        self.pos = (p - torch.floor(p / L[:, None] + 0.5) * L[:, None]).float().contiguous()


def _same(a, b):
    fa, ea, wa = a
    fb, eb, wb = b
    np.testing.assert_allclose(ea, eb, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(eb).max()))
    assert np.abs(fa - fb).max() <= 1e-5 * np.abs(fb).max()
    np.testing.assert_allclose(wa, wb, rtol=1e-5, atol=1e-5 * np.abs(wb).max())


def test_verlet_matches_cells_while_atoms_move_and_wrap(gpu):
    b = workloads.nacl_melts(ms=[5, 6, 5], seed=8)
    c = Case(b, gpu, NACL_LJ, 12.0, alpha=0.3, rc_cl=12.0)
    skin = 1.0
    vl = VerletList(skin, neighbor_capacity(b.sizes, b.tril, 12.0, skin))
    _same(c.verlet(vl), c.cells())
    assert vl.counters() == (3, 0)
    rng = np.random.default_rng(1)
    for k in range(4):
        # moves below skin/2: reuse; atoms near the faces get wrapped
        step = torch.as_tensor(rng.normal(0, 0.08, (c.n, 3)).clip(-0.24, 0.24), dtype=torch.float32,
                               device=gpu)
        c.pos = (c.pos + step).contiguous()
        c.wrap()
        _same(c.verlet(vl), c.cells())
    rebuilds, overflow = vl.counters()
    assert overflow == 0
    # beyond skin/2: every system rebuilds
    c.pos = (c.pos + torch.full_like(c.pos, 0.6)).contiguous()
    c.wrap()
    _same(c.verlet(vl), c.cells())
    assert vl.counters()[0] == rebuilds + 3


def test_verlet_rebuilds_only_the_triggered_systems(gpu):
    """One system of four moved beyond skin/2: only its list is rebuilt (the
    list build walks the cells of the triggered systems alone) while the
    others keep theirs, and every system still matches the cell-list path."""
    b = workloads.nacl_melts(ms=[5, 6, 5, 7], seed=13)
    c = Case(b, gpu, NACL_LJ, 12.0, alpha=0.3, rc_cl=12.0)
    skin = 1.0
    vl = VerletList(skin, neighbor_capacity(b.sizes, b.tril, 12.0, skin))
    _same(c.verlet(vl), c.cells())
    assert vl.counters() == (4, 0)
    off = np.concatenate([[0], np.cumsum(b.sizes)])
    rng = np.random.default_rng(2)
    for k, s in enumerate((2, 0, 3)):
        step = np.zeros((c.n, 3), np.float32)
        step[off[s]:off[s + 1]] = rng.normal(0, 0.5, (b.sizes[s], 3)).clip(-1.5, 1.5)
        # the other systems move a little: at most 3 x 0.05 sqrt(3) < skin/2
        # in all (reuse; wraps near the faces)
        mask = np.ones(c.n, bool)
        mask[off[s]:off[s + 1]] = False
        step[mask] = rng.normal(0, 0.02, (mask.sum(), 3)).clip(-0.05, 0.05)
        c.pos = (c.pos + torch.as_tensor(step, device=gpu)).contiguous()
        c.wrap()
        _same(c.verlet(vl), c.cells())
        assert vl.counters() == (4 + k + 1, 0)


@pytest.mark.parametrize("ms", [[7], [9, 5]])
def test_verlet_every_tile_of_few_large_systems(gpu, ms):
    """Few systems of many tiles (a single 2744-ion melt: 11 tiles of 256;
    5832 + 1000 ions: 27): every tile of the die-interleaved workgroup order
    is evaluated (a grid sized to the tile count alone left some out), and
    the list build covers every atom."""
    b = workloads.nacl_melts(ms=ms, seed=12)
    c = Case(b, gpu, NACL_LJ, 12.0, alpha=0.3, rc_cl=12.0)
    vl = VerletList(1.0, neighbor_capacity(b.sizes, b.tril, 12.0, 1.0))
    ref = c.cells()
    got = c.verlet(vl)
    assert (np.abs(got[0]).max(axis=1) > 0).all()  # no atom left without forces
    _same(got, ref)
    assert vl.counters() == (len(ms), 0)


def test_verlet_cells_of_more_than_64_atoms(gpu):
    """A long cutoff (24 A, skin 2 A) on an 8000-ion melt: cells of ~13 A hold
    ~120 atoms, so the list build walks each cell's stencil once per group of
    64 of its atoms and refills its candidate window many times per group;
    lists of ~3300 entries per atom."""
    b = workloads.nacl_melts(ms=[10], seed=14)
    lj = LennardJonesParameters.from_dict(24.0, {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)})
    c = Case(b, gpu, lj, 24.0, alpha=0.2, rc_cl=24.0)
    skin = 2.0
    vl = VerletList(skin, neighbor_capacity(b.sizes, b.tril, 24.0, skin))
    _same(c.verlet(vl), c.cells())
    assert vl.counters() == (1, 0)
    c.pos = (c.pos + 0.4).contiguous()
    c.wrap()
    _same(c.verlet(vl), c.cells())
    assert vl.counters() == (1, 0)


def test_verlet_overflowed_lists_use_the_stencil(gpu):
    b = workloads.nacl_melts(ms=[5], seed=9)
    c = Case(b, gpu, NACL_LJ, 12.0, alpha=0.3, rc_cl=12.0)
    vl = VerletList(0.5, 16)
    _same(c.verlet(vl), c.cells())
    assert vl.counters()[1] == c.n


def test_verlet_narrow_cells_with_image_shifts(gpu):
    # LJ argon boxes narrower than 2 (rc + skin): image grid, self images
    lj = LennardJonesParameters.from_dict(workloads.AR_CUTOFF, {"Ar": (workloads.AR_SIGMA, workloads.AR_EPS)})
    b = workloads.batched_lj_boxes(n_sys=6, nmin=64, nmax=120, seed=3)
    c = Case(b, gpu, lj, workloads.AR_CUTOFF)
    vl = VerletList(0.8, neighbor_capacity(b.sizes, b.tril, workloads.AR_CUTOFF, 0.8))
    _same(c.verlet(vl), c.cells())
    c.pos = (c.pos + 0.1).contiguous()
    c.wrap()
    _same(c.verlet(vl), c.cells())


def test_verlet_exclusion_groups(gpu):
    b = workloads.nacl_melts(ms=[5, 5], seed=10)
    groups = np.repeat(np.arange(b.sizes.sum() // 4), 4)
    c = Case(b, gpu, NACL_LJ, 10.0, alpha=0.3, rc_cl=12.0, lj_excl=groups)
    vl = VerletList(1.0, neighbor_capacity(b.sizes, b.tril, 12.0, 1.0))
    _same(c.verlet(vl), c.cells())
    c.pos = (c.pos - 0.2).contiguous()
    c.wrap()
    _same(c.verlet(vl), c.cells())


def test_ewald_md_with_skin_matches_rebuild_every_step(gpu):
    b = workloads.nacl_melts(ms=[5, 6], seed=11)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(0))
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    runs = []
    for skin in (None, 1.0):
        st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                           time_step_fs=1.0, temperature=1200.0, momenta=p0.astype(np.float32),
                           charges=b.charges)
        prop = EwaldLJPropagator(NACL_LJ, ep, VELOCITY_VERLET, skin=skin)
        prop.evaluate_forces(st)
        obs = prop(st, 30)
        runs.append((st.positions.cpu().numpy(), obs.potential_energy.cpu().numpy()))
    (p_a, e_a), (p_b, e_b) = runs
    assert np.abs(p_a - p_b).max() < 1e-3
    np.testing.assert_allclose(e_b, e_a, rtol=1e-5)
