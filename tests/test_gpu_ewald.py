"""GPU parity of the general pair path (LJ + Ewald real space) and the
reciprocal-space Ewald kernels against the FP64 oracle.

Tolerances: energies 1e-5 relative per system (scale max(|E|, sum|q|^2 K/a)
for the real-space and Madelung cases), forces 1e-5 in norm
(||F_gpu - F_ref|| / ||F_ref||), k-space 1e-5 relative to the
system's total Ewald energy scale.
"""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import workloads
from kups_amd.core.batch import ParticleBatch, SystemBatch
from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
from kups_amd.md.integrators import VELOCITY_VERLET, EwaldLJPropagator
from kups_amd.potential.ewald import EwaldParameters, EwaldPotential, kvec_grid
from kups_amd.potential.lennard_jones import LennardJonesParameters

pytestmark = pytest.mark.gpu

NACL_LJ = LennardJonesParameters.from_dict(12.0, {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)})


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def _batch_tensors(b, dev):
    system = np.repeat(np.arange(b.n_sys), b.sizes).astype(np.int32)
    pb = ParticleBatch(torch.as_tensor(b.positions.astype(np.float32), device=dev),
                       torch.as_tensor(system, device=dev),
                       species=torch.as_tensor(b.species, device=dev),
                       charges=None if b.charges is None else
                       torch.as_tensor(b.charges.astype(np.float32), device=dev))
    sb = SystemBatch(torch.as_tensor(b.tril, device=dev))
    return pb, sb


def _rel_norm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_pair_path_lj_large_systems(gpu):
    """LJ through the HBM path, including a 6000-atom system beyond the LDS kernel."""
    from kups_amd import _lib

    b = workloads.batched_lj_boxes(sizes=[6000, 90, 1500], seed=2)
    pb, sb = _batch_tensors(b, gpu)
    sig = torch.as_tensor(np.array([[workloads.AR_SIGMA]]), device=gpu)
    eps = torch.as_tensor(np.array([[workloads.AR_EPS]]), device=gpu)
    cut = torch.full((b.n_sys,), workloads.AR_CUTOFF, dtype=torch.float64, device=gpu)
    off = torch.as_tensor(b.sys_off, device=gpu)
    f = torch.empty((b.n, 3), dtype=torch.float32, device=gpu)
    e = torch.empty((b.n_sys, 2), dtype=torch.float64, device=gpu)
    w = torch.empty((b.n_sys, 3, 3), dtype=torch.float64, device=gpu)
    _lib.check(_lib.load().kupsmi_pair_forces(
        _lib.ctx(), _lib.stream_ptr(), b.n_sys, _lib.ptr(off), b.n, int(b.sizes.max()),
        _lib.ptr(pb.positions), None, _lib.ptr(pb.species), 1, _lib.ptr(sig), _lib.ptr(eps),
        _lib.ptr(cut), None, None, None, None, _lib.ptr(sb.tril), 0, _lib.ptr(f), _lib.ptr(e),
        _lib.ptr(w), None))
    er, fr, wr = oracle.lj_cells(_f32(b.positions), b.species, b.sys_off, b.tril,
                                 np.full(b.n_sys, workloads.AR_CUTOFF), [[workloads.AR_SIGMA]],
                                 [[workloads.AR_EPS]])
    np.testing.assert_allclose(e[:, 0].cpu().numpy(), er, rtol=1e-5)
    assert _rel_norm(f.cpu().numpy(), fr) < 1e-5
    np.testing.assert_allclose(w.cpu().numpy(), wr, rtol=1e-5, atol=1e-5 * np.abs(wr).max())


def test_ewald_real_and_reciprocal_vs_oracle(gpu):
    b = workloads.nacl_melts(ms=[5, 6], seed=3)
    pb, sb = _batch_tensors(b, gpu)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    out = EwaldPotential(ep)(pb, sb, structure_factor=True)
    pos = _f32(b.positions)
    q = b.charges
    ij, sh = oracle.neighborlist(pos, np.arange(b.n), b.sys_off, b.tril, ep.cutoff)
    esr, fsr, wsr = oracle.ewald_real_edges(pos, q, b.sys_off, b.tril, ep.alpha, ep.cutoff, ij, sh)
    kv = np.stack([oracle.kvecs(t, 8) for t in b.tril])
    elr, flr, wlr, sk = oracle.ewald_recip(pos, q, b.sys_off, b.tril, ep.alpha, kv)
    es = oracle.ewald_self(q, b.sys_off, ep.alpha)
    np.testing.assert_allclose(out.short_range.cpu().numpy(), esr, rtol=1e-5)
    np.testing.assert_allclose(out.long_range.cpu().numpy(), elr, rtol=1e-5)
    np.testing.assert_allclose(out.self_energy.cpu().numpy(), es, rtol=1e-12)
    # structure factors: same grid order as the oracle's kvecs
    skg = out.structure_factor.cpu().numpy()
    scale = np.sqrt(b.sizes)[:, None, None]
    assert np.abs(skg - sk).max() / scale.min() < 1e-5
    assert _rel_norm(out.forces.cpu().numpy(), fsr + flr) < 1e-5
    np.testing.assert_allclose(out.virial.cpu().numpy(), wsr + wlr, rtol=1e-5,
                               atol=1e-5 * np.abs(wsr + wlr).max())


def test_ewald_madelung_nacl_gpu(gpu):
    """test_ewald.py:272-333: E / K = -1.7475646 * 4 * 5^3 (rtol 5e-5)."""
    base = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1],
                     [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    qb = np.array([-1, -1, -1, -1, 1, 1, 1, 1], float)
    g = np.stack(np.meshgrid(*[np.arange(5)] * 3, indexing="ij"), -1).reshape(-1, 1, 3) * 2.0
    pos = (g + base[None]).reshape(-1, 3)
    q = np.tile(qb, 125)
    b = workloads.Batch(pos, np.array([1000]), workloads.cubic_tril(10.0), np.zeros(1000, np.int32),
                        np.ones(1000), charges=q)
    pb, sb = _batch_tensors(b, gpu)
    ep = EwaldParameters.box(1, alpha=0.7, cutoff=4.9, nmax=8)
    out = EwaldPotential(ep)(pb, sb)
    e = out.total_energies.item() / oracle.coulomb_k()
    assert e == pytest.approx(-1.7475646 * 125 * 4, rel=5e-5)


def test_ewald_spherical_kmax_grid_matches_reference_set(gpu):
    """kvecs_from_kmax keeps |k| <= kmax: the GPU grid mask gives the same energy as
    the oracle summed over exactly the reference's vector set."""
    b = workloads.nacl_melts(ms=[5], seed=4)
    pb, sb = _batch_tensors(b, gpu)
    kmax = 1.2
    kv = kvec_grid(b.tril[0], kmax)
    n = int(np.abs(kv).max())
    ep = EwaldParameters(np.array([0.3]), np.array([12.0]), np.array([kmax]), (n, n, n))
    out = EwaldPotential(ep)(pb, sb)
    elr, flr, _, _ = oracle.ewald_recip(_f32(b.positions), b.charges, b.sys_off, b.tril, [0.3],
                                        kv[None])
    np.testing.assert_allclose(out.long_range.cpu().numpy(), elr, rtol=1e-5)


def test_pair_exclusion_groups(gpu):
    """GCMC exclusion semantics: LJ excluded within a group id (host atoms share
    one id), Coulomb by particle only."""
    from kups_amd import _lib

    rng = np.random.default_rng(5)
    L = 25.0
    n = 300
    pos = rng.uniform(-L / 2, L / 2, (n, 3)).astype(np.float32)
    grp = np.where(np.arange(n) < 200, 10**6, np.arange(n) // 3).astype(np.int32)
    q = rng.normal(0, 0.3, n).astype(np.float32)
    q -= q.mean()
    sp = (np.arange(n) % 2).astype(np.int32)
    tril = workloads.cubic_tril(L)
    lj = NACL_LJ
    keep = []  # device buffers must outlive the (asynchronous) launch

    def t(a, dt=torch.float64):
        keep.append(torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=gpu))
        return _lib.ptr(keep[-1])

    f = torch.empty((n, 3), dtype=torch.float32, device=gpu)
    e = torch.empty((1, 2), dtype=torch.float64, device=gpu)
    w = torch.empty((1, 3, 3), dtype=torch.float64, device=gpu)
    _lib.check(_lib.load().kupsmi_pair_forces(
        _lib.ctx(), _lib.stream_ptr(), 1, t([0, n], torch.int32), n, n, t(pos, torch.float32),
        t(q, torch.float32), t(sp, torch.int32), 2, t(lj.sigma), t(lj.epsilon), t([11.0]),
        t(grp, torch.int32), t([0.26]), t([12.0]), None, t(tril), 0, _lib.ptr(f),
        _lib.ptr(e), _lib.ptr(w), None))
    torch.cuda.synchronize()
    p64 = _f32(pos)
    ij, sh = oracle.neighborlist(p64, grp, [0, n], tril, [11.0])
    elj, flj, _ = oracle.lj_edges(p64, sp, [0, n], tril, [11.0], lj.sigma, lj.epsilon, ij, sh)
    ij2, sh2 = oracle.neighborlist(p64, np.arange(n), [0, n], tril, [12.0])
    esr, fsr, _ = oracle.ewald_real_edges(p64, q, [0, n], tril, [0.26], [12.0], ij2, sh2)
    assert e[0, 0].item() == pytest.approx(elj[0], rel=1e-5)
    assert e[0, 1].item() == pytest.approx(esr[0], rel=1e-5)
    assert _rel_norm(f.cpu().numpy(), flj + fsr) < 1e-5


def _oracle_forces(b, pos, alpha, rc, kv):
    e_lj, f_lj, _ = oracle.lj_cells(pos, b.species, b.sys_off, b.tril, np.full(b.n_sys, rc),
                                    NACL_LJ.sigma, NACL_LJ.epsilon)
    ij, sh = oracle.neighborlist(pos, np.arange(b.n), b.sys_off, b.tril, np.full(b.n_sys, rc))
    esr, fsr, _ = oracle.ewald_real_edges(pos, b.charges, b.sys_off, b.tril, np.full(b.n_sys, alpha),
                                          np.full(b.n_sys, rc), ij, sh)
    elr, flr, _, _ = oracle.ewald_recip(pos, b.charges, b.sys_off, b.tril, np.full(b.n_sys, alpha), kv)
    es = oracle.ewald_self(b.charges, b.sys_off, np.full(b.n_sys, alpha))
    return e_lj + esr + elr + es, f_lj + fsr + flr


def test_ewald_md_velocity_verlet_vs_oracle(gpu):
    b = workloads.nacl_melts(ms=[5, 5], seed=6)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(0))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=1200.0, momenta=p0.astype(np.float32),
                       charges=b.charges)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    prop = EwaldLJPropagator(NACL_LJ, ep, VELOCITY_VERLET)
    prop.evaluate_forces(st)
    pos = _f32(st.positions.cpu().numpy())
    mom = _f32(st.momenta.cpu().numpy())
    frc = _f32(st.forces.cpu().numpy())
    kv = np.stack([oracle.kvecs(t, 8) for t in b.tril])
    m = b.masses[:, None]
    dt = st.time_step.cpu().numpy()[0]
    obs = prop(st, 4)
    for k in range(4):
        mom += frc * dt / 2
        pos = pos + mom / m * dt
        L = b.tril[0, 0]
        pos = workloads.wrap_cubic(pos, L)
        e, frc = _oracle_forces(b, pos, 0.3, 12.0, kv)
        mom += frc * dt / 2
        eg = obs.potential_energy[:, k].cpu().numpy()
        np.testing.assert_allclose(eg, e, rtol=1e-5)
    assert _rel_norm(st.forces.cpu().numpy(), frc) < 1e-4


def _oracle_forces_w(b, pos, tril, alpha, rc, kv):
    """FP64 LJ + Ewald energies, forces and virial at (pos, tril)."""
    rcs = np.full(b.n_sys, rc)
    e_lj, f_lj, w_lj = oracle.lj_cells(pos, b.species, b.sys_off, tril, rcs, NACL_LJ.sigma,
                                       NACL_LJ.epsilon)
    ij, sh = oracle.neighborlist(pos, np.arange(b.n), b.sys_off, tril, rcs)
    esr, fsr, wsr = oracle.ewald_real_edges(pos, b.charges, b.sys_off, tril,
                                            np.full(b.n_sys, alpha), rcs, ij, sh)
    elr, flr, wlr, _ = oracle.ewald_recip(pos, b.charges, b.sys_off, tril, np.full(b.n_sys, alpha),
                                          kv)
    es = oracle.ewald_self(b.charges, b.sys_off, np.full(b.n_sys, alpha))
    return e_lj + esr + elr + es, f_lj + fsr + flr, w_lj + wsr + wlr


@pytest.mark.parametrize("npt", [False, True])
def test_ewald_md_csvr_and_npt_vs_oracle(gpu, npt):
    """The HBM path's CSVR thermostat and stochastic cell rescaling
    (kupsmi_thermostat_barostat) around velocity Verlet with LJ + Ewald forces,
    step by step against the oracle's CSVR rescaling (same Philox draws), FP64
    velocity Verlet, cell rescaling from the FP64 virial and a fresh FP64
    force evaluation at the rescaled state."""
    from kups_amd.md.integrators import CSVR, CSVR_NPT

    b = workloads.nacl_melts(ms=[5, 6], seed=7)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(1))
    # a strong barostat coupling so the cells move visibly within a few steps
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=1200.0, momenta=p0.astype(np.float32),
                       charges=b.charges, thermostat_time_constant_fs=20.0,
                       target_pressure_pa=1.0e8, pressure_coupling_time_fs=20.0,
                       compressibility_per_pa=2.0e-9)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=6)
    seed = 23
    prop = EwaldLJPropagator(NACL_LJ, ep, CSVR_NPT if npt else CSVR, seed=seed)
    prop.evaluate_forces(st)
    pos = _f32(st.positions.cpu().numpy())
    mom = _f32(st.momenta.cpu().numpy())
    frc = _f32(st.forces.cpu().numpy())
    tril = np.ascontiguousarray(b.tril.astype(np.float64))
    kv = np.stack([oracle.kvecs(t, 6) for t in b.tril])
    m = b.masses.astype(np.float32).astype(np.float64)
    dt = st.time_step.cpu().numpy()
    kT = st.kT.cpu().numpy()
    tau = st.thermostat_time_constant.cpu().numpy()
    baro = st.barostat.cpu().numpy()
    step0 = st.step
    n = 5
    obs = prop(st, n)
    mus = []
    for k in range(n):
        oracle.csvr_scale(seed, step0 + k, mom, m, b.sys_off, kT, dt, tau)
        mom += frc * dt[0] / 2
        pos = np.ascontiguousarray(pos + mom / m[:, None] * dt[0])
        e, frc, w = _oracle_forces_w(b, pos, tril, 0.3, 12.0, kv)
        mom += frc * dt[0] / 2
        if npt:
            mus.append(oracle.scr_scale(seed, step0 + k, mom, pos, tril, m, b.sys_off, kT, dt, baro,
                                        w))
            e, frc, w = _oracle_forces_w(b, pos, tril, 0.3, 12.0, kv)
        np.testing.assert_allclose(obs.potential_energy[:, k].cpu().numpy(), e, rtol=1e-5)
    if npt:
        assert np.abs(np.log(np.array(mus))).max() > 1e-5  # the cells did move
        np.testing.assert_allclose(st.tril.cpu().numpy(), tril, rtol=1e-6)
    else:
        np.testing.assert_array_equal(st.tril.cpu().numpy(), b.tril)
    assert _rel_norm(st.forces.cpu().numpy(), frc) < 1e-4
    assert _rel_norm(st.momenta.cpu().numpy(), mom) < 1e-4


@pytest.mark.parametrize("integ", ["baoab", "npt"])
def test_ewald_md_graph_replays_match_eager_steps(gpu, integ):
    """Steps recorded in a HIP graph (GraphedSteps) and replayed three times,
    with eager work and synchronisations between the replays, give the same
    bits as the same steps run eagerly: positions, momenta, forces, cells and
    every energy (the random streams follow the device-side step counter)."""
    from kups_amd.md.integrators import BAOAB_LANGEVIN, CSVR_NPT

    b = workloads.nacl_melts(ms=[5, 6], seed=9)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(2))

    def state():
        return make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril,
                             12.0, time_step_fs=1.0, temperature=1200.0,
                             momenta=p0.astype(np.float32), charges=b.charges,
                             friction_per_fs=0.01, thermostat_time_constant_fs=20.0,
                             target_pressure_pa=1.0e8, pressure_coupling_time_fs=20.0,
                             compressibility_per_pa=2.0e-9)

    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=6)
    integrator = BAOAB_LANGEVIN if integ == "baoab" else CSVR_NPT
    skin = 1.0 if integ == "baoab" else None
    sa, sb = state(), state()
    pa = EwaldLJPropagator(NACL_LJ, ep, integrator, seed=5, skin=skin)
    pb = EwaldLJPropagator(NACL_LJ, ep, integrator, seed=5, skin=skin)
    pa.evaluate_forces(sa)
    pb.evaluate_forces(sb)
    graph = pb.capture(sb, 3)
    # a larger batch evaluated eagerly with the same propagator and with the
    # shared context grows their buffers; the graph's own are untouched
    big = workloads.nacl_melts(ms=[7, 7, 6], seed=10)
    sbig = make_md_state(big.positions.astype(np.float32), big.masses, big.species, big.sizes,
                         big.tril, 12.0, time_step_fs=1.0, temperature=1200.0,
                         charges=big.charges)
    epbig = EwaldParameters.box(big.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    for r in range(3):
        oa = pa(sa, 3)
        ob = graph.replay()
        torch.cuda.synchronize()
        assert float(torch.ones(4096, device=gpu).sum().item()) == 4096.0  # eager work
        if r == 0:
            pb.ewald, ep_saved = epbig, pb.ewald
            pb._dev = None
            pb.evaluate_forces(sbig)
            pb.ewald = ep_saved
            pb._dev = None
            torch.cuda.synchronize()
        for name in ("positions", "momenta", "forces", "tril"):
            torch.testing.assert_close(getattr(sb, name), getattr(sa, name), rtol=0, atol=0)
        torch.testing.assert_close(ob.potential_energy, oa.potential_energy, rtol=0, atol=0)
        torch.testing.assert_close(ob.kinetic_energy, oa.kinetic_energy, rtol=0, atol=0)
        assert sb.step == sa.step
    if integ == "npt":
        assert not torch.equal(sb.tril, torch.as_tensor(b.tril, device=gpu))


def test_frozen_context_and_verlet_handle_refuse_to_grow(gpu):
    """A frozen context (the one a captured graph records with) fails a call
    that needs more scratch instead of freeing buffers the graph holds; so
    does a frozen Verlet handle asked to resize."""
    from kups_amd import _lib
    from kups_amd.core.verlet import VerletList

    small = workloads.nacl_melts(ms=[3], seed=1)
    big = workloads.nacl_melts(ms=[5, 5], seed=1)
    ctx = _lib.Context()
    prop = EwaldLJPropagator(NACL_LJ, EwaldParameters.box(1, alpha=0.3, cutoff=12.0, nmax=4),
                             VELOCITY_VERLET, skin=1.0)
    prop.context = ctx

    def st(b):
        return make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril,
                             12.0, time_step_fs=1.0, temperature=300.0, charges=b.charges)

    e_small = prop.evaluate_forces(st(small))[0].clone()
    ctx.freeze()
    prop.verlet.freeze()
    np.testing.assert_array_equal(prop.evaluate_forces(st(small))[0].cpu().numpy(),
                                  e_small.cpu().numpy())  # same size: no growth needed
    prop.ewald = EwaldParameters.box(2, alpha=0.3, cutoff=12.0, nmax=4)
    prop._dev = None
    with pytest.raises(_lib.KupsmiError, match="frozen"):
        prop.evaluate_forces(st(big))
    prop.verlet.freeze(False)
    with pytest.raises(_lib.KupsmiError, match="frozen context"):
        prop.evaluate_forces(st(big))
    ctx.freeze(False)
    prop.evaluate_forces(st(big))
    torch.cuda.synchronize()


def test_pair_tail_corrected_lj_vs_oracle(gpu):
    """PairTailCorrectedLennardJonesPotential (kupsmi_pair_forces with
    lj_truncation) against the oracle's switched LJ: batched boxes (image
    grid for the small ones), two species, one system unswitched; and the
    reference's regime cases (test_lennard_jones.py:316-387)."""
    from kups_amd.potential.lennard_jones import (LennardJonesParameters,
                                                  PairTailCorrectedLennardJonesPotential)
    b = workloads.batched_lj_boxes(sizes=[64, 90, 700, 1500], seed=13)
    rng = np.random.default_rng(2)
    spec = rng.integers(0, 2, b.n).astype(np.int32)
    s0 = workloads.AR_SIGMA
    params = LennardJonesParameters.from_dict(workloads.AR_CUTOFF, {"A": (s0, workloads.AR_EPS),
                                                                    "B": (1.1 * s0, 0.7 * workloads.AR_EPS)})
    rc = workloads.AR_CUTOFF
    rtr = np.array([0.6, 0.8, 0.9, 1.5]) * rc  # the last system is not switched
    pot = PairTailCorrectedLennardJonesPotential(params, torch.as_tensor(rtr))
    pb = ParticleBatch(torch.as_tensor(np.asarray(b.positions, np.float32), device=gpu),
                       torch.as_tensor(np.repeat(np.arange(4), b.sizes).astype(np.int32), device=gpu),
                       species=torch.as_tensor(spec, device=gpu))
    sb = SystemBatch(torch.as_tensor(b.tril, device=gpu))
    out = pot(pb, sb)
    pos32 = np.asarray(b.positions, np.float32).astype(np.float64)
    cut = np.full(4, rc)
    ij, sh = oracle.neighborlist(pos32, np.arange(b.n), b.sys_off, b.tril, cut)
    e, f, w = oracle.lj_switched_edges(pos32, spec, b.sys_off, b.tril, cut, rtr,
                                       params.sigma, params.epsilon, ij, sh)
    np.testing.assert_allclose(out.total_energies.cpu().numpy(), e, rtol=1e-5)
    fg = -out.position_gradients.double().cpu().numpy()
    assert np.linalg.norm(fg - f) / np.linalg.norm(f) < 1e-5
    np.testing.assert_allclose(out.virial.cpu().numpy(), w, rtol=1e-5, atol=1e-5 * np.abs(w).max())
    # regimes: A-B at r = 1 (sigma 1.2, eps 0.8 by the tables below)
    lj2 = LennardJonesParameters(labels=("A", "B"), sigma=np.array([[1.0, 1.2], [1.2, 1.0]]),
                                 epsilon=np.array([[1.0, 0.8], [0.8, 1.0]]), cutoff=12.0)
    two = ParticleBatch(torch.tensor([[0.0, 0, 0], [1.0, 0, 0]], device=gpu),
                        torch.zeros(2, dtype=torch.int32, device=gpu),
                        species=torch.tensor([0, 1], dtype=torch.int32, device=gpu))
    box = SystemBatch(torch.tensor([[10.0, 0, 10, 0, 0, 10]], dtype=torch.float64, device=gpu))
    x = 1.2 ** 6
    target = 4 * 0.8 * (x * x - x)  # both directed edges: 2 x 1/2 u
    def sw(rc_, rtr_):
        return PairTailCorrectedLennardJonesPotential(lj2, rtr_)(two, box, rc_).total_energies.item()
    assert sw(2.0, 1.5) == pytest.approx(target, rel=1e-6)
    assert sw(1.2, 1.0) == pytest.approx(target, rel=1e-6)
    assert 0.0 < sw(1.2, 0.8) < target
    assert sw(1.0, 0.8) == 0.0
