"""GPU parity: neighbor-list export, LJ forces and the fused MD kernel vs the oracle.

Tolerances: pair sets bit-exact (same arrays, same order); FP32 energies and
forces within 1e-5 relative of the FP64 oracle on the same FP32 inputs
(forces as ||F_gpu - F_ref|| / ||F_ref||, energies per system).
"""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import workloads
from kups_amd.core.batch import ParticleBatch, SystemBatch
from kups_amd.core.neighborlist import CellListNeighborList
from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
from kups_amd.md.integrators import (FusedLJPropagator, make_baoab_langevin_step,
                                     make_velocity_verlet_step, normals)
from kups_amd.potential.lennard_jones import LennardJonesParameters, LennardJonesPotential

pytestmark = pytest.mark.gpu

AR = LennardJonesParameters.from_dict(workloads.AR_CUTOFF,
                                      {"Ar": (workloads.AR_SIGMA, workloads.AR_EPS)})


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def gpu_nl(pos, sizes, tril, cutoff, excl=None, dev="cuda"):
    pos = np.asarray(pos, np.float32)
    sizes = np.asarray(sizes)
    system = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    pb = ParticleBatch(torch.as_tensor(pos, device=dev), torch.as_tensor(system, device=dev),
                       exclusion=None if excl is None else torch.as_tensor(np.asarray(excl, np.int32), device=dev))
    sb = SystemBatch(torch.as_tensor(np.asarray(tril, np.float64).reshape(-1, 6), device=dev))
    cut = torch.as_tensor(np.broadcast_to(np.asarray(cutoff, np.float64), (len(sizes),)).copy(), device=dev)
    e = CellListNeighborList(avg_edges=1)(pb, None, sb, cut)
    return e.indices.cpu().numpy(), e.shifts[:, 0].cpu().numpy()


def ref_nl(pos, sizes, tril, cutoff, excl=None):
    pos = _f32(pos)
    off = np.concatenate([[0], np.cumsum(sizes)])
    if excl is None:
        excl = np.arange(len(pos))
    cut = np.broadcast_to(np.asarray(cutoff, np.float64), (len(sizes),))
    return oracle.neighborlist(pos, excl, off, tril, cut)


CASES = [
    # (positions, sizes, tril, cutoff, exclusion) -- the reference's NL test cases
    ([[0, 0, 0], [1, 0, 0], [2.1, 0, 0]], [3], [10, 0, 10, 0, 0, 10], 1.5, None),
    ([[0, 0, 0], [1, 0, 0]], [2], [10, 0, 10, 0, 0, 10], 1.0, None),
    ([[0.1, 0, 0], [2.9, 0, 0]], [2], [3, 0, 3, 0, 0, 3], 1.0, None),
    ([[0, 0, 0]], [1], [0.5, 0, 0.5, 0, 0, 0.5], 0.55, [0]),
    ([[0.1, 0.1, 0.1], [0.4, 0.1, 0.1]], [2], [0.5, 0, 0.5, 0, 0, 0.5], 0.4, [0, 0]),
    ([[0, 0, 0], [0.4, 0, 0], [0, 0, 0], [1, 0, 0]], [2, 2],
     [[1, 0, 10, 0, 0, 10], [10, 0, 10, 0, 0, 10]], [0.8, 1.5], None),
    ([[0, 0, 0], [0.3, 0, 0]], [2], [1, 0, 1, 0, 0, 1], np.inf, None),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_neighborlist_reference_cases_bit_exact(gpu, case):
    pos, sizes, tril, cut, excl = CASES[case]
    tril = np.asarray(tril, np.float64).reshape(-1, 6)
    gi, gs = gpu_nl(pos, sizes, tril, cut, excl)
    ri, rs = ref_nl(pos, sizes, tril, cut, excl)
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gs, rs)


def test_neighborlist_batched_boxes_bit_exact(gpu):
    # 64/90-atom boxes need the image grid; the others use the minimum image
    b = workloads.batched_lj_boxes(sizes=[64, 90, 333, 1000, 2100], seed=21)
    gi, gs = gpu_nl(b.positions, b.sizes, b.tril, workloads.AR_CUTOFF)
    ri, rs = ref_nl(b.positions, b.sizes, b.tril, workloads.AR_CUTOFF)
    assert len(ri) > 100000
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gs, rs)


def test_neighborlist_triclinic_and_capacity_regrow(gpu):
    rng = np.random.default_rng(4)
    lat = np.array([[9.0, 0, 0], [2.0, 8.5, 0], [1.5, -1.0, 9.5]])
    pos = rng.uniform(-0.5, 0.5, (120, 3)) @ lat
    tril = lat[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]][None]
    gi, gs = gpu_nl(pos, [120], tril, 4.2)  # avg_edges=1 forces a regrow
    ri, rs = ref_nl(pos, [120], tril, 4.2)
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gs, rs)


def test_neighborlist_empty_and_single(gpu):
    gi, gs = gpu_nl(np.zeros((1, 3)), [1], [[20, 0, 20, 0, 0, 20]], 5.0)
    assert len(gi) == 0


def _lj_batch(sizes, seed):
    return workloads.batched_lj_boxes(sizes=sizes, seed=seed)


def _check_energy(eg, er, sizes, rtol):
    """Per-system energies within rtol of max(|E|, N eps): the relative error of a
    system whose pair energies cancel to a small total is measured against the
    energy scale N eps, not against the near-zero total."""
    scale = np.maximum(np.abs(er), np.asarray(sizes) * workloads.AR_EPS)
    err = np.abs(np.asarray(eg) - er) / scale
    assert err.max() < rtol, (err.max(), eg, er)


def _check_forces(fg, fr, tol=1e-5):
    err = np.linalg.norm(fg - fr) / np.linalg.norm(fr)
    assert err < tol, err


@pytest.mark.parametrize("sizes", [[864], [64, 65, 99, 100, 101], [4096, 2000, 150, 77]])
def test_lj_forces_vs_oracle(gpu, sizes):
    b = _lj_batch(sizes, seed=len(sizes))
    pos32 = b.positions.astype(np.float32)
    system = np.repeat(np.arange(b.n_sys), b.sizes).astype(np.int32)
    pb = ParticleBatch(torch.as_tensor(pos32, device=gpu), torch.as_tensor(system, device=gpu))
    sb = SystemBatch(torch.as_tensor(b.tril, device=gpu))
    out = LennardJonesPotential(AR)(pb, sb)
    cut = np.full(b.n_sys, workloads.AR_CUTOFF)
    e, f, w = oracle.lj_cells(_f32(pos32), b.species, b.sys_off, b.tril, cut, AR.sigma, AR.epsilon)
    _check_energy(out.total_energies.cpu().numpy(), e, b.sizes, 1e-5)
    _check_forces(-out.position_gradients.cpu().numpy(), f)
    np.testing.assert_allclose(out.virial.cpu().numpy(), w, rtol=1e-5,
                               atol=1e-5 * np.abs(w).max())


def _md_state(b, seed=0, temperature=94.4, dt_fs=2.0, friction=0.0):
    rng = np.random.default_rng(seed)
    p = maxwell_boltzmann_momenta(b.masses, b.sizes, temperature, rng)
    return make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril,
                         workloads.AR_CUTOFF, time_step_fs=dt_fs, temperature=temperature,
                         friction_per_fs=friction, momenta=p.astype(np.float32))


def test_md_velocity_verlet_vs_oracle(gpu):
    b = _lj_batch([864, 64, 300], seed=3)
    st = _md_state(b)
    prop = make_velocity_verlet_step(AR)
    prop.evaluate_forces(st)
    pos = _f32(st.positions.cpu().numpy()); mom = _f32(st.momenta.cpu().numpy())
    frc = np.ascontiguousarray(_f32(st.forces.cpu().numpy()))
    n_steps = 10
    obs = prop(st, n_steps)
    cut = np.full(b.n_sys, workloads.AR_CUTOFF)
    dt = st.time_step.cpu().numpy()
    eta = np.zeros_like(pos)
    for k in range(n_steps):
        e, _ = oracle.md_step(0, pos, mom, b.masses.astype(np.float32), b.species, b.sys_off,
                              b.tril, cut, AR.sigma, AR.epsilon, dt, dt * 0, dt * 0, eta, frc)
        _check_energy(obs.potential_energy[:, k].cpu().numpy(), e, b.sizes, 1e-5)
    box = np.repeat(b.tril[:, 0], b.sizes)[:, None]
    d = st.positions.cpu().numpy() - pos
    d -= np.round(d / box) * box  # compare modulo the periodic box
    assert np.abs(d).max() < 1e-4
    _check_forces(st.forces.cpu().numpy(), frc, tol=1e-4)
    _check_forces(st.momenta.cpu().numpy(), mom, tol=1e-4)


def test_md_baoab_shared_noise_vs_oracle(gpu):
    b = _lj_batch([500, 80], seed=4)
    st = _md_state(b, friction=0.01)
    prop = make_baoab_langevin_step(AR, seed=1234)
    prop.evaluate_forces(st)
    pos = _f32(st.positions.cpu().numpy()); mom = _f32(st.momenta.cpu().numpy())
    frc = np.ascontiguousarray(_f32(st.forces.cpu().numpy()))
    cut = np.full(b.n_sys, workloads.AR_CUTOFF)
    dt = st.time_step.cpu().numpy()
    kT = st.kT.cpu().numpy(); gam = st.friction.cpu().numpy()
    step0 = st.step
    obs = prop(st, 5)
    for k in range(5):
        eta = normals(1234, step0 + k, b.n).cpu().numpy().astype(np.float64)
        e, _ = oracle.md_step(1, pos, mom, b.masses.astype(np.float32), b.species, b.sys_off,
                              b.tril, cut, AR.sigma, AR.epsilon, dt, kT, gam, eta, frc)
        _check_energy(obs.potential_energy[:, k].cpu().numpy(), e, b.sizes, 1e-5)
    _check_forces(st.momenta.cpu().numpy(), mom, tol=1e-4)


def test_normals_statistics(gpu):
    eta = normals(7, 3, 200000).cpu().numpy()
    assert abs(eta.mean()) < 0.01 and abs(eta.std() - 1) < 0.01
    eta2 = normals(7, 3, 200000).cpu().numpy()
    np.testing.assert_array_equal(eta, eta2)
    assert not np.array_equal(eta, normals(7, 4, 200000).cpu().numpy())


def test_md_full_batch_properties(gpu):
    """The full 2048-box batch: NVE energy conservation, momentum conservation,
    determinism, and oracle parity of sampled systems after the run."""
    b = workloads.batched_lj_boxes(n_sys=2048, seed=0)
    st = _md_state(b)
    prop = make_velocity_verlet_step(AR)
    prop.evaluate_forces(st)
    st2 = _md_state(b)
    st2.forces.copy_(st.forces)
    obs = prop(st, 20)
    prop(st2, 20, record=False)  # no observables: no reductions between the steps
    assert torch.equal(st.positions, st2.positions) and torch.equal(st.momenta, st2.momenta)
    etot = (obs.potential_energy + obs.kinetic_energy).cpu().numpy()
    n = b.sizes
    drift = np.abs(etot[:, -1] - etot[:, 0]) / n
    # hard-cutoff LJ (no shift, as the reference) jumps by u(rc) = -0.0163 eps
    # whenever a pair crosses rc, so small boxes drift in steps; the bulk must
    # conserve energy to FP32 accuracy
    assert np.median(drift) < 2e-5, np.median(drift)   # eV per atom over 20 steps
    assert drift.max() < 2e-3, drift.max()
    p = st.momenta.cpu().numpy().astype(np.float64)
    ptot = np.add.reduceat(p, b.sys_off[:-1], axis=0)
    assert np.abs(ptot).max() < 1e-3
    # sampled oracle parity of the final state's energies and forces
    pos = st.positions.cpu().numpy()
    for s in [0, 17, 1000, 2047]:
        a, z = b.sys_off[s], b.sys_off[s + 1]
        e, f, _ = oracle.lj_cells(_f32(pos[a:z]), b.species[a:z], [0, z - a], b.tril[s:s + 1],
                                  [workloads.AR_CUTOFF], AR.sigma, AR.epsilon)
        assert obs.potential_energy[s, -1].item() == pytest.approx(e[0], rel=1e-5)
        _check_forces(st.forces.cpu().numpy()[a:z], f)
