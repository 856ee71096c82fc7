"""The benchmark configurations, each run through the exact GPU path the
benchmark drives and compared with the FP64 oracle.

Long runs are compared in segments: every segment starts the oracle from the
GPU state at the segment's start (FP32 values widened to FP64), runs the same
steps with the same random draws (the Philox normals of the O step are read
back from the device stream), and compares every step's energy and the
segment's end state.  So a trajectory of any length is checked step by step
without the chaotic growth of FP32-vs-FP64 differences over many steps.

Tolerances: per-system potential energies 1e-5 relative; end-of-segment
momenta 1e-5 relative in the 2-norm over the batch (FP32 integration against
FP64 over the segment); positions 1e-4 A absolute modulo the periodic box;
forces 1e-5 relative in the 2-norm against the FP64 forces at the GPU's own
end positions.  Forces are compared at the same positions because the net
force on an atom near a lattice site is a small residual of cancelling pair
forces: the 1e-6 A by which FP32 and FP64 trajectories part over a segment
moves it by more than 1e-5 of itself (at the same positions the oracle's own
FP32 restatement is 6.5e-6 off on the 864-atom crystal).
"""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import workloads
from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
from kups_amd.md.integrators import (BAOAB_LANGEVIN, EwaldLJPropagator,
                                     make_baoab_langevin_step, normals)
from kups_amd.potential.ewald import EwaldParameters
from kups_amd.potential.lennard_jones import LennardJonesParameters

pytestmark = pytest.mark.gpu

AR = LennardJonesParameters.from_dict(workloads.AR_CUTOFF,
                                      {"Ar": (workloads.AR_SIGMA, workloads.AR_EPS)})
NACL_LJ = LennardJonesParameters.from_dict(12.0, {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)})


def _f64(t):
    return np.ascontiguousarray(t.detach().cpu().numpy().astype(np.float64))


def _rel_norm(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pos_diff(a, b, tril, sizes):
    """max |a - b| modulo the (lower-triangular) periodic cells."""
    d = np.asarray(a, float) - np.asarray(b, float)
    out = 0.0
    o = 0
    for t, n in zip(np.asarray(tril, float), sizes):
        L = np.array([[t[0], 0, 0], [t[1], t[2], 0], [t[3], t[4], t[5]]])
        f = d[o:o + n] @ np.linalg.inv(L)
        f -= np.round(f)
        out = max(out, float(np.abs(f @ L).max()))
        o += n
    return out


def test_cfg1_argon_fcc_baoab_1000_steps_vs_oracle(gpu):
    """864 Ar on the 6x6x6 FCC lattice (a = 5.26 A), rc = 2.5 sigma,
    Maxwell-Boltzmann at 94.4 K, BAOAB with gamma = 1/ps, dt = 2 fs, 1000
    FP32 steps through the fused LDS kernel with its Verlet lists: 50
    segments of 20 steps (one launch each), every step's energy and each
    segment's end state against the oracle's FP64 BAOAB (orc_md_step)."""
    b = workloads.argon_fcc()
    assert b.n == 864
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 94.4, np.random.default_rng(0))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril,
                       workloads.AR_CUTOFF, time_step_fs=2.0, temperature=94.4,
                       friction_per_fs=0.001, momenta=p0.astype(np.float32))
    seed = 77
    prop = make_baoab_langevin_step(AR, seed=seed)
    prop.evaluate_forces(st)
    cut = np.full(b.n_sys, workloads.AR_CUTOFF)
    mass = b.masses.astype(np.float32).astype(np.float64)
    dt = st.time_step.cpu().numpy()
    kT = st.kT.cpu().numpy()
    gam = st.friction.cpu().numpy()
    seg = 20
    for s in range(1000 // seg):
        pos, mom, frc = _f64(st.positions), _f64(st.momenta), _f64(st.forces)
        step0 = st.step
        eta = [normals(seed, step0 + k, b.n).double().cpu().numpy() for k in range(seg)]
        obs = prop(st, seg)
        pe = obs.potential_energy.cpu().numpy()
        for k in range(seg):
            e, _ = oracle.md_step(1, pos, mom, mass, b.species, b.sys_off, b.tril, cut, AR.sigma,
                                  AR.epsilon, dt, kT, gam, np.ascontiguousarray(eta[k]), frc)
            err = np.abs(pe[:, k] - e) / np.abs(e)
            assert err.max() < 1e-5, (s, k, pe[:, k], e)
        gpos = _f64(st.positions)
        assert _pos_diff(gpos, pos, b.tril, b.sizes) < 1e-4, s
        assert _rel_norm(_f64(st.momenta), mom) < 1e-5, s
        _, f_at, _ = oracle.lj_cells(gpos, b.species, b.sys_off, b.tril, cut, AR.sigma,
                                     AR.epsilon)
        assert _rel_norm(_f64(st.forces), f_at) < 1e-5, s
    assert st.step == 1000
    # the thermostat holds the lattice near 94.4 K
    kin = 0.5 * (_f64(st.momenta) ** 2 / mass[:, None]).sum()
    T = 2 * kin / (3 * b.n * float(kT[0])) * 94.4
    assert 60.0 < T < 130.0, T


def _ewald_forces(b, pos, alpha, rc, kv):
    rcs = np.full(b.n_sys, rc)
    e_lj, f_lj, _ = oracle.lj_cells(pos, b.species, b.sys_off, b.tril, rcs, NACL_LJ.sigma,
                                    NACL_LJ.epsilon)
    ij, sh = oracle.neighborlist(pos, np.arange(b.n), b.sys_off, b.tril, rcs)
    esr, fsr, _ = oracle.ewald_real_edges(pos, b.charges, b.sys_off, b.tril,
                                          np.full(b.n_sys, alpha), rcs, ij, sh)
    elr, flr, _, _ = oracle.ewald_recip(pos, b.charges, b.sys_off, b.tril,
                                        np.full(b.n_sys, alpha), kv)
    es = oracle.ewald_self(b.charges, b.sys_off, np.full(b.n_sys, alpha))
    return e_lj + esr + elr + es, f_lj + fsr + flr


@pytest.mark.parametrize("skin", [1.0, 1.5])
def test_cfg3_bench_path_baoab_verlet_vs_oracle(gpu, skin):
    """The LJ + Ewald benchmark path exactly as the benchmark builds it
    (EwaldLJPropagator, BAOAB with gamma = 1/ps at 1200 K, dt = 1 fs, rc =
    12 A, alpha = 0.3 / A, |n| <= 8 per axis, Verlet lists with the given
    skin) on melts of m = 5, 7 and 10 (1000, 2744 and 8000 ions): three
    checked segments of 4 steps, separated by 100 unchecked steps so that the
    later segments run on reused and on rebuilt lists, against the oracle's
    FP64 BAOAB (B, A, O, A, B with the same normals) and FP64 LJ + Ewald
    forces."""
    b = workloads.nacl_melts(ms=[5, 7, 10], seed=3)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(0))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=1200.0, momenta=p0.astype(np.float32),
                       charges=b.charges, friction_per_fs=0.001)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    seed = 11
    prop = EwaldLJPropagator(NACL_LJ, ep, BAOAB_LANGEVIN, seed=seed, skin=skin)
    prop.evaluate_forces(st)
    kv = np.stack([oracle.kvecs(t, 8) for t in b.tril])
    mass = b.masses.astype(np.float32).astype(np.float64)[:, None]
    dt = float(st.time_step[0])
    h = 0.5 * dt
    gam = float(st.friction[0])
    kT = float(st.kT[0])
    c1 = np.exp(-gam * dt)
    c2 = np.sqrt(kT * (1.0 - c1 * c1))
    seg = 4
    for r in range(3):
        if r:
            prop(st, 100)
        pos, mom, frc = _f64(st.positions), _f64(st.momenta), _f64(st.forces)
        step0 = st.step
        eta = [normals(seed, step0 + k, b.n).double().cpu().numpy() for k in range(seg)]
        obs = prop(st, seg)
        pe = obs.potential_energy.cpu().numpy()
        for k in range(seg):
            mom += frc * h
            pos = pos + mom / mass * h
            mom = c1 * mom + c2 * np.sqrt(mass) * eta[k]
            pos = np.ascontiguousarray(pos + mom / mass * h)
            e, frc = _ewald_forces(b, pos, 0.3, 12.0, kv)
            mom += frc * h
            np.testing.assert_allclose(pe[:, k], e, rtol=1e-5, err_msg=f"segment {r} step {k}")
        gpos = _f64(st.positions)
        assert _pos_diff(gpos, pos, b.tril, b.sizes) < 1e-4, r
        assert _rel_norm(_f64(st.momenta), mom) < 1e-5, r
        _, f_at = _ewald_forces(b, gpos, 0.3, 12.0, kv)
        assert _rel_norm(_f64(st.forces), f_at) < 1e-5, r
    rebuilds, overflow = prop.verlet.counters()
    assert rebuilds > b.n_sys and overflow == 0  # lists were reused and rebuilt


def _gcmc_trace_errors(g, n_steps, n_cmp=None):
    """Trial-by-trial comparison of a GCMC batch with the oracle.  Delta E
    within 1e-5 of max(|Delta E|, 1 eV) for every trial with |Delta E| <=
    50 kT (the trials whose acceptance Delta E can decide: beyond 50 kT the
    Metropolis ratio is below 2e-22).  The 1 eV floor: Delta E is the
    difference of two guest-host interaction energies, each a sum of several
    hundred FP32 pair terms of 0.1-1 eV in a charged host, so 1e-5 of those
    sums is about 1e-5 eV (4e-4 kT at 298 K).  Beyond 50 kT, overlaps whose
    r^-12 terms carry the FP32 cancellation of r = |r_j - r_i| at
    sub-angstrom separations: within 1e-3 and with the same decision.
    Returns the relative errors of the decision-relevant trials and the
    accepted count."""
    import math

    n_cmp = g.n_chain if n_cmp is None else n_cmp
    st = oracle.gcmc_state(g, chains=np.arange(n_cmp))
    kT = g.d["kT"].cpu().numpy()
    tr_gpu = g.run(n_steps, n_trace=n_cmp).cpu().numpy()
    tr_orc = oracle.gcmc_run(st, n_steps, 0, g.seed, n_trace=n_cmp)
    errs, accepted = [], 0
    for c in range(n_cmp):
        kt = float(kT[c])
        lu = [math.log((float(oracle.philox(g.seed, c, t, 0x47434D41)[1]) + 0.5) * 2.0**-32)
              for t in range(n_steps)]
        for t in range(n_steps):
            a, b = tr_gpu[c, t], tr_orc[c, t]
            assert a[0] == b[0], (c, t)  # same move type and coin
            if abs(b[1]) <= 50 * kt:
                e = abs(a[1] - b[1]) / max(abs(b[1]), 1.0)
                errs.append(e)
                assert e <= 1e-5, (c, t, a, b)
            else:
                assert abs(a[1] - b[1]) <= 1e-3 * abs(b[1]), (c, t, a, b)
            if a[3] != b[3]:
                # a near tie that FP32 resolves the other way; the chain's
                # trajectories part here
                assert abs(b[2] - lu[t]) < 1e-3, (c, t, a, b, lu[t])
                break
            accepted += int(a[3])
    return np.array(errs), accepted


def test_cfg4_dense_hosts_trial_by_trial_vs_oracle(gpu):
    """The plain-GCMC hosts at their configured defaults (1500-3000 atoms on a
    jittered cubic lattice with 30 % vacancies, triclinic cells of 25-30 A,
    Normal(0, 0.3) charges, fugacities 0.1-10 bar) on 6 chains x 400 trials,
    trial by trial against the oracle (tolerances in _gcmc_trace_errors).
    The GPU sums FP32 pair terms in FP64 and keeps S(k) in FP64, so its
    error is the FP32 rounding of the individual terms (about 6e-8 of each),
    a few 1e-7 eV for the few hundred host atoms within the cutoff of a
    guest."""
    from kups_amd.mcmc.gcmc import make_co2_gcmc

    h = workloads.gcmc_hosts(6, seed=5)
    assert h.sizes.min() >= 1500 and h.sizes.max() <= 3000
    g = make_co2_gcmc(h, gmax=64, seed=21)
    errs, accepted = _gcmc_trace_errors(g, 400)
    assert len(errs) > 100
    assert accepted > 0
    print(f"dense hosts: {len(errs)} trials, {accepted} accepted, Delta E rel err "
          f"median {np.median(errs):.2e} max {errs.max():.2e}")


def test_loaded_charged_triclinic_hosts_trial_by_trial_vs_oracle(gpu):
    """Chains that load: the plain-GCMC host generator (30 % vacancies,
    triclinic 25-30 A cells, Normal(0, 0.3) charges) with 500-800 atoms at
    1e7 Pa, so that insertions are accepted and deletions, translations and
    rotations move real guests (the configured 1500-3000-atom cells are
    about 0.1 atoms per cubic angstrom, with no cavity a CO2 fits: the
    benchmark's uniform sampler accepts about one insertion in 10^7 trials
    there), on 4 chains x 600 trials against the oracle."""
    from kups_amd.mcmc.gcmc import make_co2_gcmc

    h = workloads.gcmc_hosts(4, nmin=500, nmax=800, seed=6)
    h.pressure[:] = 1e7
    g = make_co2_gcmc(h, gmax=64, seed=22)
    errs, accepted = _gcmc_trace_errors(g, 600)
    assert int(g.loading.sum()) > 0
    s = g.stats
    assert s["insertions"].sum() > 0 and s["accepted"][:, :3].sum() > 0
    print(f"loaded hosts: {len(errs)} trials, {accepted} accepted, Delta E rel err "
          f"median {np.median(errs):.2e} max {errs.max():.2e}")


@pytest.mark.parametrize("chains", [1024, 4096])
def test_full_size_gcmc_vs_oracle_and_bookkeeping(gpu, chains):
    """The plain (1024 chains) and learned-proposal (4096 chains) GCMC batch
    sizes at full size with the configured hosts: 200 trials per chain, the
    first 4 chains trial by trial against the oracle; for every chain the
    loading equals accepted insertions minus accepted deletions and the
    tracked energy equals a fresh full evaluation (up to the intramolecular
    real-space term the reference's full evaluation holds for every
    guest)."""
    from kups_amd.mcmc.gcmc import make_co2_gcmc

    h = workloads.gcmc_hosts(chains, seed=0)
    g = make_co2_gcmc(h, gmax=64, seed=1)
    e0 = g.d["energy"].cpu().numpy().copy()
    ng0 = g.loading.cpu().numpy().copy()
    nt0 = g.motif_counts()
    errs, _ = _gcmc_trace_errors(g, 200, n_cmp=4)
    assert len(errs) > 100
    ng = g.loading.cpu().numpy()
    s = g.stats
    assert int(s["trials"].sum()) == 200 * chains
    np.testing.assert_array_equal(ng - ng0, (s["insertions"] - s["deletions"]).astype(ng.dtype))
    tracked = g.d["energy"].cpu().numpy()
    full = g.full_energy() - ((g.motif_counts() - nt0) * g.sr_intramolecular()).sum(1)
    scale = np.maximum(np.abs(tracked - e0), 1.0)
    assert (np.abs(tracked - full) / scale).max() < 1e-4
