"""GPU parity of the GCMC kernel against the FP64 oracle under the shared
Philox stream.

Trial by trial: identical move sequence; energy differences within
1e-5 * max(|dE|, 1 eV) where |dE| <= 50 kT (FP32 pair sums vs FP64; 1e-5 eV
is 4e-4 kT at 298 K; beyond 50 kT the Metropolis ratio is below 2e-22) and
within 1e-3 relative beyond; identical accept decisions, except that a decision may differ where
|log p - log u| is below 1e-3 (a near tie that FP32 may resolve either way),
after which the chain's trajectories diverge and the comparison of that chain
stops.
"""

import math

import numpy as np
import pytest
import torch

import oracle
from kups_amd import workloads
from kups_amd.mcmc.gcmc import make_co2_gcmc, make_mixture_gcmc

pytestmark = pytest.mark.gpu


def oracle_state(g):
    return oracle.gcmc_state(g)


def _small_gcmc(seed=0, n_chain=4, pressure=1e6, nmin=300, nmax=400, gmax=64):
    h = workloads.gcmc_hosts(n_chain, nmin=nmin, nmax=nmax, seed=seed, vacancy=0.95)
    h.pressure[:] = pressure
    return make_co2_gcmc(h, gmax=gmax, seed=seed + 100)


def _compare_with_oracle(g, n_steps):
    """Run g and the oracle from g's state for n_steps trials; compare trial
    by trial.  Returns (compared, accepted, chains that never diverged, the
    oracle's final state)."""
    st = oracle_state(g)
    kT = g.d["kT"].cpu().numpy()
    step0 = g.step
    tr_gpu = g.run(n_steps, n_trace=g.n_chain).cpu().numpy()
    tr_orc = oracle.gcmc_run(st, n_steps, step0, g.seed, n_trace=g.n_chain)
    compared = 0
    accepted = 0
    same = []
    for c in range(g.n_chain):
        lu = [math.log((float(oracle.philox(g.seed, c, step0 + t, 0x47434D41)[1]) + 0.5) * 2.0**-32)
              for t in range(n_steps)]
        for t in range(n_steps):
            a, b = tr_gpu[c, t], tr_orc[c, t]
            assert a[0] == b[0]  # same move type and coin
            if abs(b[1]) <= 50 * kT[c]:
                assert abs(a[1] - b[1]) <= 1e-5 * max(abs(b[1]), 1.0), (c, t, a, b)
            else:
                assert abs(a[1] - b[1]) <= 1e-3 * abs(b[1]), (c, t, a, b)
            compared += 1
            if a[3] != b[3]:
                assert abs(b[2] - lu[t]) < 1e-3, (c, t, a, b, lu[t])
                break
            accepted += int(a[3])
        else:
            same.append(c)
    return compared, accepted, same, st


def test_gcmc_trial_by_trial_vs_oracle(gpu):
    g = _small_gcmc()
    n_steps = 300
    compared, accepted, _, _ = _compare_with_oracle(g, n_steps)
    assert compared > 0.9 * g.n_chain * n_steps
    assert accepted > 20


def _mixture(seed=0, n_chain=4, pressure=3e6, ads=("CO2", "N2", "CH4"), comp=(0.5, 0.3, 0.2)):
    h = workloads.gcmc_hosts(n_chain, nmin=300, nmax=400, seed=seed, vacancy=0.95)
    h.pressure[:] = pressure
    return make_mixture_gcmc(h, ads, comp, gmax=64, seed=seed + 100)


def test_gcmc_mixture_trial_by_trial_vs_oracle(gpu):
    """CO2 / N2 / CH4 (3, 3 and 1 sites): the species drawn per exchange
    trial, the k-th guest of it deleted, any guest displaced, per-species
    fugacity and molecule constants; trial by trial against the oracle from
    a start that holds guests of every species, and the final species table
    of every chain that never met a near tie equal to the oracle's."""
    g = _mixture()
    g.insert_random(np.array([[3, 2, 2]] * g.n_chain), seed=5)
    n_steps = 400
    compared, accepted, same, st = _compare_with_oracle(g, n_steps)
    assert compared > 0.9 * g.n_chain * n_steps
    assert accepted > 20 and len(same) >= 2
    ng = g.loading.cpu().numpy()
    gm = g.d["guest_motif"].cpu().numpy()
    for c in same:
        assert ng[c] == st["n_guest"][c]
        np.testing.assert_array_equal(gm[c, :ng[c]], st["guest_motif"][c, :ng[c]])
    counts = g.motif_counts()
    assert (counts.sum(0) > 0).all(), counts
    # species counts per LJ label follow the molecules
    cnt = g.d["counts"].cpu().numpy()
    exp = np.zeros_like(cnt)
    for t, sp in enumerate(g.guest_species_per_motif):
        for x in sp:
            exp[:, x] += counts[:, t]
    host = np.zeros_like(cnt)
    off = g.d["host_off"].cpu().numpy()
    hs = g.d["host_species"].cpu().numpy()
    for c in range(g.n_chain):
        host[c] = np.bincount(hs[off[c]:off[c + 1]], minlength=cnt.shape[1])
    np.testing.assert_array_equal(cnt, exp + host)


def test_gcmc_mixture_tracked_energy_matches_full_evaluation(gpu):
    g = _mixture(seed=3, n_chain=6)
    g.insert_random(np.array([[2, 2, 3]] * g.n_chain), seed=1)
    nt0 = g.motif_counts()
    e0 = g.d["energy"].cpu().numpy().copy()
    g.run(800)
    tracked = g.d["energy"].cpu().numpy()
    full = g.full_energy() - ((g.motif_counts() - nt0) * g.sr_intramolecular()).sum(1)
    scale = np.maximum(np.abs(tracked - e0), 1.0)
    assert (np.abs(tracked - full) / scale).max() < 1e-4, (tracked, full)
    p = g.guest_pressure()
    assert np.isfinite(p["potential"]).all()


def test_gcmc_mixture_ideal_gas_limit(gpu):
    """No charges and no LJ: each species of a mixture is an independent
    Poisson gas, <N_t> = a_t V, whatever its site count."""
    from kups_amd.mcmc.gcmc import Motif, RigidGCMC
    from kups_amd.potential.lennard_jones import LennardJonesParameters

    n = 64
    L = 25.0
    lat = np.repeat(np.eye(3)[None] * L, n, 0)
    lj = LennardJonesParameters.from_dict(12.0, {"X": (3.0, 0.0)})
    motifs = [Motif.from_dict({**workloads.ADSORBATES[a], "charges": (0.0,) * k})
              for a, k in (("CO2", 3), ("CH4", 1), ("N2", 3))]
    g = RigidGCMC(np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.int32), np.zeros(n, np.int64),
                  lat, 1e5, 298.0, motifs, lj, [(0, 0, 0), (0,), (0, 0, 0)],
                  weights=(0, 0, 0, 1), gmax=128, seed=7)
    aV = np.array([6.0, 4.0, 3.0])
    g.log_activity[:] = np.log(aV / L**3)[None]
    g.d["log_activity"].copy_(torch.as_tensor(g.log_activity, device=gpu))
    samples = []
    for _ in range(60):
        g.run(150)
        samples.append(g.motif_counts())
    s = np.array(samples[10:]).astype(float).reshape(-1, 3)
    np.testing.assert_allclose(s.mean(0), aV, rtol=0.08)
    np.testing.assert_allclose(s.var(0), aV, rtol=0.25)


def test_gcmc_tracked_energy_matches_full_evaluation(gpu):
    """Incremental == full (test_potential_update.py:600-645): after a run the
    tracked energy equals a fresh full evaluation, up to the intramolecular
    real-space term the reference's full evaluation holds for every molecule
    and its incremental updates never add."""
    g = _small_gcmc(seed=1)
    nt0 = g.motif_counts()
    e0 = g.d["energy"].cpu().numpy().copy()
    g.run(600)
    ng = g.loading.cpu().numpy()
    assert ng.sum() > 0
    tracked = g.d["energy"].cpu().numpy()
    full = g.full_energy() - ((g.motif_counts() - nt0) * g.sr_intramolecular()).sum(1)
    scale = np.maximum(np.abs(tracked - e0), 1.0)
    assert (np.abs(tracked - full) / scale).max() < 1e-4, (tracked, full)


def test_gcmc_structure_factor_cache_matches_recomputation(gpu):
    g = _small_gcmc(seed=2)
    g.run(400)
    sk_inc = g.d["sk"].clone()
    g.recompute_state()
    sk_full = g.d["sk"]
    live = (g.d["pk"] != 0)[:, :, None]  # padding and k = 0 carry no energy
    err = ((sk_inc - sk_full).abs() * live).max().item()
    assert err < 1e-4, err


def test_gcmc_ideal_gas_limit(gpu):
    """No charges and no LJ: <N> = a V on the GPU as well."""
    from kups_amd.core.constants import BOLTZMANN_CONSTANT
    from kups_amd.mcmc.gcmc import Motif, RigidGCMC
    from kups_amd.potential.lennard_jones import LennardJonesParameters

    n = 64
    L = 25.0
    lat = np.repeat(np.eye(3)[None] * L, n, 0)
    lj = LennardJonesParameters.from_dict(12.0, {"X": (3.0, 0.0)})
    motif = Motif.from_dict({**workloads.CO2, "charges": (0.0, 0.0, 0.0)})
    g = RigidGCMC(np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.int32), np.zeros(n, np.int64),
                  lat, 1e5, 298.0, motif, lj, (0, 0, 0), weights=(0, 0, 0, 1), gmax=128, seed=3)
    aV = 6.0
    V = L**3
    g.log_activity[:] = math.log(aV / V)
    g.d["log_activity"].copy_(torch.as_tensor(g.log_activity, device=gpu))
    samples = []
    for _ in range(60):
        g.run(100)
        samples.append(g.loading.cpu().numpy().copy())
    s = np.array(samples[10:]).astype(float)
    assert s.mean() == pytest.approx(aV, abs=0.25)
    assert s.var() == pytest.approx(aV, rel=0.2)


def test_gcmc_graph_replays_match_eager_runs(gpu):
    """A cycle of two 40-trial launches recorded in a HIP graph and replayed
    three times, with eager work between the replays, leaves the chains in
    the state the same launches run eagerly produce (same random numbers:
    the recorded kernels read the trial counter from device memory)."""
    g_eager = _small_gcmc(seed=2, n_chain=6)
    g_graph = _small_gcmc(seed=2, n_chain=6)
    cycle = g_graph.capture(40, launches=2)
    keys = ("guest", "n_guest", "counts", "sk", "energy", "width", "hist", "hist_idx", "stats")
    for rep in range(3):
        g_eager.run(40)
        g_eager.run(40)
        cycle.replay()
        torch.cuda.synchronize()
        assert g_graph.step == g_eager.step == 80 * (rep + 1)
        for k in keys:
            assert torch.equal(g_graph.d[k], g_eager.d[k]), (rep, k)
        # unrelated eager work between replays
        (torch.arange(1000, device=gpu, dtype=torch.float64) ** 2).sum().item()
    assert int(g_graph.stats["trials"].sum()) == 6 * 240


def test_mixture_deck_per_species_loading(gpu):
    """A two-adsorbate deck (CO2 / CH4 at mole fractions 0.6 / 0.4, 10 bar,
    RUBTAK 2x2x2) through the YAML application: per-species loading samples
    that sum to the total, both species adsorbed, per-species heats of
    adsorption, and the tracked energies equal to a full re-evaluation."""
    from pathlib import Path

    from kups_amd.application import mcmc_rigid
    from kups_amd.application.config import McmcConfig, load_yaml

    ci = Path(__file__).resolve().parent / "golden" / "statistical" / "inputs" / "ci"
    d = load_yaml(ci / "gcmc_co2_rubtak.yaml")
    d["adsorbates"].append({**workloads.CH4, "positions": [list(p) for p in workloads.CH4["positions"]],
                            "symbols": list(workloads.CH4["symbols"]),
                            "charges": list(workloads.CH4["charges"])})
    d["hosts"][0].update(pressure=1e6, adsorbate_composition=[0.6, 0.4],
                         adsorbate_interaction=[[0.0, 0.0], [0.0, 0.0]], init_adsorbates=[2, 2])
    d["run"].update(num_warmup_cycles=30, num_cycles=60)
    cfg = McmcConfig.from_dict(d, ci)
    r = mcmc_rigid.build(cfg, replicas=8, device=gpu)
    assert r.gcmc.n_motif == 2
    np.testing.assert_array_equal(r.gcmc.motif_counts(), np.full((8, 2), 2))
    mcmc_rigid.run(r)
    per = r.loading_per_species
    assert per.shape == (60, 8, 2)
    np.testing.assert_array_equal(per.sum(-1), r.loading)
    np.testing.assert_array_equal(per[-1], r.gcmc.motif_counts())
    assert (per.mean(axis=(0, 1)) > 0.5).all(), per.mean(axis=(0, 1))
    res = mcmc_rigid.analyze(r)[0]
    assert np.asarray(res.loading.mean).shape == (2,)
    assert np.isfinite(np.asarray(res.heat_of_adsorption.mean)).all()
    assert np.abs(r.gcmc.energy_error()).max() < 1e-2
