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
from kups_amd.mcmc.gcmc import make_co2_gcmc

pytestmark = pytest.mark.gpu


def oracle_state(g):
    return oracle.gcmc_state(g)


def _small_gcmc(seed=0, n_chain=4, pressure=1e6, nmin=300, nmax=400, gmax=64):
    h = workloads.gcmc_hosts(n_chain, nmin=nmin, nmax=nmax, seed=seed, vacancy=0.95)
    h.pressure[:] = pressure
    return make_co2_gcmc(h, gmax=gmax, seed=seed + 100)


def test_gcmc_trial_by_trial_vs_oracle(gpu):
    g = _small_gcmc()
    st = oracle_state(g)
    kT = g.d["kT"].cpu().numpy()
    n_steps = 300
    tr_gpu = g.run(n_steps, n_trace=g.n_chain).cpu().numpy()
    tr_orc = oracle.gcmc_run(st, n_steps, 0, g.seed, n_trace=g.n_chain)
    compared = 0
    accepted = 0
    for c in range(g.n_chain):
        lu = [math.log((float(oracle.philox(g.seed, c, t, 0x47434D41)[1]) + 0.5) * 2.0**-32)
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
    assert compared > 0.9 * g.n_chain * n_steps
    assert accepted > 20


def test_gcmc_tracked_energy_matches_full_evaluation(gpu):
    """Incremental == full (test_potential_update.py:600-645): after a run the
    tracked energy equals a fresh full evaluation, up to the intramolecular
    real-space term the reference's full evaluation holds for every molecule
    and its incremental updates never add."""
    g = _small_gcmc(seed=1)
    ng0 = g.loading.cpu().numpy().copy()
    e0 = g.d["energy"].cpu().numpy().copy()
    g.run(600)
    ng = g.loading.cpu().numpy()
    assert ng.sum() > 0
    tracked = g.d["energy"].cpu().numpy()
    full = g.full_energy() - (ng - ng0) * g.sr_intramolecular()
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
