"""Benchmark on one MI355X: batched LJ+Ewald MD (the headline), batched LJ MD
and learned-proposal GCMC.

Headline (``value``, ``ms_per_step``, ``config``, ``roofline``,
``cpu_baseline``): the charged-melt configuration, 512 NaCl-like melts of
N = 8 m^3 ions (m ~ U{5..10}, seed 0, ~2.1 M ions), rock-salt a = 5.64 A with
0.1 A jitter, charges +-1, LJ sigma 2.5 / 4.4 A (eps 0.01 eV), real-space
cutoff 12 A, Ewald alpha = 0.3 / A, k grid |n| <= 8 per axis (2600 nonzero
k), BAOAB Langevin at 1200 K (gamma 1/ps, dt 1 fs); real space over a Verlet
list (skin autotuned), k-space structure factors on the matrix cores.  A
"step" is one MD step of every melt (one force evaluation).  ``--steps`` /
``--warmup`` apply to it.

``batched_lj_md`` (``--no-lj`` skips it): 2048 independent cubic PBC boxes,
N ~ U{64..4096} (seed 0, ~4.36 M atoms), FCC sites at density 0.8 sigma^-3
with 0.05 sigma jitter, argon LJ (cutoff 2.5 sigma), Maxwell-Boltzmann
momenta at 94.4 K, NVE velocity Verlet, dt = 2 fs; all ``--lj-steps`` steps in
one persistent launch.

``gcmc`` (``--no-gcmc`` skips it): the learned-proposal workload, 4096 rigid
hosts of 1500-3000 atoms (jittered cubic lattice, 30 % vacancies, 25-30 A
triclinic cells, charges N(0, 0.3) neutralised), CO2 at 298 K and 0.1-10 bar,
LJ + Ewald; 10^6 insertion samples from the plain kernel train the MLP
proposal on the GPU; after equilibration with it, the uniform and the
learned chains run the same number of trials from the same state.

Prints one JSON line (rank 0, N = 1 only).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "atom-steps/s (batched LJ+Ewald MD) + GCMC accepted moves/s, learned vs plain"
# algorithmic work per unit (conventions of the hot-path measurement table)
# Algorithmic work per atom-step of this workload: 30 FLOP per unique in-cutoff
# pair (26.18 unique pairs per atom at rho* = 0.8, rc = 2.5 sigma) + 12 FLOP of
# integration; 48 B of FP32 position and momentum read + write.
FLOP_PER_ATOM_STEP = 797.0
BYTES_PER_ATOM_STEP = 48.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed LJ+Ewald steps (headline; enough to average over the list "
                         "rebuilds, which come in bursts after the first build)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed LJ+Ewald steps")
    ap.add_argument("--no-lj", action="store_true")
    ap.add_argument("--lj-steps", type=int, default=500)
    ap.add_argument("--lj-warmup", type=int, default=20)
    ap.add_argument("--steps-per-launch", type=int, default=0,
                    help="LJ leg: 0 runs all timed steps in one launch")
    ap.add_argument("--n-sys", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-every", type=int, default=8,
                    help="CPU baseline runs every k-th system of the batch")
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--skin-sigma", type=float, default=None,
                    help="Verlet skin of the MD leg in sigma (default: autotuned; the traffic "
                         "probe defaults to 0.3)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="start with a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="1 GiB device copy (counter calibration) + one MD launch, no JSON line")
    ap.add_argument("--ewald-cpu-systems", type=int, default=16)
    ap.add_argument("--no-gcmc", action="store_true")
    ap.add_argument("--gcmc-chains", type=int, default=4096)
    ap.add_argument("--gcmc-samples", type=int, default=1_000_000)
    ap.add_argument("--gcmc-equil", type=int, default=20000, help="learned equilibration trials")
    ap.add_argument("--gcmc-steps", type=int, default=20000, help="timed trials per chain and proposal")
    ap.add_argument("--gcmc-chunk", type=int, default=500, help="trials per launch (loading samples)")
    ap.add_argument("--gcmc-cpu-chains", type=int, default=64)
    ap.add_argument("--gcmc-cpu-steps", type=int, default=25000)
    return ap.parse_args()


def cpu_baseline(batch, params, args):
    """FP32 velocity-Verlet LJ on the host (the oracle's cell-list port, OpenMP
    over systems), on a bounded sample of the same workload."""
    import oracle
    from kups_amd import workloads
    from kups_amd.core.constants import FEMTO_SECOND
    from kups_amd.md.data import maxwell_boltzmann_momenta

    idx = np.arange(0, batch.n_sys, args.cpu_sample_every)
    idx = np.array([s for s in idx if batch.sizes[s] >= 100])  # the port needs no image grid
    sizes = batch.sizes[idx]
    off = batch.sys_off
    pos = np.concatenate([batch.positions[off[s]:off[s + 1]] for s in idx]).astype(np.float32)
    masses = np.full(len(pos), workloads.AR_MASS, np.float32)
    mom = maxwell_boltzmann_momenta(masses, sizes, 94.4, np.random.default_rng(1)).astype(np.float32)
    sys_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    tril = batch.tril[idx]
    cut = np.full(len(idx), workloads.AR_CUTOFF)
    dt = np.full(len(idx), 2.0 * FEMTO_SECOND)
    frc = np.zeros_like(pos)
    species = np.zeros(len(pos), np.int32)
    sig = params.sigma.astype(np.float32)
    eps = params.epsilon.astype(np.float32)
    oracle.md_vv_f32(1, pos, mom, masses, species, sys_off, tril, cut, sig, eps, dt, frc)
    t0 = time.perf_counter()
    oracle.md_vv_f32(args.cpu_steps, pos, mom, masses, species, sys_off, tril, cut, sig, eps, dt, frc)
    el = time.perf_counter() - t0
    atoms = int(sizes.sum())
    return {
        "value": atoms * args.cpu_steps / el,
        "unit": "atom-steps/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"{len(idx)} of {batch.n_sys} systems (every {args.cpu_sample_every}th, N>=100), "
                   f"{atoms} atoms x {args.cpu_steps} FP32 VV steps, cell list rebuilt every step, "
                   f"OpenMP over systems, {el:.1f} s; CPU restatement of kUPS, not kUPS"),
        "cpu": cpu_info(),
    }


# LJ + Ewald ion-step of the charged melts: 45 FLOP per unique real-space pair
# (161 per ion within 12 A) + 36 FLOP per k vector (S(k) 16 + force 20, 2600 k)
FLOP_PER_ION_STEP = 45.0 * 161 + 36.0 * 2600
NACL = {"Na": (2.5, 0.01), "Cl": (4.4, 0.01)}


def ewald_cpu_baseline(b, args):
    """One force evaluation (cell-list LJ, the dense O(N^2) neighbor list the
    reference's applications use, real-space and k-space Ewald, self term) of
    the oracle on the first systems of the batch, FP64, OpenMP over systems;
    the integrator is negligible beside it."""
    import oracle
    from kups_amd.potential.lennard_jones import LennardJonesParameters

    k = args.ewald_cpu_systems
    n = int(b.sys_off[k])
    off = b.sys_off[:k + 1]
    tril = b.tril[:k]
    pos = b.positions[:n].astype(np.float32).astype(np.float64)
    sp, q = b.species[:n], b.charges[:n]
    rc, al = np.full(k, 12.0), np.full(k, 0.3)
    lj = LennardJonesParameters.from_dict(12.0, NACL)
    kv = np.stack([oracle.kvecs(t, 8) for t in tril])
    t0 = time.perf_counter()
    oracle.lj_cells(pos, sp, off, tril, rc, lj.sigma, lj.epsilon)
    ij, sh = oracle.neighborlist(pos, np.arange(n), off, tril, rc)
    oracle.ewald_real_edges(pos, q, off, tril, al, rc, ij, sh)
    oracle.ewald_recip(pos, q, off, tril, al, kv)
    oracle.ewald_self(q, off, al)
    el = time.perf_counter() - t0
    return {
        "value": n / el, "unit": "atom-steps/s", "cores": oracle.num_threads(), "kind": "port",
        "sample": f"first {k} of {b.n_sys} melts ({n} ions), one FP64 force evaluation (LJ over "
                  f"a cell list, real-space Ewald over the dense neighbor list the reference's "
                  f"applications use, k-space over 2600 k, self term), {el:.1f} s; the "
                  f"integrator is negligible beside it; CPU restatement of kUPS, not kUPS",
        "cpu": cpu_info(),
    }


# real-space part of an ion-step: 45 FLOP per unique pair within 12 A (161
# per ion at the rock-salt density 8 / 5.64^3 A^-3): the algorithmic work of
# vl_pair_kernel, the dominant kernel of the step
FLOP_PAIR_PER_ION_STEP = 45.0 * 161


def cpu_info() -> dict:
    """Host CPU model and thread counts (lscpu when present)."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Thread(s) per core", "Socket(s)", "Core(s) per socket"):
                info[k.lower().replace(" ", "_").replace("(s)", "s")] = v
    except Exception:
        pass
    return info


def load_profile(name: str) -> dict | None:
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text())
    except Exception:
        return None


def ewald_leg(args):
    """The headline: LJ + Ewald MD of the charged melts."""
    import torch

    from kups_amd import workloads
    from kups_amd.core.autotune import tune_ewald_skin
    from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
    from kups_amd.md.integrators import BAOAB_LANGEVIN, EwaldLJPropagator
    from kups_amd.potential.ewald import EwaldParameters
    from kups_amd.potential.lennard_jones import LennardJonesParameters

    b = workloads.nacl_melts(n_sys=512, seed=0)
    lj = LennardJonesParameters.from_dict(12.0, NACL)
    p0 = maxwell_boltzmann_momenta(b.masses, b.sizes, 1200.0, np.random.default_rng(0))
    st = make_md_state(b.positions.astype(np.float32), b.masses, b.species, b.sizes, b.tril, 12.0,
                       time_step_fs=1.0, temperature=1200.0, momenta=p0.astype(np.float32),
                       charges=b.charges, friction_per_fs=0.001)
    ep = EwaldParameters.box(b.n_sys, alpha=0.3, cutoff=12.0, nmax=8)
    EwaldLJPropagator(lj, ep, BAOAB_LANGEVIN).evaluate_forces(st)
    tuned = tune_ewald_skin(lj, ep, st, BAOAB_LANGEVIN)
    prop = EwaldLJPropagator(lj, ep, BAOAB_LANGEVIN, skin=tuned.best)
    prop.evaluate_forces(st)
    if args.warmup:
        prop(st, args.warmup)
    torch.cuda.synchronize()
    prop.verlet.profile(True)  # HIP events around the main pair pass of every step
    t0 = time.perf_counter()
    obs = prop(st, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pair_ms, pair_calls = prop.verlet.profile(False)
    ions = int(b.sizes.sum())
    rebuilds, overflow = prop.verlet.counters()
    value = ions * args.steps / wall
    pair_avg_ms = pair_ms / max(pair_calls, 1)
    # the same kernel with the k-space kernels serialised after it (not
    # sharing the GPU), a few untimed-for-value steps
    prop.overlap = False
    prop.verlet.profile(True)
    prop(st, 5)
    torch.cuda.synchronize()
    iso_ms, iso_calls = prop.verlet.profile(False)
    prop.overlap = True
    iso_avg_ms = iso_ms / max(iso_calls, 1)
    achieved = FLOP_PAIR_PER_ION_STEP * ions / (pair_avg_ms * 1e-3) / 1e12
    step_tflops = FLOP_PER_ION_STEP * value / 1e12
    traffic = load_profile("vl_pair_kernel_traffic.json")
    res = {
        "value": value, "ms_per_step": wall * 1e3 / args.steps,
        "config": {
            "workload": f"LJ+Ewald MD: {b.description}, rock-salt a=5.64 A, jitter 0.1 A, "
                        "charges +-1, LJ sigma 2.5/4.4 A eps 0.01 eV (LB mixing), rc 12 A, "
                        "Ewald alpha 0.3/A, |n| <= 8 per axis (2600 k), BAOAB 1200 K, "
                        f"gamma 1/ps, dt 1 fs, Verlet skin {tuned.best:g} A (autotuned), the "
                        "exact cutoff pair set every step",
            "n_systems": b.n_sys, "n_ions": ions, "seed": 0,
        },
        "roofline": {
            "bound": "valu", "kernel": "vl_pair_kernel (main pass)",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "kernel_ms_per_launch": pair_avg_ms, "launches_timed": pair_calls,
            "isolated": {
                "kernel_ms_per_launch": iso_avg_ms, "launches": iso_calls,
                "achieved": FLOP_PAIR_PER_ION_STEP * ions / (iso_avg_ms * 1e-3) / 1e12,
                "frac": FLOP_PAIR_PER_ION_STEP * ions / (iso_avg_ms * 1e-3) / 1e12
                        / PEAK_FP32_TFLOPS,
            },
            "note": f"{FLOP_PAIR_PER_ION_STEP:.0f} algorithmic FLOP per ion per launch (45 per "
                    "unique real-space pair, 161 pairs); kernel time from HIP events around each "
                    "launch on the library's stream over the timed steps, where the k-space "
                    "kernels run concurrently on a second stream; 'isolated': 5 further steps "
                    "with the k-space kernels serialised after it; peak = FP32 vector peak (its "
                    "packed-FP32 half is not used)",
        },
        "step_roofline": {
            "achieved": step_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": step_tflops / PEAK_FP32_TFLOPS,
            "note": f"{FLOP_PER_ION_STEP:.0f} algorithmic FLOP per ion-step (real space + "
                    "2600 k vectors), whole step",
        },
        "autotune": {"skin_A": tuned.best, "ms_per_step": tuned.table, "cached": tuned.cached},
        "verlet": {"system_rebuilds_total": rebuilds, "overflowed_lists": overflow},
        "final_potential_energy_eV": float(obs.potential_energy[:, -1].sum().item()),
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = ewald_cpu_baseline(b, args)
    del prop, st
    torch.cuda.empty_cache()
    return res


# GCMC trial work (the hot-path measurement table): per site evaluation (a
# guest site of the old or the new configuration) 45 FLOP per host atom
# within 12 A plus 16 FLOP per k vector (structure-factor change); 6 FLOP per
# k vector per trial for the long-range energy change
GCMC_FLOP_PER_PAIR = 45.0
GCMC_FLOP_PER_SITE_K = 16.0
GCMC_FLOP_PER_TRIAL_K = 6.0
SPHERE_12A = 4.0 / 3.0 * np.pi * 12.0**3

# MLP 32 -> 128 -> 128 -> 128 -> 1: multiply-adds per descriptor row
MLP_FLOP_PER_ROW = 2.0 * (32 * 128 + 128 * 128 + 128 * 128 + 128)
PEAK_BF16_TFLOPS = 2500.0


def gcmc_cpu_baseline(g, snap, args):
    """The oracle's FP64 GCMC (OpenMP over chains) on a sample of the
    equilibrated chains, uniform insertion, trials/s."""
    import oracle

    g.restore(snap)
    idx = np.linspace(0, g.n_chain - 1, min(args.gcmc_cpu_chains, g.n_chain)).astype(int)
    st = oracle.gcmc_state(g, idx)
    t0 = time.perf_counter()
    oracle.gcmc_run(st, args.gcmc_cpu_steps, g.step, g.seed)
    el = time.perf_counter() - t0
    trials = len(idx) * args.gcmc_cpu_steps
    return {
        "value": trials / el, "unit": "trials/s", "cores": oracle.num_threads(), "kind": "port",
        "sample": f"{len(idx)} of {g.n_chain} equilibrated chains x {args.gcmc_cpu_steps} trials, "
                  f"uniform insertion, FP64, OpenMP over chains, {el:.1f} s; CPU restatement of "
                  "kUPS, not kUPS",
        "cpu": cpu_info(),
    }


def gcmc_leg(args):
    import torch

    from kups_amd import workloads
    from kups_amd.mcmc.gcmc import make_co2_gcmc
    from kups_amd.mcmc.proposal import InsertionSamples, LearnedProposal

    sync = torch.cuda.synchronize
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    hosts = workloads.gcmc_hosts(args.gcmc_chains, seed=0)
    g = make_co2_gcmc(hosts, gmax=64, seed=1)
    lp = LearnedProposal(g)
    sync()
    setup_s = time.perf_counter() - t0
    # 1. training set from the plain kernel
    samples = InsertionSamples(args.gcmc_samples, lp.nv, g.dev)
    t0 = time.perf_counter()
    while samples.n < args.gcmc_samples:
        g.run(200, samples=samples)
    sync()
    collect_s = time.perf_counter() - t0
    # 2. training (GPU forward/backward + Adam) and the proposal table
    t0 = time.perf_counter()
    hist = lp.train(samples)
    sync()
    train_s = time.perf_counter() - t0
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    logits = lp.infer()
    ev1.record(stream)
    sync()
    infer_ms = ev0.elapsed_time(ev1)
    del logits
    t0 = time.perf_counter()
    lp.build_table()
    sync()
    table_s = time.perf_counter() - t0
    # 3. equilibrate with the learned proposal, then run both from that state
    for _ in range(0, args.gcmc_equil, 1000):
        g.run(1000, proposal=lp)
    sync()
    snap = g.snapshot()
    out = {}
    loads = {}
    for name, prop in (("uniform", None), ("learned", lp)):
        g.restore(snap)
        s0 = g.stats
        acc = torch.zeros(g.n_chain, dtype=torch.float64, device=g.dev)
        nchunk = max(1, args.gcmc_steps // args.gcmc_chunk)
        sync()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        kev = []  # events around each gcmc_kernel launch
        e0.record(stream)
        for _ in range(nchunk):
            a0 = torch.cuda.Event(enable_timing=True)
            a1 = torch.cuda.Event(enable_timing=True)
            a0.record(stream)
            g.run(args.gcmc_chunk, proposal=prop)
            a1.record(stream)
            kev.append((a0, a1))
            acc += g.loading
        e1.record(stream)
        sync()
        el = e0.elapsed_time(e1) * 1e-3
        kern_s = sum(a.elapsed_time(b) for a, b in kev) * 1e-3
        s1 = g.stats
        work = s1["work_trials"] - s0["work_trials"]
        sites = s1["site_evaluations"] - s0["site_evaluations"]
        rho = g.sizes / g.volume
        flop = float(np.sum(sites * (rho * SPHERE_12A * GCMC_FLOP_PER_PAIR +
                                     g.k_counts * GCMC_FLOP_PER_SITE_K) +
                            work * g.k_counts * GCMC_FLOP_PER_TRIAL_K))
        trials = int((s1["trials"] - s0["trials"]).sum())
        moved = int((s1["accepted"] - s0["accepted"]).sum())
        ins = int((s1["insertions"] - s0["insertions"]).sum())
        dele = int((s1["deletions"] - s0["deletions"]).sum())
        loads[name] = (acc / nchunk).cpu().numpy()
        err = np.abs(g.energy_error())
        energy = np.abs(g.d["energy"].cpu().numpy())
        out[name] = {
            "trials_per_s": trials / el,
            "accepted_moves_per_s": moved / el,
            "accepted_insertions_per_s": ins / el,
            "accepted_insertions": ins, "accepted_deletions": dele,
            "seconds": el,
            "loading_mean": float(loads[name].mean()),
            "loading_sem": float(loads[name].std(ddof=1) / np.sqrt(g.n_chain)),
            "energy_error_max_eV": float(err.max()),
            "energy_error_max_rel": float((err / np.maximum(energy, 1.0)).max()),
            "work_trials": int(work.sum()),
            "noop_trials": trials - int(work.sum()),
            "work_trials_per_s": int(work.sum()) / el,
            "roofline": {
                "bound": "valu", "kernel": "gcmc_kernel",
                "achieved": flop / kern_s / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": flop / kern_s / 1e12 / PEAK_FP32_TFLOPS, "traffic": None,
                "kernel_ms_per_launch": kern_s * 1e3 / len(kev),
                "note": "algorithmic FLOP of the trials that evaluated sites (45 per host "
                        "atom within 12 A and 16 per k vector per site evaluation, 6 per k "
                        "vector per trial); trials on empty chains that evaluate nothing "
                        "(noop_trials) count no work; host LDS-resident",
            },
        }
    d = loads["learned"] - loads["uniform"]
    z = float(d.mean() / (d.std(ddof=1) / np.sqrt(len(d)))) if d.std() > 0 else 0.0
    u, l = out["uniform"], out["learned"]
    res = {
        "workload": f"learned-proposal GCMC: {hosts.description}; CO2 (3 sites, q C +0.70 / "
                    "O -0.35), 298 K, 0.1-10 bar log-uniform; LJ (TraPPE, LB mixing, rc 12 A, "
                    "tail) + Ewald (rc 12 A, precision 1e-6); moves translation/rotation/"
                    "reinsertion/exchange weights 1/1/1/3",
        "n_chains": g.n_chain, "nk_max": g.nk,
        "trials_per_chain": nchunk * args.gcmc_chunk,
        "equilibration_trials": args.gcmc_equil,
        "uniform": u, "learned": l,
        "insertion_rate_ratio": l["accepted_insertions_per_s"] / max(u["accepted_insertions_per_s"], 1e-30),
        "loading_paired_z": z,
        "training": {"samples": samples.n, "epochs": len(hist), "loss": hist, "seconds": train_s,
                     "collect_seconds": collect_s, "table_seconds": table_s,
                     "voxels_per_chain": lp.nvox, "temperature": lp.temperature, "mix": lp.mix},
        "setup_seconds": setup_s,
        "mlp_roofline": {
            "kernel": "ml_infer_kernel", "bound": "mfma", "rows": int(lp.desc.shape[0]),
            "ms": infer_ms,
            "achieved": MLP_FLOP_PER_ROW * lp.desc.shape[0] / (infer_ms * 1e-3) / 1e12,
            "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
        },
    }
    res["mlp_roofline"]["frac"] = res["mlp_roofline"]["achieved"] / PEAK_BF16_TFLOPS
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = gcmc_cpu_baseline(g, snap, args)
    return res


def load_traffic(skin_sigma: float):
    """Measured HBM bytes of the LJ leg's launch at this skin (the profile
    holds one entry per autotuner candidate), or None."""
    data = load_profile("md_lj_kernel_traffic.json")
    if data is None:
        return None
    return data.get("by_skin_sigma", {}).get(f"{skin_sigma:.2f}")


def lj_leg(args, traffic_probe=False):
    """Batched LJ MD of 2048 boxes, all timed steps in one persistent launch."""
    import torch

    from kups_amd import workloads
    from kups_amd.core.autotune import tune_md_skin
    from kups_amd.md.data import make_md_state, maxwell_boltzmann_momenta
    from kups_amd.md.integrators import make_velocity_verlet_step
    from kups_amd.potential.lennard_jones import LennardJonesParameters

    params = LennardJonesParameters.from_dict(
        workloads.AR_CUTOFF, {"Ar": (workloads.AR_SIGMA, workloads.AR_EPS)})
    batch = workloads.batched_lj_boxes(n_sys=args.n_sys, seed=0)
    p0 = maxwell_boltzmann_momenta(batch.masses, batch.sizes, 94.4, np.random.default_rng(0))
    st = make_md_state(batch.positions.astype(np.float32), batch.masses, batch.species,
                       batch.sizes, batch.tril, workloads.AR_CUTOFF, time_step_fs=2.0,
                       temperature=94.4, momenta=p0.astype(np.float32))
    prop = make_velocity_verlet_step(params)
    prop.evaluate_forces(st)
    tuned = None
    if args.skin_sigma is not None or traffic_probe:
        skin_sigma = 0.3 if args.skin_sigma is None else args.skin_sigma
    else:
        # skin chosen by measurement (a few hundred steps on this state,
        # restored afterwards)
        tuned = tune_md_skin(params, st)
        skin_sigma = tuned.best / workloads.AR_SIGMA
    prop = make_velocity_verlet_step(params, skin=skin_sigma * workloads.AR_SIGMA)
    steps = args.lj_steps
    if traffic_probe:
        prop(st, steps, record=True)
        torch.cuda.synchronize()
        print(f"traffic probe: {int(batch.sizes.sum())} atoms, {steps} steps in one launch, "
              f"skin {skin_sigma:.2f} sigma", flush=True)
        return None
    if args.lj_warmup:
        prop(st, args.lj_warmup, record=False)
    torch.cuda.synchronize()
    spl = args.steps_per_launch or steps
    launches = []
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(spl, steps - done)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        obs = prop(st, k, record=True)
        ev1.record(stream)
        launches.append((ev0, ev1, k))
        done += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    atoms = int(batch.sizes.sum())
    value = atoms * steps / wall
    # per-launch kernel time (events bracket the order + geometry kernels too,
    # which take a few microseconds)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in launches]))
    steps_per_launch = float(np.mean([k for _, _, k in launches]))
    achieved_tflops = FLOP_PER_ATOM_STEP * atoms * steps_per_launch / (kern_ms * 1e-3) / 1e12
    etot = (obs.potential_energy[:, -1] + obs.kinetic_energy[:, -1]).double().sum().item()
    traffic = load_traffic(skin_sigma) if spl == 500 else None
    res = {
        "metric": "atom-steps/s (batched LJ MD)", "value": value, "unit": "atom-steps/s",
        "steps": steps, "warmup": args.lj_warmup, "ms_per_step": wall * 1e3 / steps,
        "workload": "batched LJ MD: 2048 cubic PBC boxes, N~U{64..4096} (seed 0), rho*=0.8, "
                    "jitter 0.05 sigma, argon LJ rc=2.5 sigma, NVE velocity Verlet dt=2 fs, the "
                    "exact cutoff pair set every step (Verlet lists, "
                    f"skin {skin_sigma:.2f} sigma{' autotuned' if tuned else ''}, rebuilt on the "
                    "skin/2 displacement bound)",
        "n_systems": batch.n_sys, "n_atoms": atoms, "steps_per_launch": spl,
        "roofline": {
            "bound": "valu", "kernel": "md_lj_kernel",
            "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "kernel_ms_per_launch": kern_ms,
            "note": f"{FLOP_PER_ATOM_STEP:.0f} algorithmic FLOP per atom-step; state "
                    "LDS-resident, Verlet lists in HBM; traffic from counters of a 500-step "
                    "launch (null for other launch lengths)",
        },
        "final_total_energy_eV": etot,
    }
    if tuned is not None:
        res["autotune"] = {"skin_sigma": skin_sigma,
                           "ms_per_step": {f"{k / workloads.AR_SIGMA:.2f}": v
                                           for k, v in tuned.table.items()},
                           "cached": tuned.cached}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(batch, params, args)
    del st, prop
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every configuration is a batch of independent "
              "systems, each resident on one GPU, and there is no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.traffic_probe or args.copy_calibration:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
        # calibrate the counters against 1 GiB read + 1 GiB written
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.traffic_probe:
        lj_leg(args, traffic_probe=True)
        return
    head = ewald_leg(args)
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (pairs, k-space MFMA; fp64 energies, S(k) and self terms)",
        "data": "synthetic (seeded jittered rock-salt melts, Maxwell-Boltzmann momenta)",
        "config": head["config"],
        "roofline": head["roofline"],
    }
    if "cpu_baseline" in head:
        line["cpu_baseline"] = head["cpu_baseline"]
    for k in ("step_roofline", "autotune", "verlet", "final_potential_energy_eV"):
        line[k] = head[k]
    if not args.no_lj:
        line["batched_lj_md"] = lj_leg(args)
    if not args.no_gcmc:
        line["gcmc"] = gcmc_leg(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
