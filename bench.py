"""Benchmark: batched LJ molecular dynamics on one MI355X.

Workload (the batched-MD configuration of the reference path): 2048
independent cubic PBC boxes, atom counts N ~ U{64..4096} (seed 0, ~4.26 M
atoms), FCC sites at density 0.8 sigma^-3 with 0.05 sigma Gaussian jitter,
argon LJ (sigma 3.405 A, eps/kB 119.8 K, cutoff 2.5 sigma, hard, no shift),
Maxwell-Boltzmann momenta at 94.4 K, NVE velocity Verlet, dt = 2 fs, FP32
positions/forces with FP64 per-system energies, neighbor set rebuilt every
step.  A "step" is one MD step of every system; all timed steps run in one
persistent launch (``--steps-per-launch`` splits them).

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
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--steps-per-launch", type=int, default=0, help="0: all timed steps in one launch")
    ap.add_argument("--n-sys", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-every", type=int, default=8,
                    help="CPU baseline runs every k-th system of the batch")
    ap.add_argument("--cpu-steps", type=int, default=20)
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
    }


def load_traffic():
    f = ROOT / "profiles" / "md_lj_kernel_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text())
        except Exception:
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every configuration is a batch of independent "
              "systems and this build runs no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from kups_amd import workloads
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
    if args.warmup:
        prop(st, args.warmup, record=False)
    torch.cuda.synchronize()
    spl = args.steps_per_launch or args.steps
    launches = []
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(spl, args.steps - done)
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
    value = atoms * args.steps / wall
    # per-launch kernel time (events bracket the order + geometry kernels too,
    # which take a few microseconds)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in launches]))
    steps_per_launch = float(np.mean([k for _, _, k in launches]))
    achieved_tflops = FLOP_PER_ATOM_STEP * atoms * steps_per_launch / (kern_ms * 1e-3) / 1e12
    etot = (obs.potential_energy[:, -1] + obs.kinetic_energy[:, -1]).double().sum().item()
    traffic = load_traffic()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 per-system energies)",
        "data": "synthetic (seeded FCC boxes with Gaussian jitter, Maxwell-Boltzmann momenta)",
        "config": {
            "workload": "batched LJ MD: 2048 cubic PBC boxes, N~U{64..4096} (seed 0), "
                        "rho*=0.8, jitter 0.05 sigma, argon LJ rc=2.5 sigma, NVE velocity "
                        "Verlet dt=2 fs, neighbor set rebuilt every step",
            "n_systems": batch.n_sys,
            "n_atoms": atoms,
            "steps_per_launch": spl,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "kernel": "md_lj_kernel",
            "note": "FP32 FLOP ceiling (157.3 TF/s is the FP32 peak of both the vector "
                    "ALUs and the matrix cores; this kernel is VALU work, state LDS-resident, "
                    f"{FLOP_PER_ATOM_STEP:.0f} algorithmic FLOP per atom-step)",
            "kernel_ms_per_launch": kern_ms,
        },
        "final_total_energy_eV": etot,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(batch, params, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
