"""Measured choice of the Verlet skin of the MD paths.

The skin trades list rebuilds (more frequent for a thin skin) against list
entries evaluated every step (more for a thick one).  The balance depends on
the batch's density, temperature, time step, cutoff and system sizes, so it
is measured rather than modelled: each candidate runs a few steps on the
caller's state, timed with HIP events on the current stream, the state is
put back, and the fastest candidate wins unless it beats the default (the
value the kernels were tuned and profiled at) by less than ``margin``, a
timing-noise guard that keeps the choice stable between runs.  Results are
cached per workload
signature in memory and, when the environment variable KUPS_AMD_TUNE_CACHE
names a JSON file, on disk.
"""

from __future__ import annotations

import hashlib
import json
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

_CACHE: dict[str, dict] = {}


@dataclass
class TuneResult:
    best: float                      # the chosen skin (length units of the state)
    table: dict = field(default_factory=dict)  # skin -> ms per step
    key: str = ""
    cached: bool = False


def workload_key(kind: str, state, extra=()) -> str:
    """Signature of a batch: kind, size distribution, density, cutoff, time
    step, temperature and anything in ``extra``."""
    sizes = np.asarray(state.sizes, np.int64)
    tril = state.tril.double().cpu().numpy()
    vol = np.abs(tril[:, 0] * tril[:, 2] * tril[:, 5])
    rho = float(sizes.sum() / vol.sum()) if vol.sum() > 0 else 0.0
    q = np.quantile(sizes, [0.0, 0.25, 0.5, 0.75, 1.0]).tolist() if sizes.size else []
    parts = [kind, len(sizes), int(sizes.sum()), [round(x) for x in q], round(rho, 5),
             round(float(state.cutoff.max().item()), 4) if len(sizes) else 0.0,
             round(float(state.time_step.max().item()), 6) if len(sizes) else 0.0,
             round(float(state.temperature.max().item()), 3) if len(sizes) else 0.0, list(extra)]
    return hashlib.sha1(json.dumps(parts).encode()).hexdigest()[:16]


def _disk_path() -> Path | None:
    p = os.environ.get("KUPS_AMD_TUNE_CACHE")
    return Path(p) if p else None


def _lookup(key: str) -> dict | None:
    if key in _CACHE:
        return _CACHE[key]
    p = _disk_path()
    if p and p.exists():
        try:
            data = json.loads(p.read_text())
        except (OSError, ValueError):
            return None
        if key in data:
            _CACHE[key] = data[key]
            return data[key]
    return None


def _store(key: str, entry: dict) -> None:
    _CACHE[key] = entry
    p = _disk_path()
    if p:
        data = {}
        if p.exists():
            try:
                data = json.loads(p.read_text())
            except (OSError, ValueError):
                data = {}
        data[key] = entry
        p.write_text(json.dumps(data, indent=1))


_FIELDS = ("positions", "momenta", "forces", "tril")


def _snapshot(state):
    return {k: getattr(state, k).clone() for k in _FIELDS}, state.step


def _restore(state, snap):
    tensors, step = snap
    for k in _FIELDS:
        getattr(state, k).copy_(tensors[k])
    state.step = step


def _time_steps(run, steps: int) -> float:
    """ms per step of ``run(steps)`` on the current stream."""
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    run(steps)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


def _tune(kind: str, state, candidates, make_runner, steps: int, extra=(), prefer=None,
          margin: float = 0.03) -> TuneResult:
    key = workload_key(kind, state, tuple(extra) + tuple(candidates) + (steps, prefer, margin))
    hit = _lookup(key)
    if hit is not None:
        return TuneResult(hit["best"], {float(k): v for k, v in hit["table"].items()}, key, True)
    table = {}
    snap = _snapshot(state)
    for c in candidates:
        run = make_runner(c)
        run(2)                      # sizes lists and scratch outside the timing
        _restore(state, snap)
        table[float(c)] = _time_steps(run, steps)
        _restore(state, snap)
    best = min(table, key=table.get)
    if prefer is not None and float(prefer) in table and \
            table[best] > (1.0 - margin) * table[float(prefer)]:
        best = float(prefer)
    _store(key, {"best": best, "table": {str(k): v for k, v in table.items()}})
    return TuneResult(best, table, key, False)


def tune_md_skin(params, state, integrator: int = 0, seed: int = 0,
                 skins_sigma=(0.15, 0.2, 0.25, 0.3, 0.4), steps: int = 100,
                 default_sigma: float = 0.3, margin: float = 0.03) -> TuneResult:
    """Skin of the LDS MD kernel's Verlet lists (FusedLJPropagator), candidates
    in units of the largest LJ sigma; each runs ``steps`` steps in one launch."""
    from ..md.integrators import FusedLJPropagator

    sig = float(np.max(params.sigma))
    cands = [s * sig for s in skins_sigma]

    def make(skin):
        prop = FusedLJPropagator(params, integrator, seed, skin)
        return lambda n: prop(state, n, record=False)

    prefer = default_sigma * sig if default_sigma in skins_sigma else None
    return _tune("md_lj", state, cands, make, steps, (integrator, sig), prefer, margin)


def tune_ewald_skin(params, ewald, state, integrator: int = 1, seed: int = 0,
                    skins=(0.5, 1.0, 1.5, 2.0), steps: int = 12, default: float = 1.0,
                    margin: float = 0.03) -> TuneResult:
    """Skin (A) of the HBM path's Verlet lists (EwaldLJPropagator)."""
    from ..md.integrators import EwaldLJPropagator

    def make(skin):
        prop = EwaldLJPropagator(params, ewald, integrator, seed=seed, skin=skin)
        return lambda n: prop(state, n)

    return _tune("ewald_lj", state, list(skins), make, steps,
                 (integrator, tuple(np.asarray(ewald.alpha).ravel()[:1].tolist())),
                 default if default in skins else None, margin)
