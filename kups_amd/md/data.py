"""MD state of a batch of systems (application/md/data.py:148-218).

Momenta are drawn from Maxwell-Boltzmann, p ~ N(0, sqrt(m kT)) per component,
minus the mean momentum of the system (171-178); gradients start at zero
(188), so the first half kick of the first step uses F = 0 unless the caller
evaluates forces first.  Time step, friction, the CSVR and barostat coupling
times, the target pressure and the compressibility are converted from fs,
1/fs, Pa and 1/Pa to internal units (195-215).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .._lib import require_gpu
from ..core.batch import offsets_from_sizes
from ..core.constants import BOLTZMANN_CONSTANT, FEMTO_SECOND, PASCAL


@dataclass
class MDState:
    positions: torch.Tensor  # (n, 3) f32
    momenta: torch.Tensor  # (n, 3) f32
    forces: torch.Tensor  # (n, 3) f32, F at the current positions
    masses: torch.Tensor  # (n,) f32
    species: torch.Tensor  # (n,) int32
    sys_off: torch.Tensor  # (n_sys + 1,) int32
    sizes: np.ndarray  # host copy of the system sizes
    tril: torch.Tensor  # (n_sys, 6) f64
    cutoff: torch.Tensor  # (n_sys,) f64
    time_step: torch.Tensor  # (n_sys,) f64, internal units
    temperature: torch.Tensor  # (n_sys,) f64, K
    friction: torch.Tensor  # (n_sys,) f64, 1 / internal time
    step: int = 0
    # (n_sys,) f64, CSVR coupling time tau, internal units
    thermostat_time_constant: torch.Tensor | None = None
    # (n_sys, 4) f64, stochastic cell rescaling: target pressure (eV/A^3),
    # coupling time (internal units), compressibility (A^3/eV), minimum scale
    barostat: torch.Tensor | None = None
    charges: torch.Tensor | None = None  # (n,) f32, for Ewald systems
    extras: dict = field(default_factory=dict)

    @property
    def n_sys(self) -> int:
        return self.tril.shape[0]

    @property
    def n(self) -> int:
        return self.positions.shape[0]

    @property
    def kT(self) -> torch.Tensor:
        return self.temperature * BOLTZMANN_CONSTANT


def maxwell_boltzmann_momenta(masses: np.ndarray, sizes, temperature, rng: np.random.Generator):
    """p ~ N(0, sqrt(m kT)) minus the per-system mean (md/data.py:171-178)."""
    masses = np.asarray(masses, np.float64)
    temperature = np.broadcast_to(np.asarray(temperature, np.float64), (len(sizes),))
    sys_of = np.repeat(np.arange(len(sizes)), sizes)
    std = np.sqrt(masses * temperature[sys_of] * BOLTZMANN_CONSTANT)
    p = rng.standard_normal((masses.size, 3)) * std[:, None]
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in range(len(sizes)):
        a, b = off[s], off[s + 1]
        if b > a:
            p[a:b] -= p[a:b].sum(axis=0) / (b - a)
    return p


def make_md_state(positions, masses, species, sizes, tril, cutoff, *, time_step_fs,
                  temperature, friction_per_fs=0.0, thermostat_time_constant_fs=100.0,
                  target_pressure_pa=1.0e5, pressure_coupling_time_fs=1000.0,
                  compressibility_per_pa=2.0e-9, minimum_scale_factor=0.9, momenta=None,
                  charges=None, device=None) -> MDState:
    dev = device or require_gpu()
    n_sys = len(sizes)
    n = int(np.sum(sizes))

    def per_sys(x):
        return torch.as_tensor(np.broadcast_to(np.asarray(x, np.float64), (n_sys,)).copy(),
                               dtype=torch.float64, device=dev)

    if momenta is None:
        momenta = np.zeros((n, 3))
    return MDState(
        positions=torch.as_tensor(np.asarray(positions, np.float32), device=dev).contiguous(),
        momenta=torch.as_tensor(np.asarray(momenta, np.float32), device=dev).contiguous(),
        forces=torch.zeros((n, 3), dtype=torch.float32, device=dev),
        masses=torch.as_tensor(np.asarray(masses, np.float32), device=dev).contiguous(),
        species=torch.as_tensor(np.asarray(species, np.int32), device=dev).contiguous(),
        sys_off=offsets_from_sizes(sizes, dev),
        sizes=np.asarray(sizes, np.int64),
        tril=torch.as_tensor(np.asarray(tril, np.float64).reshape(n_sys, 6), device=dev).contiguous(),
        cutoff=per_sys(cutoff),
        time_step=per_sys(np.asarray(time_step_fs, np.float64) * FEMTO_SECOND),
        temperature=per_sys(temperature),
        friction=per_sys(np.asarray(friction_per_fs, np.float64) / FEMTO_SECOND),
        thermostat_time_constant=per_sys(
            np.asarray(thermostat_time_constant_fs, np.float64) * FEMTO_SECOND),
        barostat=torch.stack([
            per_sys(np.asarray(target_pressure_pa, np.float64) * PASCAL),
            per_sys(np.asarray(pressure_coupling_time_fs, np.float64) * FEMTO_SECOND),
            per_sys(np.asarray(compressibility_per_pa, np.float64) / PASCAL),
            per_sys(minimum_scale_factor)], dim=1).contiguous(),
        charges=None if charges is None else
        torch.as_tensor(np.asarray(charges, np.float32), device=dev).contiguous(),
    )
