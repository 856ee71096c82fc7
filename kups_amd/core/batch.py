"""Batched particle/system containers at the drop-in boundary.

kUPS keeps particles in a keyed Table whose ``system`` column is a foreign
key into the systems table (core/data/table.py, core/data/index.py).  The
kernels take the same data as structure-of-arrays device tensors with the
particles argsorted by system, described by a system-offset table
``sys_off[n_sys + 1]`` (potential/common/graph.py:122-164 sorts the same way).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .._lib import require_gpu


def _dev(t, dtype, device):
    if t is None:
        return None
    return torch.as_tensor(t, dtype=dtype, device=device).contiguous()


@dataclass
class SystemBatch:
    """Per-system unit cells: tril (n_sys, 6) FP64, [L00, L10, L11, L20, L21, L22]
    (TriclinicUnitCell.tril, core/unitcell.py:251-300)."""

    tril: torch.Tensor

    @classmethod
    def from_matrices(cls, lattice, device=None) -> "SystemBatch":
        lat = np.asarray(lattice, dtype=np.float64).reshape(-1, 3, 3)
        if np.any(np.triu(lat, 1) != 0):
            raise ValueError("lattice vectors must be lower triangular (rows a, b, c)")
        tril = lat[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]
        return cls(_dev(tril, torch.float64, device or require_gpu()))

    @property
    def n_sys(self) -> int:
        return self.tril.shape[0]

    @property
    def volume(self) -> torch.Tensor:
        t = self.tril
        return (t[:, 0] * t[:, 2] * t[:, 5]).abs()

    def lattice(self) -> torch.Tensor:
        t = self.tril
        z = torch.zeros_like(t[:, 0])
        return torch.stack(
            [torch.stack([t[:, 0], z, z], -1), torch.stack([t[:, 1], t[:, 2], z], -1),
             torch.stack([t[:, 3], t[:, 4], t[:, 5]], -1)], -2)


@dataclass
class ParticleBatch:
    """Particles of a batch of systems, contiguous per system.

    positions (n, 3) f32; system (n,) int32 sorted ascending; optional
    species (n,) int32, exclusion (n,) int32 (default: the particle id, i.e.
    only self pairs on the minimum image are excluded), masses (n,) f32,
    momenta (n, 3) f32, charges (n,) f32.
    """

    positions: torch.Tensor
    system: torch.Tensor
    species: torch.Tensor | None = None
    exclusion: torch.Tensor | None = None
    masses: torch.Tensor | None = None
    momenta: torch.Tensor | None = None
    charges: torch.Tensor | None = None

    @property
    def n(self) -> int:
        return self.positions.shape[0]


def system_offsets(system: torch.Tensor, n_sys: int) -> tuple[torch.Tensor, np.ndarray]:
    """sys_off (device int32) and host sizes; checks that ``system`` is sorted."""
    sys_h = system.detach().to("cpu", torch.int64).numpy()
    if sys_h.size and (np.any(np.diff(sys_h) < 0) or sys_h[0] < 0 or sys_h[-1] >= n_sys):
        raise ValueError("particles must be sorted by system index in [0, n_sys)")
    sizes = np.bincount(sys_h, minlength=n_sys).astype(np.int64)
    off = np.zeros(n_sys + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    return torch.as_tensor(off.astype(np.int32), device=system.device), sizes


def segment_sum(values: torch.Tensor, system: torch.Tensor, n_sys: int) -> torch.Tensor:
    """Per-system FP64 sums of ``values`` (n, ...) over particles sorted by
    system: differences of one running sum at the system boundaries, found by
    binary search (no atomics, which serialise when millions of particles add
    into a few thousand systems)."""
    dev = values.device
    bounds = torch.arange(n_sys + 1, dtype=system.dtype, device=dev)
    off = torch.searchsorted(system.contiguous(), bounds)
    c = torch.zeros((values.shape[0] + 1,) + tuple(values.shape[1:]), dtype=torch.float64,
                    device=dev)
    torch.cumsum(values.double(), 0, out=c[1:])
    return c[off[1:]] - c[off[:-1]]


def offsets_from_sizes(sizes, device) -> torch.Tensor:
    sizes = np.asarray(sizes, np.int64)
    off = np.zeros(len(sizes) + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    return torch.as_tensor(off.astype(np.int32), device=device)
