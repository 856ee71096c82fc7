"""MD observables on device tensors (md/observables.py:17-61,
observables/stress.py:62-190).

The force kernels return the per-system virial W = -sum_pairs (u'(r)/r) d (x) d
(including the strain term of the periodic images), so the Cauchy stress of
the reference's virial theorem, sigma = -(1/V)(sum_i dU/dr_i (x) r_i + h^T dU/dh),
is W / V, and the instantaneous pressure is P = 2K / (3V) + tr(sigma) / 3.
"""

from __future__ import annotations

import torch

from ..core.batch import segment_sum


def particle_kinetic_energy(momenta: torch.Tensor, masses: torch.Tensor) -> torch.Tensor:
    """K_i = |p_i|^2 / (2 m_i) (md/observables.py:17-34)."""
    return 0.5 * (momenta.double() ** 2).sum(-1) / masses.double()


def kinetic_energy(momenta: torch.Tensor, masses: torch.Tensor, system: torch.Tensor,
                   n_sys: int) -> torch.Tensor:
    """Per-system kinetic energy (FP64)."""
    return segment_sum(particle_kinetic_energy(momenta, masses), system, n_sys)


def volume(tril: torch.Tensor) -> torch.Tensor:
    """|det L| of lower-triangular cells, tril = [L00, L10, L11, L20, L21, L22]."""
    return (tril[..., 0] * tril[..., 2] * tril[..., 5]).abs()


def stress_via_virial(virial: torch.Tensor, tril: torch.Tensor) -> torch.Tensor:
    """Cauchy stress (n_sys, 3, 3) from the kernels' virial W: sigma = W / V
    (observables/stress.py:69-82)."""
    w = virial.reshape(-1, 3, 3).double()
    return w / volume(tril).double()[:, None, None]


def instantaneous_pressure(kinetic: torch.Tensor, cauchy_stress: torch.Tensor,
                           vol: torch.Tensor) -> torch.Tensor:
    """P = 2K / (d V) + tr(sigma) / d, d = 3 (md/observables.py:37-61)."""
    d = cauchy_stress.shape[-1]
    tr = torch.diagonal(cauchy_stress, dim1=-2, dim2=-1).sum(-1)
    return 2.0 * kinetic / (d * vol) + tr / d
