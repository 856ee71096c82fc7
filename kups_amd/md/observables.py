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


def lattice_gradient_stress(positions: torch.Tensor, forces: torch.Tensor, virial: torch.Tensor,
                            tril: torch.Tensor, system: torch.Tensor) -> torch.Tensor:
    """-(dU/dh) / V with the Cartesian positions held fixed, per system (n_sys, 3, 3).

    This is the stress that the reference's MD application logs
    (application/md/data.py:101-107, observables/stress.py:63-66): the lattice
    vectors enter the energy only through the image shifts, so it is the box
    part of the virial alone and depends on where the atoms sit (positions
    are taken wrapped to [-0.5, 0.5) fractional, unitcell.py:139-171).  From the
    virial theorem h^T dU/dh = -W + sum_i r_i (x) F_i with the kernels' pair
    virial W, so dU/dh = h^-T (sum_i r_i (x) F_i - W).  The physical stress is
    ``stress_via_virial``."""
    n_sys = tril.shape[0]
    L = torch.zeros((n_sys, 3, 3), dtype=torch.float64, device=tril.device)
    L[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]] = tril.double()
    Linv = torch.linalg.inv(L)
    sys = system.long()
    r = positions.double()
    s = torch.einsum("ic,icd->id", r, Linv[sys])
    s = s - torch.floor(s + 0.5)
    rw = torch.einsum("ic,icd->id", s, L[sys])
    rF = segment_sum((rw[:, :, None] * forces.double()[:, None, :]).reshape(-1, 9), system, n_sys)
    G = Linv.transpose(-1, -2) @ (rF.reshape(n_sys, 3, 3) - virial.reshape(n_sys, 3, 3).double())
    return -G / volume(tril).double()[:, None, None]
