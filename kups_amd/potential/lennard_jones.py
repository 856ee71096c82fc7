"""Lennard-Jones potential (potential/classical/lennard_jones.py).

``LennardJonesParameters.from_dict`` keeps the reference's YAML schema and
Lorentz-Berthelot mixing (98-141); ``LennardJonesPotential`` returns what the
reference's PotentialFromEnergy returns for it (energy/common/energy.py:263-355):
per-system energies, position gradients (= -forces) and the virial, computed
by the HIP kernel ``kupsmi_lj_forces``.  The tail corrections (276-340) are
closed forms on per-system species counts and are evaluated on the host.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.batch import ParticleBatch, SystemBatch, system_offsets


@dataclass
class LennardJonesParameters:
    """labels (n_species), sigma/epsilon (n_species, n_species) mixed tables,
    cutoff per system (lennard_jones.py:84-96)."""

    labels: tuple[str, ...]
    sigma: np.ndarray
    epsilon: np.ndarray
    cutoff: float

    @classmethod
    def from_dict(cls, cutoff: float, parameters: dict, mixing_rule: str = "lorentz_berthelot"):
        """``parameters`` maps label -> (sigma [A], epsilon [eV]); None -> sigma 1,
        epsilon 0 (lennard_jones.py:98-123)."""
        if mixing_rule != "lorentz_berthelot":
            raise ValueError(f"unknown mixing rule {mixing_rule!r}")
        labels = tuple(parameters.keys())
        raw = np.array([(s or 1.0, e or 0.0) for s, e in parameters.values()], dtype=np.float64)
        sigma, eps = raw[:, 0], raw[:, 1]
        return cls(labels, (sigma[:, None] + sigma) / 2, np.sqrt(eps[:, None] * eps), float(cutoff))

    @property
    def n_species(self) -> int:
        return len(self.labels)


def global_tail_correction_energy(params: LennardJonesParameters, counts: np.ndarray,
                                  volume: np.ndarray) -> np.ndarray:
    """E_tail = sum_ab (8/3) pi N_a N_b / V eps sigma^3 [(sigma/rc)^9 / 3 - (sigma/rc)^3]
    for pairs with eps > 0 (lennard_jones.py:276-313).  counts (n_sys, n_species)."""
    counts = np.asarray(counts, np.float64)
    vol = np.asarray(volume, np.float64)[:, None, None]
    dens = counts[:, :, None] * counts[:, None, :] / vol
    t1 = (params.sigma / params.cutoff) ** 3
    res = (8 / 3) * math.pi * dens * params.epsilon * params.sigma**3 * (t1**3 / 3 - t1)
    return (res * (params.epsilon > 0)).sum(axis=(1, 2))


def global_tail_correction_pressure(params: LennardJonesParameters, counts: np.ndarray,
                                    volume: np.ndarray) -> np.ndarray:
    """Pressure analogue (lennard_jones.py:316-340)."""
    counts = np.asarray(counts, np.float64)
    vol = np.asarray(volume, np.float64)[:, None, None]
    dens = counts[:, :, None] * counts[:, None, :] / vol
    t1 = (params.sigma / params.cutoff) ** 3
    res = (16 / 3) * math.pi * dens / vol * params.epsilon * params.sigma**3 * (t1**3 / 3 * 2 - t1)
    return (res * (params.epsilon > 0)).sum(axis=(1, 2))


@dataclass
class PotentialOut:
    """Per-system energies, position gradients (dE/dr) and virial (W, stress = W/V)."""

    total_energies: torch.Tensor
    position_gradients: torch.Tensor
    virial: torch.Tensor


class LennardJonesPotential:
    """LJ potential over a particle batch (Potential protocol,
    core/potential.py:186-198).  Systems up to 4096 atoms (one LDS tile each)."""

    def __init__(self, params: LennardJonesParameters):
        self.params = params
        self._tables = None

    def tables(self, device):
        if self._tables is None or self._tables[0].device != device:
            self._tables = (
                torch.as_tensor(self.params.sigma, dtype=torch.float64, device=device).contiguous(),
                torch.as_tensor(self.params.epsilon, dtype=torch.float64, device=device).contiguous(),
            )
        return self._tables

    def __call__(self, particles: ParticleBatch, systems: SystemBatch, cutoffs=None) -> PotentialOut:
        dev = particles.positions.device
        n_sys = systems.n_sys
        sys_off, sizes = system_offsets(particles.system, n_sys)
        if sizes.size and sizes.max() > 4096:
            raise ValueError("LennardJonesPotential holds systems of at most 4096 atoms")
        sig, eps = self.tables(dev)
        if cutoffs is None:
            cutoffs = torch.full((n_sys,), self.params.cutoff, dtype=torch.float64, device=dev)
        species = particles.species
        if species is None:
            species = torch.zeros(particles.n, dtype=torch.int32, device=dev)
        # converted inputs are bound to names so they outlive the launch
        pos = particles.positions.to(torch.float32).contiguous()
        species = species.to(torch.int32).contiguous()
        cutoffs = torch.as_tensor(cutoffs, dtype=torch.float64, device=dev).contiguous()
        forces = torch.empty((particles.n, 3), dtype=torch.float32, device=dev)
        energy = torch.empty(n_sys, dtype=torch.float64, device=dev)
        virial = torch.empty((n_sys, 3, 3), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().kupsmi_lj_forces(
            _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), particles.n,
            _lib.ptr(pos), _lib.ptr(species),
            self.params.n_species, _lib.ptr(sig), _lib.ptr(eps), _lib.ptr(systems.tril),
            _lib.ptr(cutoffs), _lib.ptr(forces), _lib.ptr(energy), _lib.ptr(virial)))
        return PotentialOut(energy, -forces, virial)


class PairTailCorrectedLennardJonesPotential(LennardJonesPotential):
    """LJ with the smooth pairwise switch of
    ``make_pair_tail_corrected_lennard_jones_potential``
    (potential/classical/lennard_jones.py:182-222, 516-540): pair energies
    between ``truncation_radius`` and the cutoff are multiplied by
    S = (rc^2 - r^2)^2 (2 r^2 + rc^2 - 3 r_tr^2) / (rc^2 - r_tr^2)^3, so energy
    and force go to zero smoothly at the cutoff.  Evaluated by the HBM pair
    path (``kupsmi_pair_forces`` with ``lj_truncation``); systems up to 16384
    atoms.  ``truncation_radius``: a scalar or one value per system."""

    def __init__(self, params: LennardJonesParameters, truncation_radius):
        super().__init__(params)
        self.truncation_radius = truncation_radius

    def __call__(self, particles: ParticleBatch, systems: SystemBatch, cutoffs=None) -> PotentialOut:
        dev = particles.positions.device
        n_sys = systems.n_sys
        sys_off, sizes = system_offsets(particles.system, n_sys)
        sig, eps = self.tables(dev)
        if cutoffs is None:
            cutoffs = self.params.cutoff
        cut = torch.as_tensor(cutoffs, dtype=torch.float64, device=dev)
        cut = (cut.expand(n_sys) if cut.dim() == 0 or cut.numel() == 1 else cut).contiguous()
        rtr = torch.as_tensor(self.truncation_radius, dtype=torch.float64, device=dev)
        rtr = (rtr.expand(n_sys) if rtr.dim() == 0 or rtr.numel() == 1 else rtr).contiguous()
        species = particles.species
        if species is None:
            species = torch.zeros(particles.n, dtype=torch.int32, device=dev)
        pos = particles.positions.to(torch.float32).contiguous()
        species = species.to(torch.int32).contiguous()
        forces = torch.empty((particles.n, 3), dtype=torch.float32, device=dev)
        e2 = torch.empty((n_sys, 2), dtype=torch.float64, device=dev)
        virial = torch.empty((n_sys, 3, 3), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().kupsmi_pair_forces(
            _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), particles.n,
            int(sizes.max()) if sizes.size else 0, _lib.ptr(pos), None, _lib.ptr(species),
            self.params.n_species, _lib.ptr(sig), _lib.ptr(eps), _lib.ptr(cut), None, None, None,
            None, _lib.ptr(systems.tril), 0, _lib.ptr(forces), _lib.ptr(e2), _lib.ptr(virial),
            _lib.ptr(rtr)))
        return PotentialOut(e2[:, 0].clone(), -forces, virial)
