"""Morse bond potential over explicit bonds (potential/classical/morse.py).

``MorseBondParameters`` keeps the reference's layout (labels, and r0 / D /
alpha as (n_species, n_species) tables) and constructors (``from_harmonic``,
``from_uff`` with UFF equations 2 and 4, uff_utils.py:16-39);
``MorseBondPotential`` evaluates energies, position gradients and the virial
of a bond list with the HIP kernel ``kupsmi_morse_bonds``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.batch import ParticleBatch, SystemBatch, system_offsets
from .lennard_jones import PotentialOut


def uff_bond_length(bond_radius, electronegativity) -> np.ndarray:
    """r_ij = r_i + r_j - r_EN, r_EN = r_i r_j (sqrt(chi_i) - sqrt(chi_j))^2 /
    (chi_i r_i + chi_j r_j) (uff_utils.py:16-39)."""
    r = np.asarray(bond_radius, float)
    chi = np.asarray(electronegativity, float)
    ri, rj = r[:, None], r[None, :]
    ci, cj = chi[:, None], chi[None, :]
    denom = ci * ri + cj * rj
    safe = np.where(np.abs(denom) > 1e-10, denom, 1.0)
    r_en = np.where(np.abs(denom) > 1e-10, ri * rj * (np.sqrt(ci) - np.sqrt(cj)) ** 2 / safe, 0.0)
    return ri + rj - r_en


@dataclass
class MorseBondParameters:
    """labels (n_species,); r0 [A], D [eV], alpha [1/A]: (n_species, n_species)."""

    labels: tuple[str, ...]
    r0: np.ndarray
    D: np.ndarray
    alpha: np.ndarray

    @classmethod
    def from_harmonic(cls, labels, r0, k, D) -> "MorseBondParameters":
        """alpha = sqrt(k / D) (morse.py from_harmonic)."""
        r0, k, D = (np.asarray(x, float) for x in (r0, k, D))
        return cls(tuple(labels), r0, D, np.sqrt(k / D))

    @classmethod
    def from_uff(cls, labels, bond_radius, electronegativity, effective_charge,
                 dissociation_energy) -> "MorseBondParameters":
        """UFF: r0 from uff_bond_length, k = 664.12 Z_i Z_j / r0^3,
        alpha = sqrt(k / (2 D)) (morse.py from_uff)."""
        r0 = uff_bond_length(bond_radius, electronegativity)
        z = np.asarray(effective_charge, float)
        k = 664.12 * z[:, None] * z[None, :] / r0**3
        D = np.asarray(dissociation_energy, float)
        return cls(tuple(labels), r0, D, np.sqrt(k / (2.0 * D)))

    @property
    def n_species(self) -> int:
        return len(self.labels)


class MorseBondPotential:
    """Morse energy over a bond list (Potential protocol, core/potential.py:
    186-198): returns per-system energies, dE/dr and the virial."""

    def __init__(self, params: MorseBondParameters):
        self.params = params
        self._dev = None

    def _tables(self, dev):
        if self._dev != dev:
            t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
            self._D, self._alpha, self._r0 = t(self.params.D), t(self.params.alpha), t(self.params.r0)
            self._dev = dev

    def __call__(self, particles: ParticleBatch, systems: SystemBatch, bonds: torch.Tensor,
                 shifts: torch.Tensor | None = None) -> PotentialOut:
        dev = particles.positions.device
        self._tables(dev)
        n, n_sys = particles.n, systems.n_sys
        bonds = bonds.to(device=dev, dtype=torch.int32).contiguous()
        if bonds.numel() and (int(bonds.min()) < 0 or int(bonds.max()) >= n):
            raise IndexError("bond index out of range")
        if shifts is not None:
            shifts = shifts.to(device=dev, dtype=torch.int32).contiguous()
        sys_off, _ = system_offsets(particles.system, n_sys)
        species = particles.species
        if species is None:
            species = torch.zeros(n, dtype=torch.int32, device=dev)
        species = species.to(torch.int32).contiguous()
        forces = torch.empty((n, 3), dtype=torch.float32, device=dev)
        energy = torch.empty(n_sys, dtype=torch.float64, device=dev)
        virial = torch.empty((n_sys, 9), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().kupsmi_morse_bonds(
            _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), n, _lib.ptr(particles.positions),
            _lib.ptr(species), self.params.n_species, _lib.ptr(self._D),
            _lib.ptr(self._alpha), _lib.ptr(self._r0), int(bonds.shape[0]), _lib.ptr(bonds),
            _lib.ptr(shifts), _lib.ptr(systems.tril), 0, _lib.ptr(forces), _lib.ptr(energy),
            _lib.ptr(virial)))
        return PotentialOut(energy, -forces, virial.view(n_sys, 3, 3))
