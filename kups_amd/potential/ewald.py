"""Ewald summation (potential/classical/ewald.py).

``EwaldParameters`` keeps the reference's fields (alpha, real-space cutoff,
integer k grid); ``estimate_ewald_parameters`` restates its estimator
(1166-1319) with scipy on the host, ``kvec_grid`` the grid of
kvecs_from_kmax (1322-1339).  ``EwaldPotential`` evaluates, per system,
  sr    K/2 sum_directed q_i q_j erfc(alpha r)/r [r < rc]   (305-325)
  lr    K sum_k P(k) |S(k)|^2                              (349-439, 585-601)
  self  -K alpha / sqrt(pi) sum q^2                        (277-302)
  excl  E_sr - E_vac over intramolecular pairs (optional)  (328-346)
with K = HARTREE*BOHR, on the GPU (kupsmi_pair_forces, kupsmi_ewald_recip);
the self term and the exclusion correction over explicit small edge lists are
FP64 torch reductions.  There is no neutralising-background term, as in the
reference.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.batch import ParticleBatch, SystemBatch, segment_sum, system_offsets
from ..core.constants import TO_STANDARD_UNITS


@dataclass
class EwaldParameters:
    """alpha (n_sys,) [1/A], cutoff (n_sys,) [A], kmax (n_sys,) [1/A] (< 0: whole
    grid), and the half widths (mx, my, mz) of the integer k grid."""

    alpha: np.ndarray
    cutoff: np.ndarray
    kmax: np.ndarray
    grid: tuple[int, int, int]

    @classmethod
    def box(cls, n_sys: int, alpha: float, cutoff: float, nmax: int) -> "EwaldParameters":
        """Explicit parameters with the full n in [0..nmax] x [-nmax..nmax]^2 grid."""
        return cls(np.full(n_sys, alpha, float), np.full(n_sys, cutoff, float),
                   np.full(n_sys, -1.0), (nmax, nmax, nmax))

    @classmethod
    def make(cls, charges_per_system, trils, epsilon_total: float = 1e-8,
             real_cutoff: float | None = None) -> "EwaldParameters":
        """Per-system estimates padded to one k grid (EwaldParameters.make,
        ewald.py:170-223)."""
        ests = [estimate_ewald_parameters(q, t, real_cutoff=real_cutoff,
                                          epsilon_total=epsilon_total)
                for q, t in zip(charges_per_system, trils)]
        n = max(kvec_box_extent(t, e["k_max"]) for t, e in zip(trils, ests))
        return cls(np.array([e["alpha"] for e in ests]), np.array([e["real_cutoff"] for e in ests]),
                   np.array([e["k_max"] for e in ests]), (n, n, n))


def _lattice(tril) -> np.ndarray:
    t = np.asarray(tril, float)
    return np.array([[t[0], 0, 0], [t[1], t[2], 0], [t[3], t[4], t[5]]])


def kvec_box_extent(tril, kmax: float) -> int:
    """n = ceil(kmax / smallest singular value of 2 pi Linv^T) (ewald.py:1333-1335)."""
    rvecs = np.linalg.inv(_lattice(tril)).T * 2 * np.pi
    return int(np.ceil(kmax / np.linalg.svd(rvecs)[1].min()))


def kvec_grid(tril, kmax: float) -> np.ndarray:
    """Integer k vectors with |k| <= kmax, the set of kvecs_from_kmax (1322-1339)."""
    n = kvec_box_extent(tril, kmax)
    rvecs = np.linalg.inv(_lattice(tril)).T * 2 * np.pi
    g = np.stack(np.meshgrid(np.arange(0, n + 1), np.arange(-n, n + 1), np.arange(-n, n + 1)),
                 -1).reshape(-1, 3)
    return g[np.linalg.norm(g @ rvecs, axis=-1) <= kmax]


def estimate_ewald_parameters(charges, tril, real_cutoff: float | None = None,
                              epsilon_total: float = 1e-8) -> dict:
    """Host estimator of ewald.py:1166-1319 (scipy).  With a fixed real-space
    cutoff: alpha from minimize_scalar(|erfc(z) - rc eps/2|) / rc and
    k_max = 2 alpha sqrt(-ln(eps/2)); otherwise the grid search of the
    reference's cost model."""
    import scipy.optimize
    import scipy.special

    q = np.asarray(charges, float)
    lat = _lattice(tril)
    volume = abs(np.linalg.det(lat))
    Q2 = float(q @ q)
    N = q.size
    perp = volume / np.linalg.norm(np.cross(lat[[1, 0, 0]], lat[[2, 2, 1]]), axis=-1)
    max_radius = perp.min() / 2
    eps_target = epsilon_total / 2
    lattice_length = volume ** (1 / 3)

    def real_err(rc, a):
        return np.inf if rc <= 0 or a <= 0 else Q2 / np.sqrt(N) * scipy.special.erfc(a * rc) / rc

    def recip_err(kc, a):
        if kc <= 0 or a <= 0:
            return np.inf
        x = -(kc**2) / (4 * a**2)
        return 0.0 if x < -700 else Q2 * a / (np.sqrt(N) * np.pi) * np.exp(x)

    if real_cutoff is not None:
        target = real_cutoff * eps_target
        res = scipy.optimize.minimize_scalar(lambda z: abs(scipy.special.erfc(z) - target))
        if not res.success:
            raise ValueError("Failed to find alpha for given real space cutoff.")
        alpha = res.x / real_cutoff
        kmax = 2 * alpha * np.sqrt(-np.log(eps_target))
        return {"alpha": float(alpha), "real_cutoff": float(real_cutoff), "k_max": float(kmax),
                "error_real": real_err(real_cutoff, alpha), "error_recip": recip_err(kmax, alpha)}

    def minimize(f, bounds, n=200):
        xs = np.linspace(bounds[0], bounds[1], n)
        return xs[np.argmin([f(x) for x in xs])]

    def opt_rc(a):
        rc = minimize(lambda r: abs(real_err(r, a) - eps_target), (0.01, min(20.0, max_radius)))
        return np.inf if real_err(rc, a) > 2 * eps_target else rc

    def opt_kc(a):
        kc = minimize(lambda k: abs(recip_err(k, a) - eps_target), (0.1, 5.0))
        return np.inf if recip_err(kc, a) > 2 * eps_target else kc

    def cost(a):
        rc = opt_rc(a)
        if rc == np.inf:
            return np.inf
        kc = opt_kc(a)
        if kc == np.inf:
            return np.inf
        return float(rc**3 * N / volume * 4 / 3 * np.pi + (2 * kc * lattice_length / 2 / np.pi + 1) ** 3)

    alpha = minimize(cost, (0.001, 2), n=400)
    rc, kc = opt_rc(alpha), opt_kc(alpha)
    if rc <= 0 or kc <= 0 or rc == np.inf or kc == np.inf:
        raise ValueError("Invalid cutoff values computed")
    return {"alpha": float(alpha), "real_cutoff": float(rc), "k_max": float(kc),
            "error_real": real_err(rc, alpha), "error_recip": recip_err(kc, alpha)}


@dataclass
class EwaldOut:
    short_range: torch.Tensor
    long_range: torch.Tensor
    self_energy: torch.Tensor
    exclusion: torch.Tensor
    forces: torch.Tensor
    virial: torch.Tensor
    structure_factor: torch.Tensor | None

    @property
    def total_energies(self) -> torch.Tensor:
        return self.short_range + self.long_range + self.self_energy + self.exclusion


def self_energy(q: torch.Tensor, system: torch.Tensor, alpha: torch.Tensor, n_sys: int):
    """-K alpha/sqrt(pi) sum q^2 per system (ewald.py:277-302), FP64."""
    q2 = segment_sum(q.double() ** 2, system, n_sys)
    return -TO_STANDARD_UNITS * alpha * q2 / math.sqrt(math.pi)


def exclusion_correction(positions, q, system, systems: SystemBatch, alpha, edges_ij, shifts):
    """E_sr - E_vac over explicit intramolecular directed edges (ewald.py:328-346,
    coulomb.py:64-83), FP64; edges carry image shifts like kUPS Edges."""
    n_sys = systems.n_sys
    out = torch.zeros(n_sys, dtype=torch.float64, device=positions.device)
    if edges_ij.numel() == 0:
        return out
    i, j = edges_ij[:, 0].long(), edges_ij[:, 1].long()
    s = system[i].long()
    lat = systems.lattice()[s]
    d = positions[j].double() - positions[i].double() + torch.einsum("ej,ejk->ek", shifts.double(), lat)
    r = d.norm(dim=-1)
    qq = q[i].double() * q[j].double()
    e = qq * (torch.special.erfc(alpha[s] * r) - 1.0) / r
    out.index_add_(0, s, 0.5 * TO_STANDARD_UNITS * e)
    return out


class EwaldPotential:
    """Ewald energy, forces and virial of a particle batch (Potential protocol,
    core/potential.py:186-198)."""

    def __init__(self, params: EwaldParameters):
        self.params = params

    def __call__(self, particles: ParticleBatch, systems: SystemBatch, *,
                 structure_factor: bool = False, intramolecular_edges=None) -> EwaldOut:
        dev = particles.positions.device
        n_sys = systems.n_sys
        sys_off, sizes = system_offsets(particles.system, n_sys)
        maxn = int(sizes.max()) if sizes.size else 0
        q = particles.charges.to(torch.float32).contiguous()
        alpha = torch.as_tensor(self.params.alpha, dtype=torch.float64, device=dev).contiguous()
        rc = torch.as_tensor(self.params.cutoff, dtype=torch.float64, device=dev).contiguous()
        kmax = torch.as_tensor(self.params.kmax, dtype=torch.float64, device=dev).contiguous()
        species = torch.zeros(particles.n, dtype=torch.int32, device=dev)
        forces = torch.zeros((particles.n, 3), dtype=torch.float32, device=dev)
        e2 = torch.zeros((n_sys, 2), dtype=torch.float64, device=dev)
        vir = torch.zeros((n_sys, 3, 3), dtype=torch.float64, device=dev)
        lib = _lib.load()
        pos = particles.positions.contiguous()
        _lib.check(lib.kupsmi_pair_forces(
            _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), particles.n, maxn,
            _lib.ptr(pos), _lib.ptr(q), _lib.ptr(species), 1, None, None, None, None,
            _lib.ptr(alpha), _lib.ptr(rc), None, _lib.ptr(systems.tril), 0, _lib.ptr(forces),
            _lib.ptr(e2), _lib.ptr(vir), None))
        mx, my, mz = self.params.grid
        nk = (mx + 1) * (2 * my + 1) * (2 * mz + 1)
        elr = torch.zeros(n_sys, dtype=torch.float64, device=dev)
        sk = torch.empty((n_sys, nk, 2), dtype=torch.float64, device=dev) if structure_factor else None
        _lib.check(lib.kupsmi_ewald_recip(
            _lib.ctx(), _lib.stream_ptr(), n_sys, _lib.ptr(sys_off), particles.n, maxn,
            _lib.ptr(pos), _lib.ptr(q), _lib.ptr(systems.tril), _lib.ptr(alpha), _lib.ptr(kmax),
            mx, my, mz, 1, _lib.ptr(forces), _lib.ptr(elr), _lib.ptr(vir), _lib.ptr(sk)))
        es = self_energy(q, particles.system, alpha, n_sys)
        if intramolecular_edges is not None:
            ij, sh = intramolecular_edges
            ex = exclusion_correction(pos, q, particles.system, systems, alpha, ij, sh)
        else:
            ex = torch.zeros(n_sys, dtype=torch.float64, device=dev)
        return EwaldOut(e2[:, 1].clone(), elr, es, ex, forces, vir, sk)
