"""Seeded synthetic inputs of the benchmark workloads (host numpy).

- ``argon_fcc``: 864 Ar, 6x6x6 FCC, a = 5.26 A (cubic PBC), LJ sigma 3.405 A,
  eps/kB 119.8 K, cutoff 2.5 sigma, mass 39.948 u.
- ``batched_lj_boxes``: independent cubic boxes with N ~ U{64..4096} atoms,
  FCC sites at density 0.8 sigma^-3 (N of the 4 m^3 sites chosen without
  replacement, m = ceil((N/4)^(1/3))), Gaussian jitter 0.05 sigma, same LJ.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .core.constants import BOLTZMANN_CONSTANT

AR_SIGMA = 3.405
AR_EPS = 119.8 * BOLTZMANN_CONSTANT
AR_MASS = 39.948
AR_CUTOFF = 2.5 * AR_SIGMA


def fcc_sites(m: int, a: float) -> np.ndarray:
    """4 m^3 FCC sites of an m x m x m cubic supercell of lattice constant a."""
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1)
    return ((g.reshape(-1, 1, 3) + base[None]) * a).reshape(-1, 3)


def wrap_cubic(pos: np.ndarray, L: float) -> np.ndarray:
    """Cartesian positions wrapped to [-L/2, L/2) (fractional [-0.5, 0.5))."""
    f = pos / L
    f = (f + 0.5) % 1.0 - 0.5
    return f * L


@dataclass
class Batch:
    positions: np.ndarray  # (n, 3)
    sizes: np.ndarray  # (n_sys,)
    tril: np.ndarray  # (n_sys, 6)
    species: np.ndarray  # (n,) int32
    masses: np.ndarray  # (n,)
    charges: np.ndarray | None = None
    description: str = ""

    @property
    def n_sys(self) -> int:
        return len(self.sizes)

    @property
    def n(self) -> int:
        return int(self.sizes.sum())

    @property
    def sys_off(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int32)


def cubic_tril(L) -> np.ndarray:
    L = np.atleast_1d(np.asarray(L, np.float64))
    t = np.zeros((L.size, 6))
    t[:, 0] = L
    t[:, 2] = L
    t[:, 5] = L
    return t


def argon_fcc(m: int = 6, a: float = 5.26) -> Batch:
    L = m * a
    pos = wrap_cubic(fcc_sites(m, a), L)
    n = len(pos)
    return Batch(pos, np.array([n]), cubic_tril(L), np.zeros(n, np.int32),
                 np.full(n, AR_MASS), description=f"argon FCC {m}x{m}x{m}, a={a} A")


def batched_lj_boxes(n_sys: int = 2048, nmin: int = 64, nmax: int = 4096, seed: int = 0,
                     density: float = 0.8, jitter: float = 0.05,
                     sizes: np.ndarray | None = None) -> Batch:
    """Independent cubic LJ boxes (density in sigma^-3, jitter in sigma)."""
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = rng.integers(nmin, nmax + 1, size=n_sys)
    sizes = np.asarray(sizes, np.int64)
    rho = density / AR_SIGMA**3
    pos_all = []
    trils = []
    for n in sizes:
        L = (n / rho) ** (1.0 / 3.0)
        m = int(np.ceil((n / 4) ** (1.0 / 3.0)))
        a = L / m
        sites = fcc_sites(m, a)
        pick = rng.choice(len(sites), size=int(n), replace=False)
        p = sites[np.sort(pick)] + rng.normal(0.0, jitter * AR_SIGMA, size=(int(n), 3))
        pos_all.append(wrap_cubic(p, L))
        trils.append(cubic_tril(L)[0])
    n = int(sizes.sum())
    return Batch(np.concatenate(pos_all) if pos_all else np.zeros((0, 3)), sizes,
                 np.array(trils).reshape(-1, 6), np.zeros(n, np.int32), np.full(n, AR_MASS),
                 description=f"{len(sizes)} LJ boxes, N in [{sizes.min() if n else 0},"
                             f"{sizes.max() if n else 0}], rho*={density}, jitter {jitter} sigma")


NA_MASS, CL_MASS = 22.98977, 35.453


def nacl_melts(n_sys: int = 512, mmin: int = 5, mmax: int = 10, seed: int = 0, a: float = 5.64,
               jitter: float = 0.1, ms=None) -> Batch:
    """Charged LJ melts: rock-salt lattices of N = 8 m^3 ions (m ~ U{mmin..mmax}),
    lattice constant a, Gaussian jitter (A), charges +1 (species 0) / -1
    (species 1) alternating on the lattice (neutral)."""
    rng = np.random.default_rng(seed)
    if ms is None:
        ms = rng.integers(mmin, mmax + 1, size=n_sys)
    pos_all, q_all, sp_all, trils = [], [], [], []
    for m in ms:
        g = np.stack(np.meshgrid(*[np.arange(2 * m)] * 3, indexing="ij"), -1).reshape(-1, 3)
        sp = (g.sum(-1) % 2).astype(np.int32)
        L = m * a
        p = g * (a / 2) + rng.normal(0.0, jitter, size=g.shape)
        pos_all.append(wrap_cubic(p, L))
        sp_all.append(sp)
        q_all.append(np.where(sp == 0, 1.0, -1.0))
        trils.append(cubic_tril(L)[0])
    sizes = np.array([8 * m**3 for m in ms], np.int64)
    sp = np.concatenate(sp_all)
    return Batch(np.concatenate(pos_all), sizes, np.array(trils), sp,
                 np.where(sp == 0, NA_MASS, CL_MASS), charges=np.concatenate(q_all),
                 description=f"{len(ms)} NaCl-like melts, N=8m^3, m in [{min(ms)},{max(ms)}]")
