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


# TraPPE/UFF-style LJ parameters (sigma [A], epsilon [eV]) of the host types
# and of the CO2 sites, as listed in the reference's statistical inputs
# (ci/statistical/inputs/lennard_jones/trappe.yaml:47,113,167,191,617,623-628).
GCMC_LJ = {
    "C_": (3.4309, 0.004553698956747378),
    "H_": (2.5711, 0.0019082130125430423),
    "O_": (3.1181, 0.002602114920617121),
    "N_": (3.2607, 0.002992428280911038),
    "Zn": (2.4616, 0.005377696700922425),
    "C_co2": (2.8, 0.0023266791910486473),
    "O_co2": (3.05, 0.006807690966401598),
    # TraPPE united-atom methane and three-site N2, whose massless centre
    # carries charge only (examples/lennard_jones/trappe.yaml:629-637)
    "CH4": (3.73, 0.012753648899081475),
    "N_n2": (3.31, 0.0031022389213981967),
    "M_n2": (None, None),
}
HOST_TYPES = ("C_", "H_", "O_", "N_", "Zn")
HOST_TYPE_P = (0.4, 0.3, 0.15, 0.1, 0.05)
# rigid linear CO2 (ci/statistical/inputs/ci/gcmc_co2_rubtak.yaml:2-12)
CO2 = {
    "symbols": ("C_co2", "O_co2", "O_co2"),
    "positions": ((0.0, 0.0, 0.0), (-1.16, 0.0, 0.0), (1.16, 0.0, 0.0)),
    "charges": (0.7, -0.35, -0.35),
    "critical_temperature": 303.75,
    "critical_pressure": 7_840_000.0,
    "acentric_factor": 0.22394,
}
# TraPPE methane and N2 (examples/adsorbate/ch4.yaml, n2.yaml)
CH4 = {
    "symbols": ("CH4",),
    "positions": ((0.0, 0.0, 0.0),),
    "charges": (0.0,),
    "critical_temperature": 190.564,
    "critical_pressure": 4_599_200.0,
    "acentric_factor": 0.01142,
}
N2 = {
    "symbols": ("N_n2", "M_n2", "N_n2"),
    "positions": ((0.0, 0.0, 0.55), (0.0, 0.0, 0.0), (0.0, 0.0, -0.55)),
    "charges": (-0.482, 0.964, -0.482),
    "critical_temperature": 126.192,
    "critical_pressure": 3_395_800.0,
    "acentric_factor": 0.0372,
}
ADSORBATES = {"CO2": CO2, "CH4": CH4, "N2": N2}


def lattice_from_lengths_angles(lengths, angles_deg) -> np.ndarray:
    """Lower-triangular lattice (rows a, b, c) of core/unitcell.py:100-136."""
    a, b, c = lengths
    al, be, ga = np.radians(angles_deg)
    cz = np.sqrt(1 - np.cos(be) ** 2 - ((np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)) ** 2)
    return np.array([[a, 0, 0], [b * np.cos(ga), b * np.sin(ga), 0],
                     [c * np.cos(be), c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga), c * cz]])


def perpendicular_widths(lat) -> np.ndarray:
    v = abs(np.linalg.det(lat))
    return v / np.linalg.norm(np.cross(lat[[1, 0, 0]], lat[[2, 2, 1]]), axis=-1)


@dataclass
class GcmcHosts:
    positions: np.ndarray  # (n_host_total, 3)
    charges: np.ndarray
    species: np.ndarray  # index into HOST_TYPES
    sizes: np.ndarray
    lattices: np.ndarray  # (n_chain, 3, 3)
    pressure: np.ndarray  # Pa
    temperature: np.ndarray  # K
    description: str = ""

    @property
    def n_chain(self) -> int:
        return len(self.sizes)

    @property
    def tril(self) -> np.ndarray:
        return self.lattices[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]


def gcmc_hosts(n_chain: int = 1024, nmin: int = 1500, nmax: int = 3000, seed: int = 0,
               lmin: float = 25.0, lmax: float = 30.0, angle_dev: float = 5.0,
               vacancy: float = 0.3, jitter: float = 0.1, qstd: float = 0.3,
               pmin: float = 1e4, pmax: float = 1e6, temperature: float = 298.0,
               min_width: float = 24.5, sizes=None) -> GcmcHosts:
    """Rigid hosts: N_h ~ U{nmin..nmax} atoms on a jittered cubic fractional
    lattice with a `vacancy` fraction of empty sites, in a triclinic box with
    lengths U(lmin, lmax) and angles 90 +- U(angle_dev) degrees (resampled
    until every perpendicular width exceeds min_width, so a 12 A cutoff needs no
    periodic images), charges N(0, qstd) neutralised, host types drawn with
    HOST_TYPE_P, pressure log-uniform in [pmin, pmax] Pa."""
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = rng.integers(nmin, nmax + 1, size=n_chain)
    sizes = np.asarray(sizes, np.int64)
    pos, q, sp, lats = [], [], [], []
    for n in sizes:
        while True:
            lat = lattice_from_lengths_angles(rng.uniform(lmin, lmax, 3),
                                              90.0 + rng.uniform(-angle_dev, angle_dev, 3))
            if perpendicular_widths(lat).min() > min_width:
                break
        side = int(np.ceil((n / (1 - vacancy)) ** (1 / 3)))
        g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
        pick = np.sort(rng.choice(len(g), size=int(n), replace=False))
        f = (g[pick] + 0.5) / side - 0.5
        r = f @ lat + rng.normal(0, jitter, (int(n), 3))
        fr = r @ np.linalg.inv(lat)
        fr = (fr + 0.5) % 1.0 - 0.5
        pos.append(fr @ lat)
        qq = rng.normal(0, qstd, int(n))
        q.append(qq - qq.mean())
        sp.append(rng.choice(len(HOST_TYPES), size=int(n), p=HOST_TYPE_P))
        lats.append(lat)
    press = np.exp(rng.uniform(np.log(pmin), np.log(pmax), len(sizes)))
    return GcmcHosts(np.concatenate(pos), np.concatenate(q), np.concatenate(sp).astype(np.int32),
                     sizes, np.array(lats), press, np.full(len(sizes), temperature),
                     description=f"{len(sizes)} rigid hosts, N_h in [{sizes.min()},{sizes.max()}], "
                                 f"{int(vacancy * 100)}% vacancies, angles 90+-{angle_dev}")
