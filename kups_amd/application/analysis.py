"""Block averaging and the per-system observables of the two applications.

The statistics are the reference's (core/utils/block_average.py,
application/md/analysis.py:80-120, application/mcmc/analysis.py:81-157):

* ``block_average``: the series is cut into n equal blocks (the tail that does
  not fill a block is dropped); mean of the block means, SEM = sample standard
  deviation (ddof 1) of the block means / sqrt(n);
* ``optimal_block_average``: block counts n_samples//2, //4, ... down to
  ``min_blocks``; the first block size after which the SEM changes by less
  than ``rtol`` relative is used, else the largest block size;
* MD: T = 2K / (k_B (3N - 3)), pressure = tr(stress) / 3 with the stress of
  the configurational virial;
* MCMC: energy with the optimal block count, loading and heat of adsorption
  with the same count; q_st per block = cov(U, N) var(N)^-1 - k_B T.

Replicas (independent chains of the same system) are combined as the mean of
the replica means with SEM = std(replica means) / sqrt(R) when R >= 2.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..core.constants import BOLTZMANN_CONSTANT


@dataclass
class BlockAverageResult:
    mean: np.ndarray | float
    sem: np.ndarray | float
    n_blocks: int


def compute_block_means(data, n_blocks: int, axis: int = 0) -> np.ndarray:
    data = np.moveaxis(np.asarray(data, np.float64), axis, 0)
    size = data.shape[0] // n_blocks
    if size < 1:
        raise ValueError(f"{data.shape[0]} samples cannot fill {n_blocks} blocks")
    blocks = data[: size * n_blocks].reshape((n_blocks, size) + data.shape[1:]).mean(axis=1)
    return np.moveaxis(blocks, 0, axis)


def block_average_from_blocks(blocks, axis: int = 0) -> BlockAverageResult:
    blocks = np.asarray(blocks, np.float64)
    n = blocks.shape[axis]
    return BlockAverageResult(blocks.mean(axis=axis), blocks.std(axis=axis, ddof=1) / np.sqrt(n), n)


def block_average(data, n_blocks: int, axis: int = 0) -> BlockAverageResult:
    return block_average_from_blocks(compute_block_means(data, n_blocks, axis), axis)


def optimal_block_average(data, axis: int = 0, min_blocks: int = 4,
                          rtol: float = 0.05) -> BlockAverageResult:
    data = np.asarray(data, np.float64)
    n_samples = data.shape[axis]
    top = n_samples // 2
    counts = [top >> i for i in range(64) if (top >> i) >= min_blocks]
    if not counts:
        raise ValueError(f"{n_samples} samples are too few for {min_blocks} blocks")
    sizes = np.array([n_samples // c for c in counts])
    sems = np.array([np.asarray(block_average(data, c, axis).sem) for c in counts])
    rel = np.abs(np.diff(sems, axis=0)) / (np.abs(sems[:-1]) + 1e-10)
    if rel.ndim > 1:
        rel = rel.reshape(len(rel), -1).max(axis=1)
    hit = np.nonzero(rel < rtol)[0]
    idx = hit[0] + 1 if len(hit) else len(sizes) - 1
    return block_average(data, n_samples // int(sizes[idx]), axis)


def combine_replicas(results: list[BlockAverageResult]) -> BlockAverageResult:
    """One result from R replicas of the same system."""
    if len(results) == 1:
        return results[0]
    means = np.array([np.asarray(r.mean, np.float64) for r in results])
    return BlockAverageResult(means.mean(axis=0), means.std(axis=0, ddof=1) / np.sqrt(len(means)),
                              int(sum(r.n_blocks for r in results)))


# ------------------------------------------------------------------ MD
@dataclass
class MDAnalysisResult:
    potential_energy: BlockAverageResult
    kinetic_energy: BlockAverageResult
    total_energy: BlockAverageResult
    temperature: BlockAverageResult
    pressure: BlockAverageResult
    energy_drift: float
    n_atoms: int
    n_steps: int


def analyze_md(potential_energy, kinetic_energy, pressure, n_atoms: int,
               n_blocks: int | None = None) -> MDAnalysisResult:
    """One system's series: potential/kinetic energy per step (eV) and the
    configurational pressure (eV/A^3) at the steps it was sampled."""
    pe = np.asarray(potential_energy, np.float64)
    ke = np.asarray(kinetic_energy, np.float64)
    p = np.asarray(pressure, np.float64)
    te = pe + ke
    temp = 2.0 * ke / (BOLTZMANN_CONSTANT * (3 * n_atoms - 3))
    if n_blocks is None:
        pres = optimal_block_average(p)
        n_blocks = pres.n_blocks
    else:
        pres = block_average(p, n_blocks)
    slope = float(np.polyfit(np.arange(len(te)), te, 1)[0]) if len(te) > 1 else 0.0
    return MDAnalysisResult(block_average(pe, n_blocks), block_average(ke, n_blocks),
                            block_average(te, n_blocks), block_average(temp, n_blocks), pres,
                            slope, int(n_atoms), len(te))


def combine_md(results: list[MDAnalysisResult]) -> MDAnalysisResult:
    r0 = results[0]
    return MDAnalysisResult(*(combine_replicas([getattr(r, k) for r in results]) for k in
                              ("potential_energy", "kinetic_energy", "total_energy",
                               "temperature", "pressure")),
                            float(np.mean([r.energy_drift for r in results])), r0.n_atoms,
                            r0.n_steps)


# ------------------------------------------------------------------ MCMC
@dataclass
class MCMCAnalysisResult:
    energy: BlockAverageResult
    loading: BlockAverageResult
    heat_of_adsorption: BlockAverageResult
    total_heat_of_adsorption: BlockAverageResult
    pressure: BlockAverageResult | None = None
    pressure_potential: BlockAverageResult | None = None
    pressure_tail_correction: BlockAverageResult | None = None
    pressure_ideal_gas: BlockAverageResult | None = None


def analyze_mcmc(energy, counts, temperature: float, n_blocks: int | None = None,
                 pressure: dict | None = None) -> MCMCAnalysisResult:
    """One system's per-cycle energy (eV) and adsorbate counts (n, n_species);
    ``pressure``: per-sample 'potential', 'tail' and 'ideal' terms (eV/A^3),
    block-averaged with the energy's block count, total = their sum
    (application/mcmc/analysis.py:142-151)."""
    u = np.asarray(energy, np.float64)
    n = np.asarray(counts, np.float64)
    if n.ndim == 1:
        n = n[:, None]
    if n_blocks is None:
        e_res = optimal_block_average(u)
        n_blocks = e_res.n_blocks
    else:
        e_res = block_average(u, n_blocks)
    load = block_average(n, n_blocks)
    ub = compute_block_means(u, n_blocks)
    nb = compute_block_means(n, n_blocks)
    unb = compute_block_means(u[:, None] * n, n_blocks)
    nnb = compute_block_means(n[:, :, None] * n[:, None, :], n_blocks)
    cov = nnb - nb[:, :, None] * nb[:, None, :]
    diff = unb - ub[:, None] * nb
    with np.errstate(invalid="ignore", divide="ignore"):
        inv = np.linalg.pinv(cov) if np.any(cov == 0) else np.linalg.inv(cov)
        hoa = np.einsum("bi,bij->bj", diff, inv) - temperature * BOLTZMANN_CONSTANT
        w = nb / nb.sum(axis=1, keepdims=True)
    hoa_res = block_average_from_blocks(hoa)
    tot = block_average_from_blocks((hoa * w).sum(axis=1))
    out = MCMCAnalysisResult(e_res, load, hoa_res, tot)
    if pressure is not None:
        nb = min(n_blocks, len(pressure["potential"]))
        parts = {k: np.asarray(v, np.float64) for k, v in pressure.items()}
        out.pressure = block_average(parts["potential"] + parts["tail"] + parts["ideal"], nb)
        out.pressure_potential = block_average(parts["potential"], nb)
        out.pressure_tail_correction = block_average(parts["tail"], nb)
        out.pressure_ideal_gas = block_average(parts["ideal"], nb)
    return out


_MCMC_FIELDS = ("energy", "loading", "heat_of_adsorption", "total_heat_of_adsorption",
                "pressure", "pressure_potential", "pressure_tail_correction", "pressure_ideal_gas")


def combine_mcmc(results: list[MCMCAnalysisResult]) -> MCMCAnalysisResult:
    vals = []
    for k in _MCMC_FIELDS:
        parts = [getattr(r, k) for r in results]
        vals.append(None if any(p is None for p in parts) else combine_replicas(parts))
    return MCMCAnalysisResult(*vals)


def z_score(mean, sem, expected_mean, expected_sem) -> np.ndarray:
    """|measured - expected| / sqrt(sem^2 + expected_sem^2), the statistic the
    reference's validator bounds (ci/statistical/validate.py)."""
    d = np.abs(np.asarray(mean, np.float64) - np.asarray(expected_mean, np.float64))
    s = np.sqrt(np.asarray(sem, np.float64) ** 2 + np.asarray(expected_sem, np.float64) ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(s > 0, d / s, np.where(d == 0, 0.0, np.inf))
