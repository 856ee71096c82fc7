"""``kups_md_lj`` on the fused HIP MD kernel (application/simulations/md_lj.py).

``build`` reads the deck's structure (CIF) and LJ table and makes an MDState
of ``replicas`` copies of the system (the kernel runs them side by side, each
with its own random stream), momenta from Maxwell-Boltzmann when
``initialize_momenta`` is set, else zero, and forces zero before the first
step (application/md/data.py:148-218); ``run`` does the warm-up and
production steps (application/md/simulation.py) and keeps the potential and
kinetic energy of every step and, at the end of every launch of ``chunk``
steps, two pressures: the one the reference logs, tr(-(dU/dh)/V)/3 with the
Cartesian positions fixed (``lattice_gradient_stress``; what its expected
values hold), and the configurational virial pressure tr(W)/(3V);
``analyze`` reduces them as application/md/analysis.py does.

md_lj.py builds its potential without the tail correction whatever the deck's
``tail_correction`` says (md_lj.py:91-100); so does this driver.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from ..md.data import MDState, make_md_state, maxwell_boltzmann_momenta
from ..md.integrators import (BAOAB_LANGEVIN, CSVR, CSVR_NPT, VELOCITY_VERLET,
                              FusedLJPropagator)
from ..md.observables import lattice_gradient_stress, volume
from ..core.storage import HDF5StorageWriter
from .analysis import MDAnalysisResult, analyze_md, combine_md
from .logging import MDRecord, make_md_logged_data
from .cif import Structure, read_cif
from .config import MdConfig
from .mcmc_rigid import lj_table

INTEGRATORS = {"verlet": VELOCITY_VERLET, "baoab_langevin": BAOAB_LANGEVIN, "csvr": CSVR,
               "csvr_npt": CSVR_NPT}


@dataclass
class MdRun:
    state: MDState
    propagator: FusedLJPropagator
    config: MdConfig
    structure: Structure
    replicas: int
    potential_energy: np.ndarray | None = None  # (n_steps, replicas)
    kinetic_energy: np.ndarray | None = None  # (n_steps, replicas)
    pressure: np.ndarray | None = None  # (n_samples, replicas), the logged definition
    virial_pressure: np.ndarray | None = None  # (n_samples, replicas), tr(W) / 3V


def build(config: MdConfig, *, replicas: int = 1, seed: int | None = None,
          skin: float | None = None, device=None) -> MdRun:
    s = read_cif(config.inp_file)
    if np.isnan(s.masses).any():
        bad = sorted({sym for sym, m in zip(s.symbols, s.masses) if np.isnan(m)})
        raise ValueError(f"no standard atomic weight for {bad}")
    lj = lj_table(config.lj, s.labels)
    index = {lab: i for i, lab in enumerate(lj.labels)}
    md = config.md
    n = len(s)
    run_seed = config.run.seed if seed is None else seed
    run_seed = 0 if run_seed is None else int(run_seed)
    sizes = np.full(replicas, n)
    masses = np.tile(s.masses, replicas)
    momenta = None
    if md.initialize_momenta:
        momenta = maxwell_boltzmann_momenta(masses, sizes, md.temperature,
                                            np.random.default_rng(run_seed))
    state = make_md_state(
        np.tile(s.positions, (replicas, 1)), masses,
        np.tile([index[lab] for lab in s.labels], replicas), sizes,
        np.tile(s.tril, (replicas, 1)), config.lj.cutoff, time_step_fs=md.time_step,
        temperature=md.temperature, friction_per_fs=md.friction_coefficient,
        thermostat_time_constant_fs=md.thermostat_time_constant,
        target_pressure_pa=md.target_pressure, pressure_coupling_time_fs=md.pressure_coupling_time,
        compressibility_per_pa=md.compressibility, minimum_scale_factor=md.minimum_scale_factor,
        momenta=momenta, device=device)
    prop = FusedLJPropagator(lj, INTEGRATORS[md.integrator], run_seed,
                             0.0 if md.integrator == "csvr_npt" else skin)
    return MdRun(state, prop, config, s, replicas)


def run(r: MdRun, *, num_warmup_steps: int | None = None, num_steps: int | None = None,
        chunk: int = 10, out_file: str | None = None) -> MdRun:
    """Warm-up, then production in launches of ``chunk`` steps;
    ``out_file``: HDF5 output of the production steps (``MDLoggedData``,
    application/md/simulation.py:150-170)."""
    cfg = r.config.run
    n_warm = cfg.num_warmup_steps if num_warmup_steps is None else int(num_warmup_steps)
    n_prod = cfg.num_steps if num_steps is None else int(num_steps)
    st, prop = r.state, r.propagator
    done = 0
    while done < n_warm:
        k = min(chunk, n_warm - done)
        prop(st, k, record=False)
        done += k
    pe, ke, pr, pv = [], [], [], []
    system = torch.repeat_interleave(torch.arange(st.n_sys, device=st.positions.device),
                                     torch.as_tensor(st.sizes, device=st.positions.device))
    writer = None
    if out_file is not None:
        z = np.zeros(st.n_sys)
        writer = HDF5StorageWriter(out_file, make_md_logged_data(chunk, st),
                                   MDRecord(st, z, z, np.zeros((st.n_sys, 3, 3)), z), n_prod)
        writer.__enter__()
    try:
        done = 0
        while done < n_prod:
            k = min(chunk, n_prod - done)
            obs = prop(st, k, record=True)
            pe.append(obs.potential_energy.T)
            ke.append(obs.kinetic_energy.T)
            vol = volume(st.tril)
            pv.append((torch.diagonal(obs.virial, dim1=-2, dim2=-1).sum(-1) / (3.0 * vol))[None, :])
            sig = lattice_gradient_stress(st.positions, st.forces, obs.virial, st.tril, system)
            pr.append((torch.diagonal(sig, dim1=-2, dim2=-1).sum(-1) / 3.0)[None, :])
            if writer is not None:
                pe_k, ke_k = pe[-1].cpu().numpy(), ke[-1].cpu().numpy()
                sig_k, pv_k = sig.cpu().numpy(), pv[-1][0].cpu().numpy()
                for j in range(k):
                    writer.log(MDRecord(st, pe_k[j], ke_k[j], sig_k, pv_k), done + j)
            done += k
    finally:
        if writer is not None:
            writer.__exit__(None, None, None)
    r.potential_energy = torch.cat(pe).cpu().numpy()
    r.kinetic_energy = torch.cat(ke).cpu().numpy()
    r.pressure = torch.cat(pr).cpu().numpy()
    r.virial_pressure = torch.cat(pv).cpu().numpy()
    return r


def run_independent(config: MdConfig, n_runs: int, *, seed: int | None = None,
                    chunk: int = 10, device=None) -> MdRun:
    """``n_runs`` one-system runs with seeds seed, seed + 1, ...  The CSVR
    thermostat and the cell-rescaling barostat draw one random number per
    step for the whole batch (md/integrators.py:549, 774), so copies run
    side by side in one launch share it and are not independent; separate
    launches are.  Returns the series stacked as replicas."""
    base = config.run.seed if seed is None else seed
    base = 0 if base is None else int(base)
    runs = [run(build(config, replicas=1, seed=base + k, device=device), chunk=chunk)
            for k in range(n_runs)]
    out = runs[-1]
    out.replicas = n_runs
    for name in ("potential_energy", "kinetic_energy", "pressure", "virial_pressure"):
        setattr(out, name, np.concatenate([getattr(x, name) for x in runs], axis=1))
    return out


def analyze(r: MdRun, n_blocks: int | None = None) -> MDAnalysisResult:
    if r.potential_energy is None:
        raise ValueError("run the system first")
    n = len(r.structure)
    res = [analyze_md(r.potential_energy[:, k], r.kinetic_energy[:, k], r.pressure[:, k], n,
                      n_blocks) for k in range(r.replicas)]
    return combine_md(res)
