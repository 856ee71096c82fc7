"""HDF5 output of the two applications (application/md/logging.py,
application/mcmc/logging.py:68-199) through ``kups_amd.core.storage``, and
the analyses that read it back (application/md/analysis.py
``analyze_md_file``, application/mcmc/analysis.py).

MD (``MDLoggedData``), groups of one file:

* ``group.init`` (once): the atoms (positions, momenta, masses, species,
  system index) and the systems (lattice ``tril``, temperature, time step,
  friction, cutoff) at the start of production;
* ``group.step`` (every step): per-system ``potential_energy`` and
  ``kinetic_energy``;
* ``group.frame`` (at the end of every launch of ``chunk`` steps): the atoms'
  positions and momenta, ``stress_tensor`` = -(dU/dh)/V at fixed Cartesian
  positions (the reference's logged MD stress, application/md/data.py:101-107)
  and the configurational ``virial_pressure`` tr(W)/3V.  The reference logs
  atoms and stress in its per-step group; the fused MD kernel keeps the state
  on the device for a whole launch, so they are logged per launch (``chunk``
  = 1 logs them every step).

MCMC (``MCMCLoggedData``):

* ``group.fixed`` (once): the systems (lattice, temperature, pressure,
  log-activities), the adsorbate motifs (site positions, charges, species)
  and the host particles (positions, charges, species, chain sizes);
* ``group.per_step`` (every cycle): ``particle_count`` (chain x species),
  the guests (molecule site positions, species per slot, count per chain)
  and per chain the ``potential_energy``, the step widths and acceptance
  rates of the four moves (translation, rotation, reinsertion, exchange;
  application/mcmc/logging.py:141-170) and, with ``compute_stress``, the
  guest pressure terms (potential, tail, ideal gas).
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from ..core.storage import EveryNStep, HDF5StorageReader, LoggingFrequency, Once, WriterGroupConfig
from .analysis import (MCMCAnalysisResult, MDAnalysisResult, analyze_mcmc, analyze_md,
                       combine_mcmc, combine_md)

MOVES = ("translation", "rotation", "reinsertion", "exchange")


@dataclass
class AtLaunchEnd(LoggingFrequency):
    """The last step of every launch of ``n`` steps: steps n - 1, 2 n - 1, ...
    at index step // n (the state a launch leaves)."""

    n: int

    def should_log(self, step: int) -> bool:
        return (step + 1) % self.n == 0

    def leading_shape(self, total_steps: int) -> tuple[int, ...]:
        return (total_steps // self.n,)

    def dataset_index(self, step: int):
        return step // self.n


# ------------------------------------------------------------------ MD
@dataclass
class MDRecord:
    """What the MD driver hands the writer at a step: the state, the step's
    energies and, at the end of a launch, the stress and virial pressure."""

    state: object
    potential_energy: np.ndarray
    kinetic_energy: np.ndarray
    stress_tensor: np.ndarray
    virial_pressure: np.ndarray


@dataclass
class MDLoggedData:
    init: WriterGroupConfig
    step: WriterGroupConfig
    frame: WriterGroupConfig


def _system_index(st) -> np.ndarray:
    return np.repeat(np.arange(st.n_sys, dtype=np.int32), st.sizes)


def make_md_logged_data(chunk: int, initial_state) -> MDLoggedData:
    """``initial_state``: the MDState at the start of production (copied here
    for the once-logged group)."""
    st = initial_state
    cp = lambda x: x.detach().cpu().numpy().copy()  # noqa: E731
    snap = {"atoms": {"positions": cp(st.positions), "momenta": cp(st.momenta),
                      "masses": cp(st.masses), "species": cp(st.species),
                      "system": _system_index(st)},
            "systems": {"tril": cp(st.tril), "temperature": cp(st.temperature),
                        "time_step": cp(st.time_step), "friction_coefficient": cp(st.friction),
                        "cutoff": cp(st.cutoff)}}

    def init(r: MDRecord):
        return snap

    def step(r: MDRecord):
        return {"potential_energy": r.potential_energy, "kinetic_energy": r.kinetic_energy}

    def frame(r: MDRecord):
        return {"atoms": {"positions": r.state.positions, "momenta": r.state.momenta},
                "stress_tensor": r.stress_tensor, "virial_pressure": r.virial_pressure}

    return MDLoggedData(WriterGroupConfig(init, Once()), WriterGroupConfig(step, EveryNStep(1)),
                        WriterGroupConfig(frame, AtLaunchEnd(int(chunk))))


def analyze_md_file(path: str | Path, n_blocks: int | None = None) -> list[MDAnalysisResult]:
    """Per-system results (application/md/analysis.py ``analyze_md_file``):
    energies from the per-step group, pressure tr(stress)/3 from the frame
    group."""
    with HDF5StorageReader(path) as r:
        init = r.focus_group("init").read()
        step = r.focus_group("step").read()
        frame = r.focus_group("frame").read()
        n_steps = int(r.attr("actual_steps"))
    system = init["atoms"]["system"]
    n_atoms = np.bincount(system, minlength=step["potential_energy"].shape[1])
    pe = step["potential_energy"][:n_steps]
    ke = step["kinetic_energy"][:n_steps]
    pr = np.trace(frame["stress_tensor"], axis1=-2, axis2=-1) / 3.0
    return [analyze_md(pe[:, k], ke[:, k], pr[:, k], int(n_atoms[k]), n_blocks)
            for k in range(pe.shape[1])]


def analyze_md_file_combined(path: str | Path, n_blocks: int | None = None) -> MDAnalysisResult:
    """The systems of the file as replicas of one system."""
    return combine_md(analyze_md_file(path, n_blocks))


# ------------------------------------------------------------------ MCMC
@dataclass
class MCMCRecord:
    gcmc: object
    pressure: dict | None


@dataclass
class MCMCLoggedData:
    fixed: WriterGroupConfig
    per_step: WriterGroupConfig


def make_mcmc_logged_data(with_stress: bool) -> MCMCLoggedData:
    def fixed(r: MCMCRecord):
        g = r.gcmc
        d = g.d
        off = d["host_off"].cpu().numpy()
        return {"systems": {"tril": d["tril"], "temperature": d["kT"] / _kB(),
                            "log_activity": d["log_activity"]},
                "motifs": {"positions": d["motif_pos"], "charges": d["motif_q"],
                           "species": d["motif_species"]},
                "host_particles": {"positions": d["host"][:, :3], "charges": d["host"][:, 3],
                                   "species": d["host_species"], "sizes": np.diff(off)}}

    def per_step(r: MCMCRecord):
        g = r.gcmc
        d = g.d
        acc = acceptance_rates(d["hist"].cpu().numpy())
        systems = {"potential_energy": d["energy"]}
        for k, mv in enumerate(MOVES):
            systems[f"{mv}_step_width"] = d["width"][:, k]
            systems[f"{mv}_acceptance"] = acc[:, k]
        if with_stress:
            p = r.pressure
            systems["guest_pressure"] = {k: np.asarray(p[k], np.float64)
                                         for k in ("potential", "tail", "ideal")}
        return {"particle_count": g.loading_per_motif,
                "guests": {"positions": d["guest"], "species": d["guest_motif"],
                           "n_guest": d["n_guest"]},
                "systems": systems}

    return MCMCLoggedData(WriterGroupConfig(fixed, Once()),
                          WriterGroupConfig(per_step, EveryNStep(1)))


def acceptance_rates(hist: np.ndarray) -> np.ndarray:
    """Mean of each move's acceptance history (parameter_scheduler.py:161-208):
    the GCMC kernel keeps the last 100 accept flags of a move as bits 0..99 of
    four 32-bit words; (n_chain, 4, 4) words -> (n_chain, 4) rates."""
    h = np.ascontiguousarray(hist).view(np.uint32).copy()
    h[..., 3] &= 0xF
    bits = np.unpackbits(h.view(np.uint8), axis=-1)
    return bits.reshape(h.shape[0], h.shape[1], -1).sum(-1) / 100.0


def _kB() -> float:
    from ..core.constants import BOLTZMANN_CONSTANT

    return BOLTZMANN_CONSTANT


def analyze_mcmc_file(path: str | Path, n_blocks: int | None = None) -> list[MCMCAnalysisResult]:
    """Per-chain results (application/mcmc/analysis.py): energy, loading,
    heats of adsorption and, when logged, the guest pressure."""
    with HDF5StorageReader(path) as r:
        fixed = r.focus_group("fixed").read()
        ps = r.focus_group("per_step").read()
        n = int(r.attr("actual_steps"))
    temp = np.asarray(fixed["systems"]["temperature"], np.float64)
    e = ps["systems"]["potential_energy"][:n]
    counts = ps["particle_count"][:n]
    gp = ps["systems"].get("guest_pressure")
    return [analyze_mcmc(e[:, c], counts[:, c], float(temp[c]), n_blocks,
                         None if gp is None else {k: v[:n, c] for k, v in gp.items()})
            for c in range(e.shape[1])]


def analyze_mcmc_file_combined(path: str | Path, groups: list[range],
                               n_blocks: int | None = None) -> list[MCMCAnalysisResult]:
    """One result per group of chains (replicas of one host), combined."""
    res = analyze_mcmc_file(path, n_blocks)
    return [combine_mcmc([res[c] for c in g]) for g in groups]
