"""``kups_mcmc_rigid`` on the HIP GCMC kernel (application/simulations/mcmc_rigid.py).

``build`` turns a YAML deck into a ``RigidGCMC`` batch: every host of the deck
(CIF, optional replication) becomes ``replicas`` independent chains with the
deck's pressure, temperature, move weights, LJ table, Ewald settings and
initial loading; ``run`` does the warm-up and production cycles of
mcmc_rigid.py:299-305 (a cycle is max(largest loading of the batch,
min_cycle_length) trial moves) and samples the energy and loading after every
production cycle, with S(k) and the running energy re-evaluated in full
between warm-up and production (keeping the reference's tracked-energy
composition); ``analyze`` reduces the samples per host as
application/mcmc/analysis.py does, replicas combined.

Up to four adsorbate species per deck (a gas mixture: each host's
composition and binary interaction parameters give the Peng-Robinson
fugacity of every species, application/mcmc/data.py:380-391; the loading,
heat of adsorption and initial insertions are per species).  The arithmetic
is all in the kernel; this module only assembles inputs and reads the
per-cycle samples.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from ..mcmc.gcmc import Motif, RigidGCMC
from ..potential.lennard_jones import LennardJonesParameters
from ..core.storage import HDF5StorageWriter
from .analysis import MCMCAnalysisResult, analyze_mcmc, combine_mcmc
from .cif import Structure, make_supercell, read_cif
from .config import McmcConfig
from .logging import MCMCRecord, make_mcmc_logged_data


def host_structure(host) -> Structure:
    s = read_cif(host.cif_file)
    if host.unitcell_replication is not None:
        s = make_supercell(s, host.unitcell_replication)
    return s


def _motif(ads) -> Motif:
    q = ads.charges if ads.charges else [0.0] * len(ads.symbols)
    return Motif.from_dict({"positions": ads.positions, "symbols": ads.symbols, "charges": q,
                            "critical_temperature": ads.critical_temperature,
                            "critical_pressure": ads.critical_pressure,
                            "acentric_factor": ads.acentric_factor})


def lj_table(config_lj, labels) -> LennardJonesParameters:
    """The deck's LJ table restricted to the labels in use, in the table's
    order (mixing is per pair, so the restriction changes no value)."""
    used = set(labels)
    missing = sorted(used - set(config_lj.parameters))
    if missing:
        raise ValueError(f"labels {missing} have no Lennard-Jones parameters")
    params = {k: v for k, v in config_lj.parameters.items() if k in used}
    return LennardJonesParameters.from_dict(config_lj.cutoff, params, config_lj.mixing_rule)


@dataclass
class McmcRun:
    gcmc: RigidGCMC
    config: McmcConfig
    replicas: int
    structures: list
    energy: np.ndarray | None = None  # (n_cycles, n_chain) eV
    loading: np.ndarray | None = None  # (n_cycles, n_chain), all species
    loading_per_species: np.ndarray | None = None  # (n_cycles, n_chain, n_species)
    # (n_samples, n_chain) eV/A^3 when the deck sets compute_stress
    pressure: dict | None = None


def build(config: McmcConfig, *, replicas: int = 1, seed: int | None = None,
          device=None) -> McmcRun:
    if not 1 <= len(config.adsorbates) <= 4:
        raise ValueError("the rigid GCMC kernel runs 1..4 adsorbate species per deck")
    motifs = [_motif(a) for a in config.adsorbates]
    mt = len(motifs)
    structs = [host_structure(h) for h in config.hosts]
    labels = [lab for s in structs for lab in s.labels] + [x for m_ in motifs for x in m_.species]
    lj = lj_table(config.lj, labels)
    index = {lab: i for i, lab in enumerate(lj.labels)}
    guest_species = [tuple(index[sym] for sym in m_.species) for m_ in motifs]
    pos, q, sp, sizes, lats, pres, temps, n_init = [], [], [], [], [], [], [], []
    comp, kij = [], []
    for h, s in zip(config.hosts, structs):
        for _ in range(replicas):
            pos.append(s.positions)
            q.append(s.charges)
            sp.append(np.array([index[lab] for lab in s.labels], np.int32))
            sizes.append(len(s))
            lats.append(s.lattice)
            pres.append(float(h.pressure))
            temps.append(float(h.temperature))
            ni = list(h.init_adsorbates or [])
            n_init.append([int(ni[t]) if t < len(ni) else 0 for t in range(mt)])
            y = np.asarray(h.adsorbate_composition if h.adsorbate_composition else
                           [1.0 / mt] * mt, float)
            if y.shape != (mt,):
                raise ValueError("adsorbate_composition needs one mole fraction per adsorbate")
            comp.append(y)
            k = np.zeros((mt, mt)) if h.adsorbate_interaction is None else \
                np.asarray(h.adsorbate_interaction, float)
            if k.shape != (mt, mt):
                raise ValueError("adsorbate_interaction must be n_ads x n_ads")
            kij.append(k)
    run_seed = config.run.seed if seed is None else seed
    run_seed = 0 if run_seed is None else int(run_seed)
    gcmc = RigidGCMC(np.concatenate(pos), np.concatenate(q), np.concatenate(sp), sizes,
                     np.stack(lats), np.array(pres), np.array(temps),
                     motifs[0] if mt == 1 else motifs, lj,
                     guest_species[0] if mt == 1 else guest_species,
                     ewald_precision=config.ewald.precision, ewald_cutoff=config.ewald.real_cutoff,
                     weights=config.run.weights, gmax=int(config.max_num_adsorbates),
                     seed=run_seed, tail_correction=bool(config.lj.tail_correction),
                     composition=np.array(comp), interaction=np.array(kij), device=device)
    n_init = np.array(n_init, np.int64)
    if n_init.any():
        gcmc.insert_random(n_init, seed=run_seed)
    return McmcRun(gcmc, config, replicas, structs)


def _cycle_length(gcmc: RigidGCMC, min_len: int) -> int:
    return max(int(gcmc.loading.max().item()) if gcmc.n_chain else 0, int(min_len))


def run(r: McmcRun, *, num_warmup_cycles: int | None = None,
        num_cycles: int | None = None, stress_every: int = 1,
        out_file: str | None = None) -> McmcRun:
    """Warm-up then production cycles; per-cycle energy and loading kept,
    and with ``compute_stress`` the guest pressure terms every
    ``stress_every`` cycles (RigidGCMC.guest_pressure).  ``out_file``: HDF5
    output of every production cycle (``MCMCLoggedData``,
    application/simulations/mcmc_rigid.py:433-465; the pressure terms are
    then evaluated every cycle)."""
    cfg = r.config.run
    n_warm = cfg.num_warmup_cycles if num_warmup_cycles is None else int(num_warmup_cycles)
    n_prod = cfg.num_cycles if num_cycles is None else int(num_cycles)
    g = r.gcmc
    for _ in range(n_warm):
        g.run(_cycle_length(g, cfg.min_cycle_length))
    if n_warm:
        # random initial molecules may overlap: the early energies are huge and
        # their FP32 pair terms leave rounding in the running FP64 energy
        g.resync()
    e = torch.empty((n_prod, g.n_chain), dtype=torch.float64, device=g.dev)
    n = torch.empty((n_prod, g.n_chain, g.n_motif), dtype=torch.int32, device=g.dev)
    terms = {"potential": [], "tail": [], "ideal": []}
    stress = bool(r.config.compute_stress)
    writer = None
    if out_file is not None:
        stress_every = 1
        z = np.zeros(g.n_chain)
        writer = HDF5StorageWriter(out_file, make_mcmc_logged_data(stress),
                                   MCMCRecord(g, {"potential": z, "tail": z, "ideal": z}), n_prod)
        writer.__enter__()
    try:
        for c in range(n_prod):
            g.run(_cycle_length(g, cfg.min_cycle_length))
            e[c].copy_(g.d["energy"])
            n[c].copy_(g.loading_per_motif)
            p = None
            if stress and c % stress_every == 0:
                p = g.guest_pressure()
                for k, v in p.items():
                    terms[k].append(v)
            if writer is not None:
                writer.log(MCMCRecord(g, p), c)
    finally:
        if writer is not None:
            writer.__exit__(None, None, None)
    r.energy = e.cpu().numpy()
    r.loading_per_species = n.cpu().numpy()
    r.loading = r.loading_per_species.sum(-1)
    if r.config.compute_stress and terms["potential"]:
        r.pressure = {k: np.stack(v) for k, v in terms.items()}
    return r


def analyze(r: McmcRun, n_blocks: int | None = None) -> list[MCMCAnalysisResult]:
    """One result per host of the deck, its replicas combined."""
    if r.energy is None:
        raise ValueError("run the chains first")
    out = []
    for h_i, h in enumerate(r.config.hosts):
        cols = range(h_i * r.replicas, (h_i + 1) * r.replicas)
        res = [analyze_mcmc(r.energy[:, c], r.loading_per_species[:, c], float(h.temperature),
                            n_blocks,
                            None if r.pressure is None else
                            {k: v[:, c] for k, v in r.pressure.items()})
               for c in cols]
        out.append(combine_mcmc(res))
    return out
