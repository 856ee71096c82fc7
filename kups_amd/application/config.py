"""YAML configurations of the two applications on the hot path.

``kups_md_lj`` (application/simulations/md_lj.py:44-56, application/md/data.py:
113-146) and ``kups_mcmc_rigid`` (application/mcmc/data.py:60-170,
application/simulations/mcmc_rigid.py:105-130) take the same YAML files here:
field names, defaults and units are the reference's, so an input deck written
for the reference runs unchanged.

The loader is ``yaml.SafeLoader`` with two additions the reference's decks
rely on:

* ``!import [a.yaml, b.yaml]`` is replaced by the files' contents (paths
  relative to the including file; mappings are merged left to right);
* floats without a dot or without an exponent sign (``1e10``, ``1.0e10``)
  are floats, as in YAML 1.2, not strings.
"""

from __future__ import annotations

import re
from dataclasses import dataclass, field, fields
from pathlib import Path

import yaml

_FLOAT = re.compile(r"""^[-+]?(?:
    [0-9][0-9_]*\.[0-9_]*(?:[eE][-+]?[0-9]+)?
  | \.[0-9_]+(?:[eE][-+]?[0-9]+)?
  | [0-9][0-9_]*[eE][-+]?[0-9]+
  | \.(?:inf|Inf|INF)
  | \.(?:nan|NaN|NAN))$""", re.X)


def _merge(a, b):
    if isinstance(a, dict) and isinstance(b, dict):
        out = dict(a)
        for k, v in b.items():
            out[k] = _merge(out[k], v) if k in out else v
        return out
    return b


def load_yaml(path: str | Path):
    """Parse a YAML deck with ``!import`` resolved relative to its directory."""
    path = Path(path)

    class Loader(yaml.SafeLoader):
        pass

    Loader.add_implicit_resolver("tag:yaml.org,2002:float", _FLOAT, list("-+0123456789."))

    def _import(loader, node):
        names = (loader.construct_sequence(node) if isinstance(node, yaml.SequenceNode)
                 else [loader.construct_scalar(node)])
        out = None
        for name in names:
            out = _merge(out, load_yaml(path.parent / str(name))) if out is not None \
                else load_yaml(path.parent / str(name))
        return out

    Loader.add_constructor("!import", _import)
    with open(path) as fh:
        return yaml.load(fh, Loader=Loader)  # noqa: S506 - SafeLoader subclass


def _build(cls, d: dict, where: str):
    names = {f.name for f in fields(cls)}
    unknown = set(d) - names
    if unknown:
        raise ValueError(f"{where}: unknown keys {sorted(unknown)}")
    return cls(**d)


def _resolve(base: Path, p: str | Path) -> Path:
    p = Path(p)
    return p if p.is_absolute() else (base / p)


# ------------------------------------------------------------------ LJ block
@dataclass
class LjConfig:
    """md_lj.py:38-42 / the ``lj`` block of both applications."""

    cutoff: float
    parameters: dict
    mixing_rule: str = "lorentz_berthelot"
    tail_correction: bool = False

    def parameters_for(self, labels) -> dict:
        """(sigma, epsilon) of the given labels in the table's order, with
        the labels that are absent from the table reported."""
        missing = sorted(set(labels) - set(self.parameters))
        if missing:
            raise ValueError(f"labels {missing} have no Lennard-Jones parameters")
        return self.parameters


# ------------------------------------------------------------------ MD
@dataclass
class MdRunConfig:
    """application/md/data.py:113-123."""

    out_file: str | None = None
    num_steps: int = 0
    num_warmup_steps: int = 0
    seed: int | None = None


@dataclass
class MdParameters:
    """application/md/data.py:126-146 (fs, 1/fs, K, Pa, 1/Pa)."""

    temperature: float
    time_step: float
    friction_coefficient: float
    thermostat_time_constant: float
    target_pressure: float
    pressure_coupling_time: float
    compressibility: float
    minimum_scale_factor: float
    integrator: str
    initialize_momenta: bool = False

    def __post_init__(self):
        if self.integrator not in ("verlet", "baoab_langevin", "csvr", "csvr_npt"):
            raise ValueError(f"unknown integrator {self.integrator!r}")


@dataclass
class MdConfig:
    """md_lj.py:44-48: run, md, lj, inp_file."""

    run: MdRunConfig
    md: MdParameters
    lj: LjConfig
    inp_file: Path
    base_dir: Path = field(default=Path("."))

    @classmethod
    def from_dict(cls, d: dict, base_dir: str | Path = ".") -> "MdConfig":
        base = Path(base_dir)
        return cls(_build(MdRunConfig, d.get("run", {}), "run"),
                   _build(MdParameters, d["md"], "md"), _build(LjConfig, d["lj"], "lj"),
                   _resolve(base, d["inp_file"]), base)

    @classmethod
    def from_yaml(cls, path: str | Path) -> "MdConfig":
        return cls.from_dict(load_yaml(path), Path(path).parent)


# ------------------------------------------------------------------ MCMC
@dataclass
class AdsorbateConfig:
    """application/mcmc/data.py:60-105: a rigid molecule and its PR constants."""

    positions: list
    symbols: list
    critical_temperature: float
    critical_pressure: float
    acentric_factor: float
    charges: list | None = None
    masses: list | None = None
    atomic_numbers: list | None = None


@dataclass
class HostConfig:
    """application/mcmc/data.py:107-130."""

    cif_file: Path
    pressure: float
    temperature: float
    unitcell_replication: int | list | None = None
    init_adsorbates: list = field(default_factory=lambda: [0])
    adsorbate_composition: list = field(default_factory=lambda: [1.0])
    adsorbate_interaction: list | None = None


@dataclass
class McmcRunConfig:
    """application/mcmc/data.py:151-167 (move weights are relative)."""

    out_file: str | None = None
    num_warmup_cycles: int = 0
    num_cycles: int = 0
    min_cycle_length: int = 1
    translation_prob: float = 1 / 6
    rotation_prob: float = 1 / 6
    reinsertion_prob: float = 1 / 6
    exchange_prob: float = 1 / 2
    seed: int | None = None

    @property
    def weights(self) -> tuple[float, float, float, float]:
        return (self.translation_prob, self.rotation_prob, self.reinsertion_prob,
                self.exchange_prob)


@dataclass
class EwaldConfig:
    """mcmc_rigid.py:97-103: real-space cutoff and target precision."""

    real_cutoff: float = 12.0
    precision: float = 1e-6


@dataclass
class McmcConfig:
    """mcmc_rigid.py:105-130."""

    run: McmcRunConfig
    adsorbates: list
    hosts: list
    lj: LjConfig
    ewald: EwaldConfig
    max_num_adsorbates: int = 100
    compute_stress: bool = False
    base_dir: Path = field(default=Path("."))

    @classmethod
    def from_dict(cls, d: dict, base_dir: str | Path = ".") -> "McmcConfig":
        base = Path(base_dir)
        hosts = []
        for h in d["hosts"]:
            h = dict(h)
            h["cif_file"] = _resolve(base, h["cif_file"])
            hosts.append(_build(HostConfig, h, "hosts"))
        return cls(_build(McmcRunConfig, d.get("run", {}), "run"),
                   [_build(AdsorbateConfig, a, "adsorbates") for a in d["adsorbates"]], hosts,
                   _build(LjConfig, d["lj"], "lj"), _build(EwaldConfig, d.get("ewald", {}), "ewald"),
                   int(d.get("max_num_adsorbates", 100)), bool(d.get("compute_stress", False)), base)

    @classmethod
    def from_yaml(cls, path: str | Path) -> "McmcConfig":
        return cls.from_dict(load_yaml(path), Path(path).parent)
