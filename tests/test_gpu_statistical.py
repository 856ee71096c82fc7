"""Statistical parity with the reference's own CI decks and expected values.

Each test runs one of the reference's statistical-CI decks
(ci/statistical/inputs/ci, copied under tests/golden/statistical) through the
YAML applications on the HIP kernels and compares every observable with the
reference's expected mean and standard error (ci/statistical/expected).

The bar is the reference validator's (ci/statistical/validate.py:103-140):
|mean - expected| <= 10 expected SEM.  The combined z-score
|mean - expected| / sqrt(sem^2 + expected_sem^2) is reported beside it, and the
GCMC loading is held to z <= 2.  Independent replicas of each deck run side by
side in one launch (each chain and system has its own random stream), so the
measured SEM is the spread of the replica means.

The MD decks' "pressure" is the one the reference's MD application logs,
tr(-(dU/dh)/V)/3 with Cartesian positions fixed (see
kups_amd.md.observables.lattice_gradient_stress); the physical
configurational pressure tr(W)/(3V) is kept beside it.

Set KUPS_STAT_OUT to a folder to keep each test's numbers as JSON.
"""

from __future__ import annotations

import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

from kups_amd.application import mcmc_rigid, md_lj
from kups_amd.application.analysis import z_score
from kups_amd.application.config import McmcConfig, MdConfig, load_yaml

STAT = Path(__file__).resolve().parent / "golden" / "statistical"
CI = STAT / "inputs" / "ci"
REF_SIGMAS = 10.0


def _compare(name: str, result, extra: dict | None = None) -> dict:
    exp = load_yaml(STAT / "expected" / f"{name}.yaml")["observables"]
    rows, bad = {}, []
    for key, e in exp.items():
        if not hasattr(result, key):
            continue
        r = getattr(result, key)
        m = float(np.asarray(r.mean).ravel()[0])
        s = float(np.asarray(r.sem).ravel()[0])
        em = float(np.asarray(e["expected_mean"], dtype=float).ravel()[0])
        es = float(np.asarray(e["expected_sem"], dtype=float).ravel()[0])
        if np.isnan(em):
            continue
        lim = REF_SIGMAS * es if es > 1e-10 else abs(em) * 0.01
        z = float(z_score(m, s, em, es))
        rows[key] = {"mean": m, "sem": s, "expected_mean": em, "expected_sem": es, "z": z,
                     "within_reference_bar": abs(m - em) <= lim}
        if abs(m - em) > lim:
            bad.append(f"{key}: {m:.6g} vs {em:.6g} (|diff| {abs(m - em):.3g} > {lim:.3g})")
    out = {"deck": name, "observables": rows, **(extra or {})}
    print(json.dumps(out), flush=True)
    sys.stdout.flush()
    folder = os.environ.get("KUPS_STAT_OUT")
    if folder:
        Path(folder).mkdir(parents=True, exist_ok=True)
        (Path(folder) / f"{name}.json").write_text(json.dumps(out, indent=1))
    assert not bad, "; ".join(bad)
    return rows


@pytest.mark.gpu
def test_gcmc_co2_rubtak(gpu):
    """CO2 in the RUBTAK Zr MOF (2x2x2, 912 host atoms), 10 kPa, 298.15 K:
    loading, energy (host-host Ewald included) and heat of adsorption."""
    cfg = McmcConfig.from_yaml(CI / "gcmc_co2_rubtak.yaml")
    r = mcmc_rigid.build(cfg, replicas=64, device=gpu)
    mcmc_rigid.run(r)
    res = mcmc_rigid.analyze(r)[0]
    err = np.abs(r.gcmc.energy_error())
    st = r.gcmc.stats
    rows = _compare("gcmc_co2_rubtak", res, {
        "replicas": 64, "max_energy_error_eV": float(err.max()),
        "accepted_fraction": (st["accepted"].sum(0) / np.maximum(st["trials"].sum(0), 1)).tolist()})
    assert err.max() < 1e-2
    assert rows["loading"]["z"] <= 2.0


@pytest.mark.gpu
def test_nvt_50co2_30box(gpu):
    """50 CO2 in an empty 30 A box without exchange moves: the energy pins the
    absolute composition (self, exclusion and intramolecular real-space terms);
    the deck's compute_stress adds the guest molecular-virial, tail and ideal
    pressures (sampled every 5th cycle)."""
    cfg = McmcConfig.from_yaml(CI / "nvt_50co2_30box.yaml")
    assert cfg.compute_stress
    r = mcmc_rigid.build(cfg, replicas=16, device=gpu)
    assert int(r.gcmc.loading.min()) == 50
    mcmc_rigid.run(r, stress_every=5)
    res = mcmc_rigid.analyze(r)[0]
    assert np.all(r.loading == 50)
    err = np.abs(r.gcmc.energy_error())
    _compare("nvt_50co2_30box", res, {"replicas": 16, "max_energy_error_eV": float(err.max())})
    assert err.max() < 1e-3


def _md(name: str, gpu, replicas: int):
    cfg = MdConfig.from_yaml(CI / f"{name}.yaml")
    r = md_lj.build(cfg, replicas=replicas, device=gpu)
    md_lj.run(r)
    return r, md_lj.analyze(r)


@pytest.mark.gpu
def test_md_nve_lj_argon(gpu):
    """256 Ar from the jittered FCC file at rest, velocity Verlet, dt 0.1 fs:
    the total energy is conserved and the averages match."""
    r, res = _md("md_nve_lj_argon", gpu, 1)
    te = r.potential_energy[:, 0] + r.kinetic_energy[:, 0]
    _compare("md_nve_lj_argon", res, {"replicas": 1, "energy_drift_eV_per_step": res.energy_drift,
                                      "total_energy_span_eV": float(te.max() - te.min())})


@pytest.mark.gpu
def test_md_nvt_lj_argon(gpu):
    """BAOAB Langevin, gamma 1/fs, dt 2 fs, 100 K: the kinetic temperature
    carries BAOAB's dt^2 bias (100.46 K), which the step order must reproduce."""
    r, res = _md("md_nvt_lj_argon", gpu, 8)
    _compare("md_nvt_lj_argon", res, {"replicas": 8})


@pytest.mark.gpu
def test_md_npt_lj_argon(gpu):
    """CSVR thermostat + stochastic cell rescaling at 42 MPa, 100 K.  The deck
    starts from the solid at a physical pressure near 340 MPa, so the box
    expands through the whole run; both integrators draw one random number per
    step for the batch, so the replicas are separate launches (own seeds)."""
    cfg = MdConfig.from_yaml(CI / "md_npt_lj_argon.yaml")
    r = md_lj.run_independent(cfg, 8, device=gpu)
    res = md_lj.analyze(r)
    _compare("md_npt_lj_argon", res, {"replicas": 8, "independent_launches": True,
                                      "final_volume_A3": float(r.state.tril[0, [0, 2, 5]].prod())})
