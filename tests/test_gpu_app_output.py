"""HDF5 output of the YAML applications on the GPU kernels: the files hold
exactly the series the runs return, and the file-based analyses give the same
results as the in-memory ones (application/md/analysis.py analyze_md_file,
application/mcmc/analysis.py)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from kups_amd.application import logging as applog
from kups_amd.application import mcmc_rigid, md_lj
from kups_amd.application.config import McmcConfig, MdConfig
from kups_amd.core import storage
from kups_amd.core.storage import HDF5StorageReader

pytestmark = pytest.mark.gpu

CI = Path(__file__).resolve().parent / "golden" / "statistical" / "inputs" / "ci"


def _same(a, b):
    for k in ("mean", "sem"):
        np.testing.assert_allclose(np.asarray(getattr(a, k)), np.asarray(getattr(b, k)),
                                   rtol=1e-12, atol=1e-300)


@pytest.fixture(autouse=True)
def _need_hdf5():
    try:
        storage._h5()
    except storage.HDF5Unavailable as e:
        pytest.skip(f"HDF5 C library not available: {e}")


def test_md_deck_output(gpu, tmp_path):
    cfg = MdConfig.from_yaml(CI / "md_nvt_lj_argon.yaml")
    r = md_lj.build(cfg, replicas=2, device=gpu)
    out = tmp_path / "md.h5"
    md_lj.run(r, num_warmup_steps=50, num_steps=205, chunk=10, out_file=str(out))
    with HDF5StorageReader(out) as f:
        assert f.list_groups() == ["group.init", "group.step", "group.frame"]
        assert int(f.attr("actual_steps")) == 205
        step = f.focus_group("step").read()
        frame = f.focus_group("frame").read()
        init = f.focus_group("init").read()
    np.testing.assert_array_equal(step["potential_energy"], r.potential_energy)
    np.testing.assert_array_equal(step["kinetic_energy"], r.kinetic_energy)
    # 20 whole launches of 10 steps; the last 5-step launch is not a frame
    assert frame["stress_tensor"].shape == (20, 2, 3, 3)
    np.testing.assert_allclose(np.trace(frame["stress_tensor"], axis1=-2, axis2=-1) / 3,
                               r.pressure[:20], rtol=1e-12)
    np.testing.assert_array_equal(init["atoms"]["system"], np.repeat([0, 1], 256))
    assert init["atoms"]["positions"].shape == (512, 3)
    assert not np.array_equal(frame["atoms"]["positions"][0], init["atoms"]["positions"])
    # the file analysis: every step's energies, the 20 launch-end stresses
    got = applog.analyze_md_file(out, n_blocks=10)
    assert len(got) == 2
    for k in range(2):
        ref = applog.analyze_md(r.potential_energy[:, k], r.kinetic_energy[:, k],
                                r.pressure[:20, k], 256, 10)
        for name in ("potential_energy", "kinetic_energy", "temperature", "pressure"):
            _same(getattr(got[k], name), getattr(ref, name))


@pytest.mark.parametrize("deck,stress", [("nvt_50co2_30box", True), ("gcmc_co2_rubtak", False)])
def test_mcmc_deck_output(gpu, tmp_path, deck, stress):
    cfg = McmcConfig.from_yaml(CI / f"{deck}.yaml")
    assert bool(cfg.compute_stress) == stress
    r = mcmc_rigid.build(cfg, replicas=2, device=gpu)
    out = tmp_path / "mc.h5"
    mcmc_rigid.run(r, num_warmup_cycles=3, num_cycles=24, out_file=str(out))
    with HDF5StorageReader(out) as f:
        assert f.list_groups() == ["group.fixed", "group.per_step"]
        assert int(f.attr("actual_steps")) == 24
        ps = f.focus_group("per_step").read()
        fixed = f.focus_group("fixed").read()
    np.testing.assert_array_equal(ps["systems"]["potential_energy"], r.energy)
    np.testing.assert_array_equal(ps["particle_count"], r.loading_per_species)
    np.testing.assert_array_equal(ps["guests"]["n_guest"][-1], r.gcmc.loading.cpu().numpy())
    np.testing.assert_array_equal(fixed["host_particles"]["sizes"],
                                  np.diff(r.gcmc.d["host_off"].cpu().numpy()))
    assert fixed["host_particles"]["positions"].shape[0] == fixed["host_particles"]["sizes"].sum()
    for mv in applog.MOVES:
        acc = ps["systems"][f"{mv}_acceptance"]
        assert acc.shape == (24, 2) and np.all((acc >= 0) & (acc <= 1))
        assert np.all(ps["systems"][f"{mv}_step_width"] > 0)
    if stress:
        for k in ("potential", "tail", "ideal"):
            np.testing.assert_array_equal(ps["systems"]["guest_pressure"][k], r.pressure[k])
    ref = mcmc_rigid.analyze(r, n_blocks=4)[0]
    got = applog.combine_mcmc(applog.analyze_mcmc_file(out, n_blocks=4))
    for k in ("energy", "loading") + (("pressure",) if stress else ()):
        _same(getattr(got, k), getattr(ref, k))
