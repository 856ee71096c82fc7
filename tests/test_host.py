"""Host-side helpers (CPU tensors)."""

import numpy as np
import torch

from kups_amd.core.batch import segment_sum


def test_segment_sum_matches_scatter_add_with_empty_systems():
    rng = np.random.default_rng(0)
    sizes = np.array([3, 0, 5, 1, 0, 7])
    system = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    v = rng.normal(size=(sizes.sum(), 3))
    ref = np.zeros((len(sizes), 3))
    np.add.at(ref, system, v)
    out = segment_sum(torch.as_tensor(v), torch.as_tensor(system), len(sizes)).numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    assert out[1].tolist() == [0.0, 0.0, 0.0] and out[4].tolist() == [0.0, 0.0, 0.0]


def test_autotune_key_and_cache(tmp_path, monkeypatch):
    """The workload signature changes with the batch and is stable for the
    same one; results round-trip through the on-disk cache."""
    import types

    import torch

    from kups_amd.core import autotune as at

    def st(sizes, L=10.0):
        n = len(sizes)
        tril = torch.zeros((n, 6), dtype=torch.float64)
        tril[:, 0] = tril[:, 2] = tril[:, 5] = L
        f = lambda v: torch.full((n,), v, dtype=torch.float64)
        return types.SimpleNamespace(sizes=np.asarray(sizes), tril=tril, cutoff=f(8.5),
                                     time_step=f(0.2), temperature=f(94.4))

    k1 = at.workload_key("md_lj", st([100, 200]))
    assert k1 == at.workload_key("md_lj", st([100, 200]))
    assert k1 != at.workload_key("md_lj", st([100, 201]))
    assert k1 != at.workload_key("md_lj", st([100, 200], L=11.0))
    assert k1 != at.workload_key("ewald_lj", st([100, 200]))
    monkeypatch.setenv("KUPS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    at._store("abc", {"best": 1.0, "table": {"1.0": 0.5, "2.0": 0.7}})
    at._CACHE.clear()
    assert at._lookup("abc")["best"] == 1.0
