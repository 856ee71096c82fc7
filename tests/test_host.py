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


def test_autotune_keeps_the_default_within_the_noise_margin(monkeypatch):
    """A candidate replaces the default only when it is faster by more than
    the margin; the state is snapshot and restored around every candidate."""
    import types

    import torch

    from kups_amd.core import autotune as at

    st = types.SimpleNamespace(
        sizes=np.array([10]), tril=torch.tensor([[5.0, 0, 5, 0, 0, 5]], dtype=torch.float64),
        cutoff=torch.tensor([2.0], dtype=torch.float64),
        time_step=torch.tensor([0.1], dtype=torch.float64),
        temperature=torch.tensor([100.0], dtype=torch.float64), step=3,
        positions=torch.zeros(10, 3), momenta=torch.zeros(10, 3), forces=torch.zeros(10, 3))
    times = {1.0: 1.00, 2.0: 0.985, 3.0: 1.2}

    def make(c):
        def run(n):
            st.positions += 1.0  # the runner moves the state
            st.step += n
        run.c = c
        return run

    monkeypatch.setattr(at, "_time_steps", lambda run, steps: (run(steps), times[run.c])[1])
    monkeypatch.delenv("KUPS_AMD_TUNE_CACHE", raising=False)
    at._CACHE.clear()
    r = at._tune("a", st, [1.0, 2.0, 3.0], make, 10, prefer=1.0, margin=0.03)
    assert r.best == 1.0 and r.table[2.0] == 0.985
    assert st.step == 3 and float(st.positions.abs().max()) == 0.0
    times[2.0] = 0.95
    r = at._tune("b", st, [1.0, 2.0, 3.0], make, 10, prefer=1.0, margin=0.03)
    assert r.best == 2.0
