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
