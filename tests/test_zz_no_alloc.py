"""The allocation-free reciprocal-space Ewald entry point
(kupsmi_ewald_recip_noalloc, scratch in a static device array) against
kupsmi_ewald_recip (context scratch) through the C ABI: two calls in a row,
the second on a smaller batch, must give the same bits (same kernels, same
inputs; the static array keeps the previous call's contents, which no result
may depend on).  A request beyond the array's 64 MiB is refused on the host.
"""

import numpy as np
import pytest
import torch

from kups_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _inputs(ms, seed, dev):
    b = workloads.nacl_melts(ms=ms, seed=seed)
    t = {
        "off": torch.as_tensor(b.sys_off, device=dev),
        "pos": torch.as_tensor(b.positions.astype(np.float32), device=dev),
        "q": torch.as_tensor(b.charges.astype(np.float32), device=dev),
        "tril": torch.as_tensor(b.tril, device=dev),
        "alpha": torch.full((b.n_sys,), 0.3, dtype=torch.float64, device=dev),
        "kmax": torch.full((b.n_sys,), -1.0, dtype=torch.float64, device=dev),
    }
    return b, t


def _call(fn_ctx, b, t, nm, with_sk, dev):
    f = torch.full((b.n, 3), 7.0, dtype=torch.float32, device=dev)
    e = torch.full((b.n_sys,), 7.0, dtype=torch.float64, device=dev)
    w = torch.full((b.n_sys, 3, 3), 7.0, dtype=torch.float64, device=dev)
    nk = (nm + 1) * (2 * nm + 1) ** 2
    sk = torch.full((b.n_sys, nk, 2), 7.0, dtype=torch.float64, device=dev) if with_sk else None
    lib = _lib.load()
    args = (_lib.stream_ptr(), b.n_sys, _lib.ptr(t["off"]), b.n, int(b.sizes.max()),
            _lib.ptr(t["pos"]), _lib.ptr(t["q"]), _lib.ptr(t["tril"]), _lib.ptr(t["alpha"]),
            _lib.ptr(t["kmax"]), nm, nm, nm, 0, _lib.ptr(f), _lib.ptr(e), _lib.ptr(w),
            _lib.ptr(sk))
    if fn_ctx:
        _lib.check(lib.kupsmi_ewald_recip(_lib.ctx(), *args))
    else:
        _lib.check(lib.kupsmi_ewald_recip_noalloc(*args))
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (f, e, w) + ((sk,) if with_sk else ())]


@pytest.mark.parametrize("with_sk", [False, True])
def test_noalloc_ewald_recip_matches_context_entry_point(gpu, with_sk):
    for ms, seed, nm in (([5, 6, 4, 5], 3, 8), ([3], 4, 6)):  # second call: smaller
        b, t = _inputs(ms, seed, gpu)
        ref = _call(True, b, t, nm, with_sk, gpu)
        got = _call(False, b, t, nm, with_sk, gpu)
        assert np.isfinite(ref[1]).all() and np.abs(ref[0]).max() > 0
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


def test_noalloc_ewald_recip_rejects_oversized_calls(gpu):
    b, t = _inputs([5], 5, gpu)
    lib = _lib.load()
    # 512 systems x 2601 k vectors x 16 atom chunks: about 370 MiB of scratch
    rc = lib.kupsmi_ewald_recip_noalloc(
        _lib.stream_ptr(), 512, _lib.ptr(t["off"]), b.n, 8000, _lib.ptr(t["pos"]),
        _lib.ptr(t["q"]), _lib.ptr(t["tril"]), _lib.ptr(t["alpha"]), _lib.ptr(t["kmax"]), 8, 8, 8,
        0, None, None, None, None)
    assert rc == -1
    assert b"static array" in lib.kupsmi_last_error()
