"""GPU known-answer checks of the shared random stream: the device's
Philox4x32-10 against Random123's published vectors, and the O-step normals
(kupsmi_normals) against Box-Muller of the oracle's Philox words."""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import _lib

from test_rng_barostat_cpu import KAT

pytestmark = pytest.mark.gpu


def test_device_philox_known_answers(gpu):
    ctr = torch.as_tensor(np.array([c for c, _, _ in KAT], np.uint32).view(np.int32), device=gpu)
    key = torch.as_tensor(np.array([k for _, k, _ in KAT], np.uint32).view(np.int32), device=gpu)
    out = torch.zeros((len(KAT), 4), dtype=torch.int32, device=gpu)
    _lib.check(_lib.load().kupsmi_philox(_lib.stream_ptr(), len(KAT), _lib.ptr(ctr),
                                         _lib.ptr(key), _lib.ptr(out)))
    got = out.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, np.array([o for _, _, o in KAT], np.uint32))


def test_device_philox_random_counters_match_oracle(gpu):
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2**32, (512, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, (512, 2), dtype=np.uint64).astype(np.uint32)
    out = torch.zeros((512, 4), dtype=torch.int32, device=gpu)
    tc = torch.as_tensor(ctr.view(np.int32), device=gpu)
    tk = torch.as_tensor(key.view(np.int32), device=gpu)
    _lib.check(_lib.load().kupsmi_philox(_lib.stream_ptr(), 512, _lib.ptr(tc), _lib.ptr(tk),
                                         _lib.ptr(out)))
    got = out.cpu().numpy().view(np.uint32)
    want = np.stack([oracle.philox_raw(c, k) for c, k in zip(ctr, key)])
    np.testing.assert_array_equal(got, want)


def test_o_step_normals_are_box_muller_of_the_stream(gpu):
    seed, step, n = 0x1234_5678_9ABC_DEF0, (3 << 32) + 17, 300
    eta = torch.empty((n, 3), dtype=torch.float32, device=gpu)
    _lib.check(_lib.load().kupsmi_normals(_lib.stream_ptr(), seed, step, n, _lib.ptr(eta)))
    got = eta.cpu().numpy()
    w = np.stack([oracle.philox_raw([i, step & 0xFFFFFFFF, step >> 32, 0x4B555053],
                                    [seed & 0xFFFFFFFF, seed >> 32]) for i in range(n)])
    u = (w.astype(np.float64) + 0.5) * 2.0**-32
    r0, r1 = np.sqrt(-2 * np.log(u[:, 0])), np.sqrt(-2 * np.log(u[:, 2]))
    want = np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                     r1 * np.cos(2 * np.pi * u[:, 3])], 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
