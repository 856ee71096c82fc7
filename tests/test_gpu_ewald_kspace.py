"""Reciprocal-space Ewald alone (kupsmi_ewald_recip) against the FP64 oracle.

The force pass has two implementations chosen by the k grid: the matrix-core
one (z half width <= 8: the n_z contraction as an FP16 hi/lo split GEMM) and
the vector-ALU one (wider grids).  Both are held to the k-space forces alone,
not to the sum with the much larger real-space forces:
||F_gpu - F_ref|| / ||F_ref|| <= 1e-5 over all atoms and, per atom,
|F_gpu - F_ref| <= 1e-4 of the batch's RMS k-space force; energies and
virials 1e-5 relative.
"""

import numpy as np
import pytest
import torch

import oracle
from kups_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _recip(b, tril, alpha, nmax, dev, accumulate=False, f0=None):
    pos = torch.as_tensor(b.positions.astype(np.float32), device=dev)
    q = torch.as_tensor(b.charges.astype(np.float32), device=dev)
    off = torch.as_tensor(b.sys_off.astype(np.int32), device=dev)
    t = torch.as_tensor(tril, dtype=torch.float64, device=dev)
    al = torch.full((b.n_sys,), alpha, dtype=torch.float64, device=dev)
    f = (torch.zeros((b.n, 3), dtype=torch.float32, device=dev) if f0 is None
         else torch.as_tensor(f0, dtype=torch.float32, device=dev).clone())
    e = torch.zeros(b.n_sys, dtype=torch.float64, device=dev)
    w = torch.zeros((b.n_sys, 3, 3), dtype=torch.float64, device=dev)
    _lib.check(_lib.load().kupsmi_ewald_recip(
        _lib.ctx(), _lib.stream_ptr(), b.n_sys, _lib.ptr(off), b.n, int(b.sizes.max()),
        _lib.ptr(pos), _lib.ptr(q), _lib.ptr(t), _lib.ptr(al), None, nmax, nmax, nmax,
        int(accumulate), _lib.ptr(f), _lib.ptr(e), _lib.ptr(w), None))
    torch.cuda.synchronize()
    return e.cpu().numpy(), f.cpu().numpy().astype(np.float64), w.cpu().numpy()


def _oracle(b, tril, alpha, nmax):
    pos = np.asarray(b.positions, np.float32).astype(np.float64)
    kv = np.stack([oracle.kvecs(t, nmax) for t in tril])
    e, f, w, _ = oracle.ewald_recip(pos, b.charges, b.sys_off, tril, np.full(b.n_sys, alpha), kv)
    return e, f, w


def _check(got, ref):
    e, f, w = got
    er, fr, wr = ref
    np.testing.assert_allclose(e, er, rtol=1e-5)
    np.testing.assert_allclose(w, wr, rtol=1e-5, atol=1e-5 * np.abs(wr).max())
    assert np.linalg.norm(f - fr) / np.linalg.norm(fr) < 1e-5
    rms = np.sqrt((fr ** 2).sum(1).mean())
    assert np.abs(f - fr).max() < 1e-4 * rms


@pytest.mark.parametrize("nmax", [8, 5, 10])
def test_kspace_forces_vs_oracle_melts(gpu, nmax):
    """Melts of 1000, 2744 and 8000 ions (the bench's sizes): nmax 8 is the
    bench grid (matrix cores), 5 leaves features of the split GEMM unused,
    10 takes the vector-ALU pass."""
    b = workloads.nacl_melts(ms=[5, 7, 10], seed=11)
    _check(_recip(b, b.tril, 0.3, nmax, gpu), _oracle(b, b.tril, 0.3, nmax))


def test_kspace_forces_triclinic_and_accumulate(gpu):
    """A sheared cell (general tril, wrapped fractional phases) and the
    accumulate flag (forces added to what the buffer holds)."""
    b = workloads.nacl_melts(ms=[5, 6], seed=12)
    tril = b.tril.copy()
    tril[:, 1] = 0.15 * tril[:, 0]
    tril[:, 3] = -0.1 * tril[:, 0]
    tril[:, 4] = 0.2 * tril[:, 2]
    ref = _oracle(b, tril, 0.32, 8)
    base = np.random.default_rng(0).normal(size=(b.n, 3)).astype(np.float32)
    e, f, w = _recip(b, tril, 0.32, 8, gpu, accumulate=True, f0=base)
    _check((e, f - base.astype(np.float64), w), ref)


def test_kspace_forces_small_and_uneven_systems(gpu):
    """Systems smaller than one workgroup's 128 atoms and ragged batches: a
    64-ion melt, a two-ion pair and an 8000-ion melt in one launch."""
    parts = [workloads.nacl_melts(ms=[2], seed=13), workloads.nacl_melts(ms=[10], seed=14)]
    pos = [parts[0].positions, np.array([[0.0, 0.0, 0.0], [2.7, 0.3, -0.4]]), parts[1].positions]
    q = [parts[0].charges, np.array([1.0, -1.0]), parts[1].charges]
    tr = [parts[0].tril[0], workloads.cubic_tril(11.0)[0], parts[1].tril[0]]
    sizes = np.array([len(p) for p in pos])
    b = workloads.Batch(np.concatenate(pos), sizes, np.stack(tr),
                        np.zeros(sizes.sum(), np.int32), np.ones(sizes.sum()),
                        charges=np.concatenate(q))
    _check(_recip(b, b.tril, 0.3, 8, gpu), _oracle(b, b.tril, 0.3, 8))
