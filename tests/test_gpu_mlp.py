"""GPU checks of the learned insertion proposal: descriptors (numpy
restatement), BF16-MFMA inference (PyTorch FP32 on the same BF16 weights and
BF16-rounded activations, tolerance 1e-2 of the logit scale), training
gradients (FP64 oracle on the same BF16-rounded weights, 1e-4 relative in
norm), training progress, and detailed balance of the biased chain
(ideal gas: <N> = a V for an arbitrary proposal table)."""

import math

import numpy as np
import pytest
import torch

import oracle
from kups_amd import workloads
from kups_amd.mcmc.gcmc import Motif, RigidGCMC, make_co2_gcmc
from kups_amd.mcmc.proposal import InsertionSamples, LearnedProposal
from kups_amd.potential.lennard_jones import LennardJonesParameters

pytestmark = pytest.mark.gpu


def bf16_to_f32(t: torch.Tensor) -> torch.Tensor:
    return (t.to(torch.int32) << 16).view(torch.float32)


def _bf16_round(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(torch.float32)


def _small(seed=0, n_chain=3):
    h = workloads.gcmc_hosts(n_chain, nmin=300, nmax=400, seed=seed, vacancy=0.95)
    h.pressure[:] = 1e6
    return make_co2_gcmc(h, gmax=64, seed=seed + 7)


def test_descriptors_match_numpy(gpu):
    g = _small()
    lp = LearnedProposal(g, nv=8, rmax=8.0, rho0=0.1)
    desc = bf16_to_f32(lp.desc).cpu().numpy().reshape(g.n_chain, 8**3, 32)
    host = g.d["host"].cpu().numpy()
    off = g.d["host_off"].cpu().numpy()
    tril = g.d["tril"].cpu().numpy()
    s = 1
    lat = np.array([[tril[s, 0], 0, 0], [tril[s, 1], tril[s, 2], 0], [tril[s, 3], tril[s, 4], tril[s, 5]]])
    hp = host[off[s]:off[s + 1], :3].astype(np.float64)
    nv = 8
    for v in [0, 77, 300, 511]:
        i = np.array([v // 64, (v // 8) % 8, v % 8])
        c = ((i + 0.5) / nv) @ lat
        d = (hp - c) @ np.linalg.inv(lat)
        d -= np.round(d)
        r = np.linalg.norm(d @ lat, axis=1)
        cnt = np.bincount(np.minimum((r[r < 8.0] / 0.25).astype(int), 31), minlength=32)
        shell = 4 / 3 * np.pi * ((np.arange(1, 33) * 0.25) ** 3 - (np.arange(32) * 0.25) ** 3) * 0.1
        ref = cnt / shell
        # counts can move by one at a bin edge (FP32 distances)
        assert np.abs(desc[s, v] - ref).max() <= np.max(1.0 / shell) * 1.01 + 1e-2 * ref.max()


def test_mfma_inference_matches_fp32(gpu):
    g = _small()
    lp = LearnedProposal(g, nv=8)
    # random descriptors and weights of realistic scale
    rng = np.random.default_rng(2)
    x = torch.as_tensor(rng.gamma(2.0, 0.5, (1000, 32)), dtype=torch.float32, device=gpu)
    lp.desc = x.to(torch.bfloat16).view(torch.int16).contiguous()
    out = lp.infer().cpu()
    w = bf16_to_f32(lp.wbf).cpu().double()
    wf = lp.wq.cpu().double()
    W1 = w[:4096].reshape(32, 128); b1 = wf[4096:4224]
    W2 = w[4224:20608].reshape(128, 128); b2 = wf[20608:20736]
    W3 = w[20736:37120].reshape(128, 128); b3 = wf[37120:37248]
    w4 = w[37248:37376]; b4 = wf[37376]
    X = _bf16_round(x).cpu().double()
    silu = torch.nn.functional.silu
    h1 = _bf16_round(silu(X @ W1 + b1).float()).double()
    h2 = _bf16_round(silu(h1 @ W2 + b2).float()).double()
    h3 = _bf16_round(silu(h2 @ W3 + b3).float()).double()
    ref = h3 @ w4 + b4
    err = (out.double() - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-2, err


def test_training_gradient_vs_oracle(gpu):
    g = _small()
    lp = LearnedProposal(g, nv=8)
    rng = np.random.default_rng(3)
    rows = torch.as_tensor(rng.integers(0, lp.desc.shape[0], 200), dtype=torch.int32, device=gpu)
    y = torch.as_tensor(rng.normal(0, 2, 200), dtype=torch.float32, device=gpu)
    grad = torch.empty(lp.master.numel(), dtype=torch.float32, device=gpu)
    loss = lp.train_step(rows, y, lr=0.0, apply=False, grad=grad)
    X = bf16_to_f32(lp.desc[rows.long()]).cpu().numpy().astype(np.float64)
    l_ref, g_ref = oracle.mlp_loss_and_grad(lp.wq.cpu().numpy().astype(np.float64), X,
                                            y.cpu().numpy().astype(np.float64))
    gg = grad.cpu().numpy()
    assert np.linalg.norm(gg - g_ref) / np.linalg.norm(g_ref) < 1e-4
    assert loss.double().sum().item() / 200 == pytest.approx(l_ref, rel=1e-4)


def test_training_reduces_loss(gpu):
    g = _small()
    lp = LearnedProposal(g, nv=12)
    n = 20000
    rng = np.random.default_rng(4)
    rows = torch.as_tensor(rng.integers(0, lp.desc.shape[0], n), dtype=torch.int32, device=gpu)
    X = bf16_to_f32(lp.desc[rows.long()]).float()
    y = (X[:, :8].sum(1) - X[:, 8:16].sum(1)) * 0.5  # a learnable target
    s = InsertionSamples(n, 12, gpu)
    s.rows.copy_(rows)
    s.y.copy_(y)
    s.count.fill_(n)
    hist = lp.train(s, epochs=8, batch=2048, lr=3e-3)
    assert hist[-1] < 0.3 * hist[0], hist


def test_detailed_balance_ideal_gas_with_biased_proposal(gpu):
    """An arbitrary (random) proposal table must leave <N> = a V unchanged."""
    n = 64
    L = 25.0
    lat = np.repeat(np.eye(3)[None] * L, n, 0)
    lj = LennardJonesParameters.from_dict(12.0, {"X": (3.0, 0.0)})
    motif = Motif.from_dict({**workloads.CO2, "charges": (0.0, 0.0, 0.0)})
    g = RigidGCMC(np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.int32), np.zeros(n, np.int64),
                  lat, 1e5, 298.0, motif, lj, (0, 0, 0), weights=(1, 0, 0, 3), gmax=128, seed=5)
    aV = 6.0
    g.log_activity[:] = math.log(aV / L**3)
    g.d["log_activity"].copy_(torch.as_tensor(g.log_activity, device=gpu))
    lp = LearnedProposal(g, nv=8, mix=0.05)
    # strongly non-uniform table: random logits with spread 3
    lp.logits.copy_(torch.randn(lp.logits.shape, generator=torch.Generator(device=gpu).manual_seed(1),
                                device=gpu) * 3.0)
    from kups_amd import _lib
    _lib.check(_lib.load().kupsmi_proposal_table(_lib.stream_ptr(), n, lp.nvox, _lib.ptr(lp.logits),
                                                 1.0, 0.05, _lib.ptr(lp.cdf), _lib.ptr(lp.guide)))
    samples = []
    for _ in range(60):
        g.run(100, proposal=lp)
        samples.append(g.loading.cpu().numpy().copy())
    s = np.array(samples[10:]).astype(float)
    assert s.mean() == pytest.approx(aV, abs=0.3)
    assert s.var() == pytest.approx(aV, rel=0.25)


def _train_learned(g, n_samples, epochs=4):
    lp = LearnedProposal(g)
    s = InsertionSamples(n_samples, lp.nv, g.dev)
    while s.n < n_samples:
        g.run(200, samples=s)
    hist = lp.train(s, epochs=epochs)
    assert hist[-1] < hist[0]
    lp.build_table()
    return lp


def _window(g, prop, n_chunks, chunk=500):
    s0 = g.stats
    acc = torch.zeros(g.n_chain, dtype=torch.float64, device=g.dev)
    for _ in range(n_chunks):
        g.run(chunk, proposal=prop)
        acc += g.loading
    s1 = g.stats
    return (int((s1["insertions"] - s0["insertions"]).sum()), (acc / n_chunks).cpu().numpy())


def test_learned_proposal_gains_over_uniform(gpu):
    """Dense hosts of the GCMC workload (256 chains): from the same
    equilibrated state and random stream, the trained proposal accepts many
    more insertions than uniform insertion, with the same loading (paired z)
    and the same tracked-energy error against a full re-evaluation."""
    h = workloads.gcmc_hosts(256, seed=3)
    g = make_co2_gcmc(h, gmax=64, seed=11)
    lp = _train_learned(g, 200_000)
    for _ in range(12):
        g.run(1000, proposal=lp)
    snap = g.snapshot()
    res = {}
    for name, prop in (("uniform", None), ("learned", lp)):
        g.restore(snap)
        ins, load = _window(g, prop, 16)
        err = np.abs(g.energy_error()) / np.maximum(np.abs(g.d["energy"].cpu().numpy()), 1.0)
        res[name] = (ins, load, err.max())
    (iu, lu, eu), (il, ll, el) = res["uniform"], res["learned"]
    assert il >= 10 and il >= 2 * iu + 5, (iu, il)
    d = ll - lu
    z = d.mean() / (d.std(ddof=1) / np.sqrt(len(d)))
    assert abs(z) < 3.0, (z, lu.mean(), ll.mean())
    assert eu < 1e-5 and el < 1e-5, (eu, el)


def test_learned_loading_matches_uniform_from_empty(gpu):
    """Uncharged open hosts (300-400 LJ atoms, 95 % vacancies, 10 bar), where
    both samplers equilibrate within a few thousand trials: both chains start
    empty, and their time-averaged loadings agree (paired over chains: the
    move type is shared by the batch, and both runs draw the same stream)."""
    h = workloads.gcmc_hosts(256, nmin=300, nmax=400, seed=4, vacancy=0.95)
    h.pressure[:] = 1e6
    h.charges[:] = 0.0
    g = make_co2_gcmc(h, gmax=64, seed=12)
    empty = g.snapshot()
    lp = _train_learned(g, 200_000)
    loads, ins = {}, {}
    for name, prop in (("uniform", None), ("learned", lp)):
        g.restore(empty)
        g.run(10000, proposal=prop)
        ins[name], loads[name] = _window(g, prop, 80)
    lu, ll = loads["uniform"], loads["learned"]
    d = ll - lu
    z = d.mean() / (d.std(ddof=1) / np.sqrt(len(d)))
    assert lu.mean() > 2.0 and min(ins.values()) > 1000
    assert abs(z) < 3.0, (lu.mean(), ll.mean(), z)
