"""Learned GCMC insertion proposal (no counterpart in the reference).

A small MLP (32 -> 128 -> 128 -> 128 -> 1, SiLU, BF16 weights, FP32
accumulation) maps a voxel's 32-bin radial host-density histogram to a logit;
the per-chain proposal q over the fractional voxel grid is
(1 - mix) softmax(logit / T) + mix / nvox.  Insertions draw a voxel from q and
a uniform point inside it; the acceptance adds -ln(q nvox) for insertions and
+ln(q nvox) for deletions (the reverse move), so the chain keeps the grand-
canonical distribution for any network, and the uniform mixture keeps every
voxel reachable.  The exact LJ + Ewald Delta E decides acceptance as before.

Training data are the insertion trials of the plain (uniform) kernel:
(voxel, y = -dE/kT) pairs recorded by kupsmi_gcmc_run.  In dense hosts y spans
many decades (a site overlap gives y ~ -1e7, a pore gives y ~ +5), and only a
handful in 1e6 trials have y > 0, so the network regresses the compressed
exponent sign(y) ln(1 + |y|) by least squares with Adam: it ranks voxels by how
badly an insertion there overlaps the host, which is a dense signal, and the
low proposal temperature turns that ranking into a sharp proposal.  Everything
runs on the GPU: descriptors, inference on BF16 MFMA, forward/backward, Adam.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib

N_IN, N_H = 32, 128
N_PARAMS = N_IN * N_H + N_H + 2 * (N_H * N_H + N_H) + N_H + 1


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation of the flat parameter vector (FP32)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(1.0 / N_H), N_H), np.zeros(1)]
    p = np.concatenate([x.ravel() for x in parts]).astype(np.float32)
    assert p.size == N_PARAMS
    return p


@dataclass
class InsertionSamples:
    """Device buffer of (descriptor row, -dE/kT) insertion samples."""

    capacity: int
    nv: int
    device: torch.device

    def __post_init__(self):
        self.count = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.rows = torch.zeros(self.capacity, dtype=torch.int32, device=self.device)
        self.y = torch.zeros(self.capacity, dtype=torch.float32, device=self.device)

    @property
    def n(self) -> int:
        return min(int(self.count.item()), self.capacity)


class LearnedProposal:
    def __init__(self, gcmc, nv: int = 32, rmax: float = 8.0, rho0: float = 0.1,
                 mix: float = 0.1, temperature: float = 0.1, seed: int = 0):
        self.gcmc = gcmc
        self.nv = int(nv)
        self.nvox = self.nv**3
        self.mix = float(mix)
        self.temperature = float(temperature)
        dev = gcmc.dev
        self.dev = dev
        lib = _lib.load()
        assert lib.kupsmi_mlp_param_count() == N_PARAMS
        n = gcmc.n_chain
        self.desc = torch.empty((n * self.nvox, N_IN), dtype=torch.int16, device=dev)
        d = gcmc.d
        _lib.check(lib.kupsmi_voxel_descriptors(
            _lib.stream_ptr(), n, self.nv, gcmc.max_host, _lib.ptr(d["host_off"]),
            _lib.ptr(d["host"]), _lib.ptr(d["tril"]), float(rmax), float(rho0), _lib.ptr(self.desc)))
        self.master = torch.as_tensor(init_params(seed), device=dev)
        self.m1 = torch.zeros_like(self.master)
        self.m2 = torch.zeros_like(self.master)
        self.wq = torch.empty_like(self.master)
        self.wbf = torch.empty(N_PARAMS, dtype=torch.int16, device=dev)
        self.adam_step = 0
        self.round()
        self.cdf = torch.empty((n, self.nvox), dtype=torch.float32, device=dev)
        self.guide = torch.empty((n, self.nvox), dtype=torch.int32, device=dev)
        self.logits = torch.empty(n * self.nvox, dtype=torch.float32, device=dev)
        self.build_table(uniform=True)

    def round(self):
        _lib.check(_lib.load().kupsmi_mlp_round_weights(
            _lib.stream_ptr(), _lib.ptr(self.master), _lib.ptr(self.wq), _lib.ptr(self.wbf)))

    def infer(self, rows: torch.Tensor | None = None) -> torch.Tensor:
        """MLP logits of all voxels (or of descriptor rows ``rows``)."""
        desc = self.desc if rows is None else self.desc[rows.long()].contiguous()
        out = torch.empty(desc.shape[0], dtype=torch.float32, device=self.dev)
        _lib.check(_lib.load().kupsmi_mlp_infer(
            _lib.stream_ptr(), desc.shape[0], _lib.ptr(desc), _lib.ptr(self.wbf), _lib.ptr(self.wq),
            _lib.ptr(out)))
        return out

    def train_step(self, rows: torch.Tensor, y: torch.Tensor, lr: float, apply: bool = True,
                   grad: torch.Tensor | None = None) -> float:
        rows = rows.to(torch.int32).contiguous()
        y = y.to(torch.float32).contiguous()
        n_part = (rows.numel() + 31) // 32
        loss = torch.zeros(n_part, dtype=torch.float32, device=self.dev)
        if apply:
            self.adam_step += 1
        _lib.check(_lib.load().kupsmi_mlp_train_step(
            _lib.ctx(), _lib.stream_ptr(), rows.numel(), _lib.ptr(rows), _lib.ptr(y),
            _lib.ptr(self.desc), _lib.ptr(self.master), _lib.ptr(self.m1), _lib.ptr(self.m2),
            _lib.ptr(self.wq), _lib.ptr(self.wbf), float(lr), max(self.adam_step, 1), int(apply),
            _lib.ptr(grad), _lib.ptr(loss)))
        return loss

    @staticmethod
    def compress(y: torch.Tensor) -> torch.Tensor:
        """Training target of a Boltzmann exponent y: sign(y) ln(1 + |y|)."""
        return torch.sign(y) * torch.log1p(y.abs())

    def train(self, samples: InsertionSamples, epochs: int = 8, batch: int = 8192,
              lr: float = 1e-3, seed: int = 0) -> list[float]:
        """Adam epochs over the recorded samples; returns the mean loss per epoch."""
        n = samples.n
        rows = samples.rows[:n]
        y = self.compress(samples.y[:n])
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        history = []
        for _ in range(epochs):
            perm = torch.randperm(n, device=self.dev, generator=gen)
            tot = torch.zeros((), dtype=torch.float64, device=self.dev)
            for a in range(0, n, batch):
                idx = perm[a:a + batch]
                loss = self.train_step(rows[idx], y[idx], lr)
                tot += loss.double().sum()
            history.append(float(tot.item()) / n)
        return history

    def build_table(self, uniform: bool = False):
        """Proposal cdf of every chain from the current network (or uniform)."""
        logits = torch.zeros_like(self.logits) if uniform else self.infer()
        self.logits.copy_(logits)
        _lib.check(_lib.load().kupsmi_proposal_table(
            _lib.stream_ptr(), self.gcmc.n_chain, self.nvox, _lib.ptr(self.logits),
            self.temperature, self.mix, _lib.ptr(self.cdf), _lib.ptr(self.guide)))
