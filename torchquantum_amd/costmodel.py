"""Learned cost model for fusion plans: an MLP from plan features to
predicted log-milliseconds per amplitude, trained and evaluated on the GPU with this
package's own HIP kernels (tqsv_mlp.hip: GEMM + bias + ReLU, ReLU backward,
column sums, MSE gradient, Adam).  torch only holds the buffers.

Plan features are computed on the host from a compiled plan (no launch), so
the planner can score every candidate option set cheaply and run only the
predicted best; exhaustive timing (``autotune.Autotuner``) is the fallback
and the reference for the prediction error.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from . import engine as E

FEATURES = ["log2_amps", "n_qubits", "log2_bsz", "stages", "solos", "ops", "mat1", "mat2", "matk", "diag",
            "swap", "remaps", "ops_per_stage", "max_ops_stage", "stage_bits", "fusion_width", "log_bytes",
            "log_flops", "launches"]

_OP_FLOPS = {"mat1": 16.0, "mat2": 32.0, "matk": 128.0, "diag": 8.0, "swap": 2.0}


def plan_features(c: E.Compiled) -> np.ndarray:
    n, bsz = c.n, c.bsz
    amps = float(bsz) * (1 << n)
    stages = [st for kind, (st) in c.steps if kind == "stage"]
    solos = sum(1 for kind, _ in c.steps if kind == "solo")
    cnt = {k: 0 for k in _OP_FLOPS}
    for k in c.kops:
        cnt[k.kind] = cnt.get(k.kind, 0) + 1
    remaps = 0
    per_stage = []
    for _, st in stages:
        h = st.words
        nops = int(h[E.H_NOPS])
        off = int(h[E.H_OFF_OPS])
        r = sum(1 for q in range(nops) if (h[off + 8 * q] & 15) == E.OP_REMAP)
        remaps += r
        per_stage.append(nops - r)
    ns = max(len(stages), 1)
    passes = len(stages) + solos
    flops = sum(_OP_FLOPS[k.kind] for k in c.kops) * amps + remaps * 4.0 * amps
    return np.array([
        math.log2(amps), n, math.log2(max(bsz, 1)), len(stages), solos, len(c.kops), cnt["mat1"], cnt["mat2"],
        cnt["matk"], cnt["diag"], cnt["swap"], remaps, sum(per_stage) / ns, max(per_stage) if per_stage else 0,
        (stages[0][1].m if stages else 0), c.opts.fusion_width, math.log2(16.0 * amps * max(passes, 1)),
        math.log2(max(flops, 1.0)), passes,
    ], dtype=np.float64)


class MLPCostModel:
    """Fully connected ReLU network, fp32, trained with Adam on the GPU."""

    def __init__(self, n_features: int = len(FEATURES), hidden: Sequence[int] = (64, 64), device="cuda",
                 seed: int = 0):
        self.device = torch.device(device)
        sizes = [n_features] + list(hidden) + [1]
        g = torch.Generator().manual_seed(seed)
        self.W: List[torch.Tensor] = []
        self.b: List[torch.Tensor] = []
        for i in range(len(sizes) - 1):
            bound = math.sqrt(6.0 / sizes[i])             # He-uniform for ReLU layers
            self.W.append((torch.rand(sizes[i + 1], sizes[i], generator=g) * 2 - 1).mul_(bound).to(self.device))
            self.b.append(torch.zeros(sizes[i + 1], device=self.device))
        self.m = [torch.zeros_like(p) for p in self.W + self.b]
        self.v = [torch.zeros_like(p) for p in self.W + self.b]
        self.t = 0
        self.mu = None
        self.sd = None
        self.x_min = None
        self.x_max = None
        self.y_mu = 0.0
        self.y_sd = 1.0

    # ------------------------------------------------------------ kernels
    def _stream(self):
        return N.stream_of(self.W[0])

    def _gemm(self, M, Nn, K, A, lda, ta, B, ldb, tb, C, bias=None, relu=False, beta=0.0):
        N.call("tqsv_gemm", self._stream(), M, Nn, K, 1.0, A.data_ptr(), lda, ta, B.data_ptr(), ldb, tb, beta,
               C.data_ptr(), Nn, bias.data_ptr() if bias is not None else None, 1 if relu else 0)

    def forward(self, X: torch.Tensor):
        """X [M, F] (standardised) -> pred [M], activations."""
        M = X.shape[0]
        acts = [X.contiguous()]
        h = acts[0]
        for i, (W, b) in enumerate(zip(self.W, self.b)):
            out = torch.empty(M, W.shape[0], device=self.device)
            self._gemm(M, W.shape[0], W.shape[1], h, h.shape[1], 0, W, W.shape[1], 1, out, bias=b,
                       relu=i < len(self.W) - 1)
            acts.append(out)
            h = out
        return h[:, 0], acts

    def grads(self, X: torch.Tensor, y: torch.Tensor):
        """MSE loss and parameter gradients (lists, same shapes as W, b)."""
        M = X.shape[0]
        pred, acts = self.forward(X)
        dz = torch.empty(M, 1, device=self.device)
        loss = torch.empty(1, device=self.device)
        s = self._stream()
        N.call("tqsv_mse_grad", s, pred.contiguous().data_ptr(), y.contiguous().data_ptr(), M, dz.data_ptr(),
               loss.data_ptr())
        gW = [None] * len(self.W)
        gb = [None] * len(self.W)
        for i in range(len(self.W) - 1, -1, -1):
            W = self.W[i]
            out_f, in_f = W.shape
            if i < len(self.W) - 1:
                N.call("tqsv_relu_backward", s, dz.data_ptr(), acts[i + 1].data_ptr(), dz.numel())
            gw = torch.empty(out_f, in_f, device=self.device)
            # dW = dz^T [out, M] x a_i [M, in]
            self._gemm(out_f, in_f, M, dz, out_f, 1, acts[i], in_f, 0, gw)
            gbi = torch.empty(out_f, device=self.device)
            N.call("tqsv_colsum", s, dz.data_ptr(), M, out_f, gbi.data_ptr())
            gW[i], gb[i] = gw, gbi
            if i > 0:
                dh = torch.empty(M, in_f, device=self.device)
                # dh = dz [M, out] x W [out, in]
                self._gemm(M, in_f, out_f, dz, out_f, 0, W, in_f, 0, dh)
                dz = dh
        return loss, gW, gb

    def adam_step(self, gW, gb, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        c1, c2 = 1 - b1 ** self.t, 1 - b2 ** self.t
        s = self._stream()
        for p, g, m, v in zip(self.W + self.b, gW + gb, self.m, self.v):
            N.call("tqsv_adam_step", s, p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel(), lr, b1,
                   b2, eps, c1, c2)

    # ------------------------------------------------------------ training API
    def _std(self, X: np.ndarray) -> torch.Tensor:
        """Standardised network input.  Features are held to the range seen in
        training: outside it a ReLU network extrapolates linearly without
        bound, while the cost per amplitude of a plan changes slowly with its
        size (the size itself enters through the per-amplitude target)."""
        X = np.asarray(X, dtype=np.float64)
        if self.x_min is not None:
            X = np.clip(X, self.x_min, self.x_max)
        Z = (X - self.mu) / self.sd
        return torch.tensor(Z, dtype=torch.float32, device=self.device)

    @staticmethod
    def _log_amps(X: np.ndarray) -> np.ndarray:
        """ln(amplitudes) of each plan (feature 0 is log2 of the state size)."""
        return np.asarray(X, dtype=np.float64)[:, 0] * math.log(2.0)

    def fit(self, X: np.ndarray, y_ms: np.ndarray, epochs: int = 1500, lr: float = 3e-3, batch: int = 0):
        """Train on log(ms per amplitude) targets -- the time of a plan grows
        with the state it sweeps, so the network learns the cost per amplitude
        and extrapolates to larger states through that factor; returns the
        loss history (host floats)."""
        X = np.asarray(X, dtype=np.float64)
        y = np.log(np.asarray(y_ms, dtype=np.float64)) - self._log_amps(X)
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-6
        self.x_min, self.x_max = X.min(0), X.max(0)
        self.y_mu, self.y_sd = float(y.mean()), float(y.std() + 1e-6)
        Xt = self._std(X)
        yt = torch.tensor((y - self.y_mu) / self.y_sd, dtype=torch.float32, device=self.device)
        hist = []
        rng = np.random.default_rng(0)
        for ep in range(epochs):
            if batch and batch < len(X):
                idx = torch.as_tensor(rng.choice(len(X), batch, replace=False), device=self.device)
                loss, gW, gb = self.grads(Xt[idx], yt[idx])
            else:
                loss, gW, gb = self.grads(Xt, yt)
            self.adam_step(gW, gb, lr=lr)
            if ep % 100 == 0 or ep == epochs - 1:
                hist.append(float(loss))
        return hist

    def predict_ms(self, X: np.ndarray) -> np.ndarray:
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        pred, _ = self.forward(self._std(X))
        return np.exp(pred.detach().cpu().numpy().astype(np.float64) * self.y_sd + self.y_mu + self._log_amps(X))

    def state_dict(self):
        return {"W": [w.cpu().tolist() for w in self.W], "b": [b.cpu().tolist() for b in self.b],
                "mu": self.mu.tolist(), "sd": self.sd.tolist(), "y_mu": self.y_mu, "y_sd": self.y_sd}
