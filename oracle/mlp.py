"""float64 numpy MLP (TEST INFRASTRUCTURE ONLY): the CPU oracle for the cost
model's HIP forward/backward/Adam kernels.  y = W_L relu(... relu(W_1 x + b_1)) + b_L,
loss = mean (y - t)^2."""
from __future__ import annotations

import numpy as np


def forward(Ws, bs, X):
    acts = [X]
    h = X
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ W.T + b
        if i < len(Ws) - 1:
            z = np.maximum(z, 0.0)
        acts.append(z)
        h = z
    return h[:, 0], acts


def loss_and_grads(Ws, bs, X, t):
    pred, acts = forward(Ws, bs, X)
    M = X.shape[0]
    loss = float(np.mean((pred - t) ** 2))
    dz = (2.0 * (pred - t) / M)[:, None]
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for i in range(len(Ws) - 1, -1, -1):
        if i < len(Ws) - 1:
            dz = dz * (acts[i + 1] > 0)
        gW[i] = dz.T @ acts[i]
        gb[i] = dz.sum(0)
        dz = dz @ Ws[i]
    return loss, gW, gb


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
