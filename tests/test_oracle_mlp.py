"""Pin the float64 MLP oracle (oracle/mlp.py) before it checks the cost
model's HIP kernels (CPU only).

* ``forward`` and ``loss_and_grads`` against torch autograd in float64 on an
  independently written torch module (``nn.Linear`` + ReLU + MSE);
* ``loss_and_grads`` against central finite differences of its own loss;
* ``adam`` against ``torch.optim.Adam`` over several steps with changing
  gradients (bias corrections are step dependent);
* planted mistakes (ReLU mask dropped, the 1/M factor dropped, Adam bias
  correction dropped) must fail.
"""
import numpy as np
import pytest
import torch

from oracle import mlp as OM


def _net(rng, F, hidden):
    dims = [F, *hidden, 1]
    Ws = [rng.normal(size=(dims[i + 1], dims[i])) / np.sqrt(dims[i]) for i in range(len(dims) - 1)]
    bs = [0.1 * rng.normal(size=dims[i + 1]) for i in range(len(dims) - 1)]
    return Ws, bs


def _torch_loss(Ws, bs, X, t):
    layers = []
    for i, (W, b) in enumerate(zip(Ws, bs)):
        lin = torch.nn.Linear(W.shape[1], W.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(W))
            lin.bias.copy_(torch.tensor(b))
        layers.append(lin)
        if i < len(Ws) - 1:
            layers.append(torch.nn.ReLU())
    net = torch.nn.Sequential(*layers)
    pred = net(torch.tensor(X)).squeeze(-1)
    loss = torch.nn.functional.mse_loss(pred, torch.tensor(t))
    loss.backward()
    lins = [m for m in net if isinstance(m, torch.nn.Linear)]
    return (pred.detach().numpy(), float(loss.detach()), [m.weight.grad.numpy() for m in lins],
            [m.bias.grad.numpy() for m in lins])


@pytest.mark.parametrize("seed,hidden", [(0, (19, 13)), (1, (32, 32)), (2, (7,))])
def test_mlp_forward_and_grads_vs_torch_autograd(seed, hidden):
    rng = np.random.default_rng(seed)
    F, M = 11, 37
    Ws, bs = _net(rng, F, hidden)
    X = rng.normal(size=(M, F))
    t = rng.normal(size=M)
    pred, _ = OM.forward(Ws, bs, X)
    loss, gW, gb = OM.loss_and_grads(Ws, bs, X, t)
    tp, tl, tgW, tgb = _torch_loss(Ws, bs, X, t)
    assert np.allclose(pred, tp, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) < 1e-12
    for a, b in zip(gW + gb, tgW + tgb):
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12)


def test_mlp_grads_vs_finite_differences():
    rng = np.random.default_rng(5)
    F, M = 6, 20
    Ws, bs = _net(rng, F, (9, 5))
    X = rng.normal(size=(M, F))
    t = rng.normal(size=M)
    _, gW, gb = OM.loss_and_grads(Ws, bs, X, t)
    h = 1e-6
    for li in range(len(Ws)):
        for arr, g in ((Ws[li], gW[li]), (bs[li], gb[li])):
            flat = arr.reshape(-1)
            for j in rng.choice(flat.size, min(6, flat.size), replace=False):
                v0 = flat[j]
                flat[j] = v0 + h
                lp = OM.loss_and_grads(Ws, bs, X, t)[0]
                flat[j] = v0 - h
                lm = OM.loss_and_grads(Ws, bs, X, t)[0]
                flat[j] = v0
                assert abs((lp - lm) / (2 * h) - g.reshape(-1)[j]) < 1e-7


def test_mlp_planted_mistakes_fail():
    rng = np.random.default_rng(8)
    Ws, bs = _net(rng, 5, (8,))
    X = rng.normal(size=(16, 5))
    t = rng.normal(size=16)
    _, tl, tgW, tgb = _torch_loss(Ws, bs, X, t)

    def no_mask(Ws, bs, X, t):          # ReLU derivative dropped
        pred, acts = OM.forward(Ws, bs, X)
        dz = (2.0 * (pred - t) / X.shape[0])[:, None]
        gW = [None, dz.T @ acts[1]]
        dz = dz @ Ws[1]
        gW[0] = dz.T @ acts[0]
        return gW

    def no_mean(Ws, bs, X, t):          # 1/M factor dropped
        _, gW, _ = OM.loss_and_grads(Ws, bs, X, t)
        return [g * X.shape[0] for g in gW]

    for bad in (no_mask, no_mean):
        gW = bad(Ws, bs, X, t)
        assert not all(np.allclose(a, b, rtol=1e-6) for a, b in zip(gW, tgW)), bad.__name__


def test_adam_vs_torch_optim():
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=(4, 3))
    p = p0.copy()
    m, v = np.zeros_like(p), np.zeros_like(p)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=3e-2, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 8):
        g = rng.normal(size=p.shape) * (1.0 + step)
        OM.adam(p, g, m, v, step, 3e-2)
        tp.grad = torch.tensor(g)
        opt.step()
        assert np.allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14), step
    # without bias correction the first steps differ
    q = p0.copy()
    mq, vq = np.zeros_like(q), np.zeros_like(q)
    g = rng.normal(size=q.shape)
    mq[:] = 0.1 * g
    vq[:] = 0.001 * g * g
    q -= 3e-2 * mq / (np.sqrt(vq) + 1e-8)
    p1 = p0.copy()
    OM.adam(p1, g, np.zeros_like(p1), np.zeros_like(p1), 1, 3e-2)
    assert not np.allclose(q, p1)
