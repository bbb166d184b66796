"""GPU: tq.CircuitReplay (a circuit recorded once and replayed) against the
per-call functional path and the float64 oracle.

* the 10-qubit bsz-32 U3 + CNOT-ring circuit: <Z_i> of the replay equals the
  functional path and the oracle (|d| <= 1e-5), and follows in-place
  parameter updates;
* an encoder input replaced per call (per-row RY angles) and trainable
  RX/RY/RZ: outputs and adjoint gradients equal the functional path's
  (|d| <= 1e-5 (1 + |g|));
* argument errors.
"""
import numpy as np
import pytest
import torch

from circuits import config1_ops
from oracle import statevector as O

gpu = pytest.mark.gpu
DEV = "cuda"


def test_replay_needs_gpu():
    import torchquantum_amd as tq
    with pytest.raises(NotImplementedError):
        tq.CircuitReplay(lambda q: None, 4, 1, device="cpu")


@gpu
def test_replay_matches_functional_and_oracle():
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    n, bsz = 10, 32
    ops = config1_ops(seed=0)
    params = [None if op["params"] is None else torch.tensor(op["params"], dtype=torch.float32, device=DEV)
              for op in ops]

    def record(q):
        for op, p in zip(ops, params):
            tqf.func_name_dict[op["name"]](q, op["wires"], params=p)

    rep = tq.CircuitReplay(record, n, bsz)
    meas = tq.MeasureAll(tq.PauliZ)
    for step in range(3):
        got = meas(rep()).double().cpu().numpy()
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        record(q)
        ref = meas(q).double().cpu().numpy()
        cur = [dict(op, params=None if p is None else p.double().cpu().numpy()) for op, p in zip(ops, params)]
        st = O.run_ops(cur, n, bsz)
        orc = O.expval_z_all(st, n)
        assert np.abs(got - ref).max() <= 1e-5, step
        assert np.abs(got - orc).max() <= 1e-5, step
        with torch.no_grad():                   # an optimizer-like in-place update between replays
            for p in params:
                if p is not None:
                    p.add_(0.05 * (step + 1))


@gpu
def test_replay_inputs_and_gradients():
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    n, bsz = 11, 6
    rng = np.random.default_rng(3)
    theta = [torch.tensor([[float(t)]], device=DEV, requires_grad=True) for t in rng.normal(0, 0.5, 3 * n)]

    def record(q, x):
        for w in range(n):
            tqf.ry(q, w, params=x[:, w:w + 1])
        k = 0
        for w in range(n):
            for g in ("rx", "ry", "rz"):
                tqf.func_name_dict[g](q, w, params=theta[k])
                k += 1
        for w in range(n):
            tqf.cnot(q, [w, (w + 1) % n])

    x0 = torch.rand(bsz, n, device=DEV)
    rep = tq.CircuitReplay(record, n, bsz, example_inputs=(x0,))
    obs, coef = ["Z" * 2 + "I" * (n - 2), "I" * (n - 1) + "X"], [1.0, 0.5]
    for seed in (1, 2):
        x = torch.rand(bsz, n, device=DEV, generator=torch.Generator(device=DEV).manual_seed(seed))
        for p in theta:
            p.grad = None
        loss = tq.expval_pauli_sum(rep(x), obs, coef).mean()
        loss.backward()
        g_rep = torch.cat([p.grad.reshape(-1) for p in theta]).double().cpu().numpy()
        for p in theta:
            p.grad = None
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        record(q, x)
        loss_ref = tq.expval_pauli_sum(q, obs, coef).mean()
        loss_ref.backward()
        g_ref = torch.cat([p.grad.reshape(-1) for p in theta]).double().cpu().numpy()
        assert abs(float(loss) - float(loss_ref)) <= 1e-5
        assert np.all(np.abs(g_rep - g_ref) <= 1e-5 * (1 + np.abs(g_ref))), float(np.abs(g_rep - g_ref).max())
    with pytest.raises(ValueError):
        rep()
    with pytest.raises(ValueError):
        rep(torch.rand(bsz, n + 1, device=DEV))


@gpu
def test_replay_general_encoder_training():
    """GeneralEncoder (its input columns are copies: advanced indexing) and a
    trainable RXYZCXLayer0: the replay re-gathers the encoder's columns on
    every call, so for new inputs its <Z_i> and gradients equal the
    functional path's (|d| <= 1e-5 (1 + |g|)), and an Adam step on the
    ansatz between calls is followed."""
    import torchquantum_amd as tq
    n, bsz = 10, 8
    enc = tq.GeneralEncoder([{"input_idx": [(3 * i) % n], "func": "ry", "wires": [i]} for i in range(n)])
    ansatz = tq.RXYZCXLayer0({"n_wires": n, "n_blocks": 2}).to(DEV)
    opt = torch.optim.Adam(ansatz.parameters(), lr=0.05)

    def record(q, x):
        enc(q, x)
        ansatz(q)

    rep = tq.CircuitReplay(record, n, bsz, example_inputs=(torch.rand(bsz, n, device=DEV),))
    assert rep._gathers, "encoder columns were not bound"
    meas = tq.MeasureAll(tq.PauliZ)
    w = torch.linspace(-1.0, 1.0, n, device=DEV)
    for seed in (1, 2, 3):
        x = torch.rand(bsz, n, device=DEV, generator=torch.Generator(device=DEV).manual_seed(seed)) * np.pi
        opt.zero_grad(set_to_none=True)
        out = meas(rep(x))
        (out * w).sum().backward()
        g_rep = torch.cat([p.grad.reshape(-1) for p in ansatz.parameters()]).double().cpu().numpy()
        opt.zero_grad(set_to_none=True)
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        record(q, x)
        ref = meas(q)
        (ref * w).sum().backward()
        g_ref = torch.cat([p.grad.reshape(-1) for p in ansatz.parameters()]).double().cpu().numpy()
        assert (out - ref).abs().max().item() <= 1e-5, seed
        assert np.all(np.abs(g_rep - g_ref) <= 1e-5 * (1 + np.abs(g_ref))), float(np.abs(g_rep - g_ref).max())
        opt.step()
