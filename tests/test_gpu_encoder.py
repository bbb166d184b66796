"""GPU: encoders and gate matrices built on the GPU from device parameters.

* every device-built gate family (RX, RY, RZ, PhaseShift, U1, U2, U3) with
  per-row parameters, inverse and not, per-gate and fused plans: states vs
  the complex128 oracle, 1e-5; the plan must hold device-built ops;
* the reference MNIST layout (encoder "4x4_u3_h_rx" + trainable ansatz,
  examples/mnist/mnist.py:82-135) from an input tensor with row and column
  strides: expectations vs the oracle (1e-5), the gradient w.r.t. the input
  (encoder parameters trainable) and w.r.t. the ansatz vs central FD of
  the oracle (2e-4 + 1e-3 rel);
* a density-matrix device fed by an encoder (the conjugated copy of each gate).
"""
import numpy as np
import pytest
import torch

from oracle import statevector as O

gpu = pytest.mark.gpu
DEV = "cuda"

FAMILIES = {"rx": 1, "ry": 1, "rz": 1, "phaseshift": 1, "u1": 1, "u2": 2, "u3": 3}


@gpu
@pytest.mark.parametrize("mode", ["pergate", "fused"])
def test_device_built_gates_vs_oracle(mode):
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E
    rng = np.random.default_rng(31)
    n, bsz = 6, 5
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    q.engine_options = tq.EngineOptions(mode=mode)
    q.set_states(torch.tensor(init, dtype=torch.complex64, device=DEV))
    ops = []
    for rep in range(3):
        for name, P in FAMILIES.items():
            w = int(rng.integers(n))
            x = rng.uniform(-np.pi, np.pi, size=(bsz, P))
            inv = bool(rep == 1)
            tqf.func_name_dict[name](q, wires=w, params=torch.tensor(x, dtype=torch.float32, device=DEV),
                                     inverse=inv)
            ops.append(dict(name=name, wires=[w], params=x, inverse=inv))
        tqf.cnot(q, wires=[int(rng.integers(n // 2)), n - 1 - int(rng.integers(n // 2))])
        ops.append(dict(name="cnot", wires=list(q._tape[-1].wires), params=None, inverse=False))
    c = E.Compiled(q._tape, n, bsz, q.engine_options, "cpu")      # the plan, built on the host only
    param_ops = {i for i, op in enumerate(q._tape) if op.params is not None}
    assert param_ops <= c.dev_tape and c.host_static
    got = q.get_states_1d().detach().cpu().numpy()
    ref = O.run_ops(ops, n, bsz, state=init)
    assert np.abs(got - ref).max() < 1e-5


def _mnist_like(bsz, seed=0):
    import torchquantum_amd as tq

    class QModel(tq.QuantumModule):
        def __init__(self):
            super().__init__()
            self.n_wires = 4
            self.encoder = tq.GeneralEncoder(tq.encoder_op_list_name_dict["4x4_u3_h_rx"])
            self.q_layer = tq.RXYZCXLayer0({"n_wires": 4, "n_blocks": 2})
            self.measure = tq.MeasureAll(tq.PauliZ)

        def forward(self, x):
            qdev = tq.QuantumDevice(n_wires=4, bsz=x.shape[0], device=x.device)
            self.encoder(qdev, x)
            self.q_layer(qdev)
            return self.measure(qdev)

    torch.manual_seed(seed)
    return QModel().to(DEV)


def _mnist_oracle_ops(model, x, theta=None):
    """Oracle op list of the encoder (per-row params from x) + ansatz; the
    ansatz parameters come from ``theta`` (flat, module order) when given."""
    import torchquantum_amd as tq
    ops = []
    for info in tq.encoder_op_list_name_dict["4x4_u3_h_rx"]:
        idx = info.get("input_idx")
        ops.append(dict(name=info["func"], wires=info["wires"], inverse=False,
                        params=None if idx is None else x[:, idx]))
    k = 0
    for layer in model.q_layer.layers_all:
        if isinstance(layer, tq.Op1QAllLayer):
            for w in range(layer.n_wires):
                o = layer.ops_all[w]
                p = o.params.detach().cpu().numpy().astype(np.float64)[0] if theta is None else theta[k:k + 1]
                k += 1
                ops.append(dict(name=o.op_name, wires=[w], inverse=False, params=p))
        else:
            for w in range(layer.n_wires):
                ops.append(dict(name="cnot", wires=[w, (w + 1) % layer.n_wires], params=None, inverse=False))
    return ops


@gpu
@pytest.mark.parametrize("mode", ["fused", "pergate"])
def test_encoder_ansatz_vs_oracle_and_fd(mode):
    """Expectations 1e-5; d(sum w * <Z>)/dx and /dtheta vs central FD of the
    float64 oracle (h = 1e-4), 2e-4 + 1e-3 rel, on seeded entries."""
    import torchquantum_amd as tq
    bsz = 6
    model = _mnist_like(bsz)
    rng = np.random.default_rng(8)
    # a strided input: columns 1, 3, 5, ... of a wider leaf tensor
    wide = torch.tensor(rng.uniform(-2, 2, size=(bsz, 33)), dtype=torch.float32, device=DEV, requires_grad=True)
    x = wide[:, 1::2]
    old = tq.get_engine_options()
    tq.set_engine_options(mode=mode)
    try:
        out = model(x)
        xnp = x.detach().cpu().numpy().astype(np.float64)
        ops = _mnist_oracle_ops(model, xnp)
        st = O.run_ops(ops, 4, bsz)
        assert np.abs(out.detach().cpu().numpy() - O.expval_z_all(st, 4)).max() < 1e-5
        w = torch.tensor(rng.normal(size=(bsz, 4)), dtype=torch.float32, device=DEV)
        (out * w).sum().backward()
        wn = w.cpu().numpy().astype(np.float64)
        params = list(model.q_layer.parameters())
        theta0 = np.concatenate([p.detach().cpu().numpy().reshape(-1) for p in params]).astype(np.float64)

        def loss(xv, th):
            return float((O.expval_z_all(O.run_ops(_mnist_oracle_ops(model, xv, th), 4, bsz), 4) * wn).sum())

        pick = np.random.default_rng(2).choice(bsz * 16, 20, replace=False)
        fd = O.central_fd(lambda v: loss(v.reshape(bsz, 16), theta0), xnp.reshape(-1).copy(), h=1e-4, idx=pick)
        gx = wide.grad.detach().cpu().numpy()[:, 1::2].reshape(-1)
        assert np.all(np.abs(gx[pick] - fd[pick]) <= 2e-4 + 1e-3 * np.abs(fd[pick])), np.abs(gx[pick] - fd[pick]).max()
        assert np.all(wide.grad.detach().cpu().numpy()[:, 0::2] == 0)
        gp = np.concatenate([p.grad.detach().cpu().numpy().reshape(-1) for p in params])
        pick_p = np.random.default_rng(3).choice(theta0.size, 12, replace=False)
        fdp = O.central_fd(lambda th: loss(xnp, th), theta0.copy(), h=1e-4, idx=pick_p)
        assert np.all(np.abs(gp[pick_p] - fdp[pick_p]) <= 2e-4 + 1e-3 * np.abs(fdp[pick_p]))
    finally:
        tq.set_engine_options(**old.__dict__)


@gpu
def test_encoder_on_density_device():
    """NoiseDevice fed by a GeneralEncoder: rho vs |psi><psi| of the oracle (1e-5)."""
    import torchquantum_amd as tq
    bsz, n = 3, 3
    rng = np.random.default_rng(4)
    x = torch.tensor(rng.uniform(-2, 2, size=(bsz, 8)), dtype=torch.float32, device=DEV)
    enc = tq.GeneralEncoder([{"input_idx": [0], "func": "ry", "wires": [0]},
                             {"input_idx": [1, 2], "func": "u2", "wires": [1]},
                             {"input_idx": [3], "func": "rz", "wires": [2]},
                             {"func": "cnot", "wires": [0, 2]},
                             {"input_idx": [4, 5, 6], "func": "u3", "wires": [2]},
                             {"input_idx": [7], "func": "phaseshift", "wires": [1]}])
    q = tq.NoiseDevice(n_wires=n, bsz=bsz, device=DEV)
    enc(q, x)
    rho = q.matrices().detach().cpu().numpy()
    xn = x.cpu().numpy().astype(np.float64)
    ops = [dict(name="ry", wires=[0], params=xn[:, [0]], inverse=False),
           dict(name="u2", wires=[1], params=xn[:, [1, 2]], inverse=False),
           dict(name="rz", wires=[2], params=xn[:, [3]], inverse=False),
           dict(name="cnot", wires=[0, 2], params=None, inverse=False),
           dict(name="u3", wires=[2], params=xn[:, [4, 5, 6]], inverse=False),
           dict(name="phaseshift", wires=[1], params=xn[:, [7]], inverse=False)]
    st = O.run_ops(ops, n, bsz)
    ref = np.einsum("bi,bj->bij", st, st.conj())
    assert np.abs(rho.reshape(bsz, 2 ** n, 2 ** n) - ref).max() < 1e-5


@gpu
def test_phase_encoders_on_strided_inputs():
    """MultiPhaseEncoder / PhaseEncoder reading column slices and single
    columns of a strided tensor (row and column strides != 1): states vs the
    oracle, 1e-5."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    bsz, n = 4, 3
    rng = np.random.default_rng(6)
    wide = torch.tensor(rng.uniform(-2, 2, size=(bsz, 40)), dtype=torch.float32, device=DEV)
    x = wide[:, 2::3]                          # [bsz, 13], strides (40, 3)
    funcs = ["u3", "rx", "u2", "ry", "rz", "phaseshift"]
    enc = tq.MultiPhaseEncoder(funcs, wires=[0, 1, 2, 0, 1, 2])
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    enc(q, x)
    tq.PhaseEncoder(tqf.ry)(q, x[:, 5:])
    xn = x.cpu().numpy().astype(np.float64)
    ops, col = [], 0
    for f, w in zip(funcs, [0, 1, 2, 0, 1, 2]):
        st = {"u3": 3, "u2": 2}.get(f, 1)
        ops.append(dict(name=f, wires=[w], params=xn[:, col:col + st], inverse=False))
        col += st
    for k in range(n):
        ops.append(dict(name="ry", wires=[k], params=xn[:, 5 + k:6 + k], inverse=False))
    ref = O.run_ops(ops, n, bsz)
    assert np.abs(q.get_states_1d().cpu().numpy() - ref).max() < 1e-5


@gpu
def test_device_built_merged_runs_and_gradients():
    """Merged single-wire runs built on the GPU: fixed gates (H, SX, Y, SH)
    between shared trainable and per-row parameterised gates (RX, RY, RZ,
    U2, U3, Rot), inverse flags; the plan must hold device-built runs; states
    vs the oracle (1e-5) and gradients of every parameter vs central FD of
    the oracle (2e-4 + 1e-3 rel)."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E
    rng = np.random.default_rng(17)
    n, bsz = 5, 3
    seq = [("hadamard", 0), ("rx", 1), ("u3", bsz), ("sx", 0), ("rot", 1), ("ry", 1), ("pauliy", 0),
           ("rz", bsz), ("u2", 1), ("shadamard", 0)]
    ops, tens = [], []
    for w in range(n):
        for j, (name, rows) in enumerate(seq[w % 3:w % 3 + 6]):
            npar = FAMILIES.get(name, 3 if name == "rot" else 0)
            p = rng.uniform(-np.pi, np.pi, size=(rows, npar)) if npar else None
            ops.append(dict(name=name, wires=[w], params=p, inverse=bool((j + w) % 4 == 1)))
        ops.append(dict(name="cnot", wires=[w, (w + 1) % n], params=None, inverse=False))
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    for op in ops:
        p = op["params"]
        t = None
        if p is not None:
            t = torch.tensor(p, dtype=torch.float32, device=DEV, requires_grad=True)
            tens.append(t)
        tqf.func_name_dict[op["name"]](q, wires=op["wires"], params=t, inverse=op["inverse"])
    c = E.Compiled(q._tape, n, bsz, q._options(), "cpu")
    runs = [c.kops[i] for i in c.dev_kops if len(c.kops[i].src) > 1]
    assert runs and c.host_static
    obs, coef = ["ZXIYZ", "XXIII", "IIZZY"], [1.0, -0.5, 0.7]
    val = tq.expval_pauli_sum(q, obs, coef)
    st = q.get_states_1d().detach().cpu().numpy()
    ref = O.run_ops([dict(o, params=(None if o["params"] is None else
                                     (o["params"] if o["params"].shape[0] > 1 else o["params"][0]))) for o in ops],
                    n, bsz)
    assert np.abs(st - ref).max() < 1e-5
    val.sum().backward()
    got = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in tens])
    flat = np.concatenate([o["params"].reshape(-1) for o in ops if o["params"] is not None])

    def f(v):
        k, o2 = 0, []
        for o in ops:
            p = o["params"]
            if p is not None:
                p = v[k:k + p.size].reshape(p.shape)
                k += p.size
                p = p if p.shape[0] > 1 else p[0]
            o2.append(dict(o, params=p))
        s = O.run_ops(o2, n, bsz)
        return float(sum(cf * O.expval_pauli(s, ob).sum() for ob, cf in zip(obs, coef)))

    fd = O.central_fd(f, flat.copy(), h=1e-4)
    assert np.all(np.abs(got - fd) <= 2e-4 + 1e-3 * np.abs(fd)), np.abs(got - fd).max()
