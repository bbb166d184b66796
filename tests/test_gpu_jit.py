"""Run-time compiled stage kernels (EngineOptions(jit="on")) on the GPU.

Each case runs the same circuit three ways: compiled stages, the program-
interpreting stage kernels (jit="off") and the complex128 oracle.  Bars:
amplitudes within 1e-5 max-abs of the oracle and 1e-6 of the interpreting
kernels (same float32 arithmetic, different instruction order); gradients
within 2e-4 + 1e-3 rel of central finite differences of the float64 oracle
and 1e-5 of the interpreting kernels.  Every case also checks that compiled
kernels actually ran.
"""
import numpy as np
import pytest
import torch

from circuits import apply_ops_tq, oracle_ops, random_ops, config3_ops, config3_observables
from oracle import statevector as O

gpu = pytest.mark.gpu
DEV = "cuda"


def _plans_with_jit(kind):
    from torchquantum_amd import engine as E
    return [c for c in E._plan_cache.values() if c.opts.jit == "on" and c._jit.get(kind)]


def _run(n, bsz, ops, **kw):
    import torchquantum_amd as tq
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    q.engine_options = tq.EngineOptions(mode="fused", **kw)
    apply_ops_tq(q, ops, DEV)
    return q.get_states_1d().detach().cpu().numpy()


@gpu
@pytest.mark.parametrize("n,stage_bits,width", [(10, 0, 1), (12, 0, 1), (14, 10, 1), (13, 0, 3), (12, 0, 4),
                                                (14, 13, 1), (13, 11, 2)])
def test_jit_random_circuits(n, stage_bits, width):
    """Random circuits over every gate kind (controls on thread and register
    bits, diagonal phases, dense 2/3/4-bit blocks, per-row parameters)."""
    from torchquantum_amd import engine as E
    E._plan_cache.clear()
    rng = np.random.default_rng(70 + n + 100 * width)
    bsz = 2
    ops = random_ops(n, 80, bsz, rng)
    got = _run(n, bsz, ops, jit="on", stage_bits=stage_bits, fusion_width=width)
    assert _plans_with_jit("f"), "compiled stage kernels did not run"
    interp = _run(n, bsz, ops, jit="off", stage_bits=stage_bits, fusion_width=width)
    ref = O.run_ops(oracle_ops(ops, bsz), n, bsz)
    assert np.abs(got - ref).max() < 1e-5
    assert np.abs(got - interp).max() < 1e-6


@gpu
@pytest.mark.parametrize("adjoint_mg,stage_bits", [(True, 0), (False, 0), (True, 13), (True, 11)])
def test_jit_adjoint_gradients(adjoint_mg, stage_bits):
    """Trainable gates of every parameter kind (1q rotations, controlled
    rotations, 2-wire rotations, diagonal phases), per-row parameters."""
    import torchquantum_amd as tq
    from torchquantum_amd import engine as E
    E._plan_cache.clear()
    rng = np.random.default_rng(5 + int(adjoint_mg) + stage_bits)
    n, bsz = (10, 2) if stage_bits == 0 else (stage_bits + 1, 2)
    names = ["rx", "ry", "rz", "u3", "rot", "cnot", "cz", "crx", "cry", "crz", "rzz", "rxx", "ryy", "rzx", "cu3",
             "u1", "phaseshift", "cu1", "h", "swap", "toffoli", "multirz", "xxplusyy", "singleexcitation", "cu",
             "u2", "r", "crot"]
    ops = random_ops(n, 50, bsz, rng, names=names)
    obs, coef = config3_observables(n)
    grads = {}
    for jit in ("on", "off"):
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        q.engine_options = tq.EngineOptions(jit=jit, adjoint_mg=adjoint_mg, stage_bits=stage_bits)
        ts = apply_ops_tq(q, ops, DEV, requires_grad=True)
        loss = tq.expval_pauli_sum(q, obs, coef).sum()
        loss.backward()
        grads[jit] = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in ts])
    assert _plans_with_jit("a"), "compiled adjoint kernels did not run"

    def f(flat):
        k = 0
        o2 = []
        for op in ops:
            p = op["params"]
            if p is not None and not np.iscomplexobj(p):
                p = flat[k:k + p.size].reshape(p.shape)
                k += p.size
            o2.append(dict(op, params=p))
        st = O.run_ops(oracle_ops(o2, bsz), n, bsz)
        return float(sum(c * O.expval_pauli(st, o).sum() for o, c in zip(obs, coef)))

    flat = np.concatenate([op["params"].reshape(-1) for op in ops
                           if op["params"] is not None and not np.iscomplexobj(op["params"])])
    fd = O.central_fd(f, flat.copy(), h=1e-4)
    got = grads["on"]
    assert np.all(np.abs(got - fd) <= 2e-4 + 1e-3 * np.abs(fd)), np.abs(got - fd).max()
    assert np.abs(got - grads["off"]).max() < 1e-5


@gpu
def test_jit_config3_family_gradient():
    """The hardware-efficient ansatz (RX, RY, RZ per wire + CNOT ring) at
    14 qubits with distinct batch rows: compiled vs interpreting kernels."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    n, bsz, layers = 14, 3, 4
    ops = config3_ops(seed=2, n=n, layers=layers)
    obs, coef = config3_observables(n)
    phi = torch.linspace(0.1, 1.3, bsz).reshape(bsz, 1).to(DEV)
    out = {}
    for jit in ("on", "off"):
        params = [torch.tensor(op["params"], dtype=torch.float32, device=DEV, requires_grad=True) for op in ops
                  if op["params"] is not None]
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        q.engine_options = tq.EngineOptions(jit=jit)
        for w in range(n):
            tqf.ry(q, w, params=phi)
        k = 0
        for op in ops:
            if op["params"] is None:
                tqf.func_name_dict[op["name"]](q, op["wires"])
            else:
                tqf.func_name_dict[op["name"]](q, op["wires"], params=params[k])
                k += 1
        loss = tq.expval_pauli_sum(q, obs, coef).mean()
        loss.backward()
        out[jit] = (float(loss.detach()), torch.cat([p.grad.reshape(-1) for p in params]).cpu().numpy())
    assert abs(out["on"][0] - out["off"][0]) < 1e-5
    assert np.abs(out["on"][1] - out["off"][1]).max() < 1e-5


@gpu
def test_io_layouts_state_and_gradients():
    """Stages that load in the register layout of their first ops and store
    in that of their last (the default; 16-qubit layered ansatz, 12-bit
    stages, where both kinds occur): compiled and interpreting kernels give
    the oracle's state (1e-5) and the gradients of canonical loads and
    stores (EngineOptions(io_layouts=False)) within 1e-5."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E
    E._plan_cache.clear()
    n, bsz, layers = 16, 2, 3
    ops = config3_ops(seed=3, n=n, layers=layers)
    obs, coef = config3_observables(n)
    phi = np.linspace(0.2, 1.1, bsz).reshape(bsz, 1)
    ref = O.run_ops(oracle_ops([dict(name="ry", wires=[w], params=phi, inverse=False) for w in range(n)] + ops,
                               bsz), n, bsz)
    out = {}
    for label, kw in (("compiled", dict(jit="on")), ("interpreting", dict(jit="off")),
                      ("canonical", dict(jit="off", io_layouts=False))):
        params = [torch.tensor(op["params"], dtype=torch.float32, device=DEV, requires_grad=True) for op in ops
                  if op["params"] is not None]
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        q.engine_options = tq.EngineOptions(stage_bits=12, **kw)
        for w in range(n):
            tqf.ry(q, w, params=torch.tensor(phi, dtype=torch.float32, device=DEV))
        k = 0
        for op in ops:
            if op["params"] is None:
                tqf.func_name_dict[op["name"]](q, op["wires"])
            else:
                tqf.func_name_dict[op["name"]](q, op["wires"], params=params[k])
                k += 1
        state = q.get_states_1d().detach().cpu().numpy()
        loss = tq.expval_pauli_sum(q, obs, coef).mean()
        loss.backward()
        out[label] = (state, torch.cat([p.grad.reshape(-1) for p in params]).cpu().numpy())
    plans = [c for c in E._plan_cache.values() if c.opts.io_layouts]
    flags = [int(st.words[E.H_FLAGS]) for c in plans for kind, (b, st) in c.steps if kind == "stage"]
    assert any((f >> E.FL_LOAD_SHIFT) & 0x1FFF for f in flags) and any(f >> E.FL_STORE_SHIFT for f in flags)
    assert _plans_with_jit("f") and _plans_with_jit("a"), "compiled stage kernels did not run"
    for label in out:
        assert np.abs(out[label][0] - ref).max() < 1e-5, label
    for label in ("compiled", "interpreting"):
        assert np.abs(out[label][1] - out["canonical"][1]).max() < 1e-5, label


@gpu
def test_jit_auto_compiles_in_background():
    """jit="auto" on a 2^22-amplitude state: the first pass runs the
    interpreting kernels while the stage kernels compile on a background
    thread; after wait_for_compiled_stages() the same plan runs compiled
    kernels, with the same amplitudes (1e-6) and gradients (1e-5)."""
    import torchquantum_amd as tq
    from torchquantum_amd import engine as E
    E._plan_cache.clear()
    rng = np.random.default_rng(9)
    n, bsz = 18, 16
    names = ["rx", "ry", "rz", "u3", "cnot", "cz", "crx", "rzz", "h", "swap"]
    ops = random_ops(n, 40, bsz, rng, names=names)
    obs, coef = config3_observables(n)
    outs = []
    for _ in range(2):
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
        ts = apply_ops_tq(q, ops, DEV, requires_grad=True)
        loss = tq.expval_pauli_sum(q, obs, coef).sum()
        st = q.get_states_1d().detach().cpu().numpy()
        loss.backward()
        g = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in ts])
        outs.append((st, g))
        if not outs[1:]:
            plans = [c for c in E._plan_cache.values() if c.bsz == bsz and c.n == n]
            assert plans and all("f" not in c._jit for c in plans)   # first pass: interpreting kernels
            assert tq.wait_for_compiled_stages(timeout=600)
    plans = [c for c in E._plan_cache.values() if c.bsz == bsz and c.n == n]
    assert any(c._jit.get("f") for c in plans) and any(c._jit.get("a") for c in plans)
    assert np.abs(outs[0][0] - outs[1][0]).max() < 1e-6
    assert np.abs(outs[0][1] - outs[1][1]).max() < 1e-5
