"""GPU: adjoint gradients of the hardware-efficient-ansatz workload checked
against references that share no code with the engine.

* every gradient entry of the ansatz (10 x [RX, RY, RZ on every wire, CNOT
  ring], per-row RY input layer, loss = mean_b [sum_i <Z_i Z_i+1> + 0.5 sum_i
  <X_i>]) at 10 qubits, bsz 4, against central finite differences of the
  float64 oracle, for the per-gate, compiled-stage and interpreting-stage paths;
* the same ansatz at 12 qubits, bsz 4, against torch autograd through the
  restated reference per-gate path (oracle/refpath.py: the reference's own
  permute + bmm chain, functional/gate_wrapper.py:96-140, and its matrix
  formulas), in complex64 as the reference runs it and in complex128;
* at the full workload size, 24 qubits, bsz 64, against a parameter-shift
  gradient computed from forward passes only (exact for RX/RY/RZ:
  util/utils.py:1017-1075) on 8 seeded parameters.
Tolerances are written in each test.
"""
import numpy as np
import pytest
import torch

from circuits import config3_ops, config3_observables
from oracle import statevector as O

gpu = pytest.mark.gpu
DEV = "cuda"


def _ansatz_inputs(n, bsz, layers=10, seed=0):
    ops = config3_ops(seed=seed, n=n, layers=layers)
    obs, coef = config3_observables(n)
    phi = np.random.default_rng(1).random(bsz) * np.pi
    theta0 = np.array([float(np.asarray(op["params"]).reshape(-1)[0]) for op in ops if op["params"] is not None])
    return ops, obs, coef, phi, theta0


def _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, theta, opts, want_grad=True):
    """loss = mean_b sum_t c_t <P_t> through the engine; returns (loss, grad [P])."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    params = [torch.tensor([[t]], dtype=torch.float32, device=DEV, requires_grad=want_grad) for t in theta]
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    q.engine_options = opts
    ph = torch.tensor(phi, dtype=torch.float32, device=DEV).reshape(bsz, 1)
    for w in range(n):
        tqf.ry(q, w, params=ph)
    k = 0
    for op in ops:
        if op["params"] is None:
            tqf.func_name_dict[op["name"]](q, op["wires"])
        else:
            tqf.func_name_dict[op["name"]](q, op["wires"], params=params[k])
            k += 1
    loss = tq.expval_pauli_sum(q, obs, coef).mean()
    if not want_grad:
        return float(loss), None
    loss.backward()
    return float(loss.detach()), torch.cat([p.grad.reshape(-1) for p in params]).cpu().numpy().astype(np.float64)


def _oracle_loss(n, bsz, ops, obs, coef, phi):
    def f(theta):
        seq = [dict(name="ry", wires=[w], params=phi.reshape(bsz, 1), inverse=False) for w in range(n)]
        k = 0
        for op in ops:
            if op["params"] is None:
                seq.append(op)
            else:
                seq.append(dict(op, params=np.array([[theta[k]]])))
                k += 1
        st = O.run_ops(seq, n, bsz)
        return float(np.mean(sum(c * O.expval_pauli(st, o) for o, c in zip(obs, coef))))
    return f


@gpu
def test_config3_ansatz_every_gradient_vs_fd():
    """10 qubits, bsz 4, all 300 parameters: engine gradient vs central FD of
    the float64 oracle (h = 1e-4), |g - fd| <= 1e-5 + 1e-4 |fd|; loss within 1e-5."""
    import torchquantum_amd as tq
    n, bsz = 10, 4
    ops, obs, coef, phi, theta0 = _ansatz_inputs(n, bsz)
    f = _oracle_loss(n, bsz, ops, obs, coef, phi)
    loss_ref = f(theta0)
    fd = O.central_fd(f, theta0.copy(), h=1e-4)
    for label, opts in (("pergate", tq.EngineOptions(mode="pergate")),
                        ("fused-compiled", tq.EngineOptions(mode="fused", jit="on")),
                        ("fused-interpreting", tq.EngineOptions(mode="fused", jit="off")),
                        ("fused-no-mg", tq.EngineOptions(mode="fused", jit="off", adjoint_mg=False))):
        loss, g = _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, theta0, opts)
        assert abs(loss - loss_ref) < 1e-5, (label, loss, loss_ref)
        err = np.abs(g - fd)
        assert np.all(err <= 1e-5 + 1e-4 * np.abs(fd)), (label, float(err.max()), int(err.argmax()))


def _refpath_grad(n, bsz, ops, obs, coef, phi, theta, cdtype):
    from oracle import refpath as R
    tparams = [torch.tensor([[t]], dtype=torch.float64, requires_grad=True) for t in theta]
    seq, plist = [], []
    for w in range(n):
        seq.append(dict(name="ry", wires=[w]))
        plist.append(torch.tensor(phi.reshape(bsz, 1)))
    k = 0
    for op in ops:
        seq.append(op)
        if op["params"] is None:
            plist.append(None)
        else:
            plist.append(tparams[k])
            k += 1
    st = R.run_autograd(seq, plist, n, bsz, cdtype)
    loss = sum(c * R.expval_pauli(st, o) for o, c in zip(obs, coef)).mean()
    loss.backward()
    return float(loss.detach()), torch.cat([p.grad.reshape(-1) for p in tparams]).numpy()


@gpu
def test_adjoint_vs_autograd_through_reference_path():
    """12 qubits, bsz 4, 360 parameters: the adjoint gradient against torch
    autograd through the restated reference per-gate path.  complex64 (the
    reference's own dtype): |g - g_ref| <= 2e-5 + 1e-3 |g_ref|; complex128:
    <= 1e-5 + 1e-4 |g_ref|."""
    import torchquantum_amd as tq
    n, bsz = 12, 4
    ops, obs, coef, phi, theta0 = _ansatz_inputs(n, bsz)
    l64, g64 = _refpath_grad(n, bsz, ops, obs, coef, phi, theta0, torch.complex64)
    l128, g128 = _refpath_grad(n, bsz, ops, obs, coef, phi, theta0, torch.complex128)
    assert abs(l64 - l128) < 1e-4 and np.abs(g64 - g128).max() < 1e-4    # the two reference runs agree
    for label, opts in (("pergate", tq.EngineOptions(mode="pergate")),
                        ("fused-compiled", tq.EngineOptions(mode="fused", jit="on"))):
        loss, g = _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, theta0, opts)
        assert abs(loss - l128) < 1e-5, (label, loss, l128)
        assert np.all(np.abs(g - g64) <= 2e-5 + 1e-3 * np.abs(g64)), (label, float(np.abs(g - g64).max()))
        assert np.all(np.abs(g - g128) <= 1e-5 + 1e-4 * np.abs(g128)), (label, float(np.abs(g - g128).max()))


@gpu
def test_adjoint_vs_parameter_shift_full_size():
    """24 qubits, bsz 64 (the workload's size): fused adjoint gradient vs the
    parameter-shift rule (f(theta + pi/2) - f(theta - pi/2)) / 2 on 8 seeded
    parameters, every forward on the GPU; |g - g_ps| <= 2e-5 + 1e-4 |g_ps|.
    Runs the program-interpreting stage kernels (no run-time compile)."""
    import torchquantum_amd as tq
    n, bsz = 24, 64
    ops, obs, coef, phi, theta0 = _ansatz_inputs(n, bsz)
    opts = tq.EngineOptions(mode="fused", jit="off")
    _, g = _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, theta0, opts)
    torch.cuda.empty_cache()
    idx = np.random.default_rng(7).choice(len(theta0), 8, replace=False)
    for i in idx:
        tp, tm = theta0.copy(), theta0.copy()
        tp[i] += np.pi / 2
        tm[i] -= np.pi / 2
        with torch.no_grad():
            fp, _ = _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, tp, opts, want_grad=False)
            fm, _ = _engine_loss_and_grad(n, bsz, ops, obs, coef, phi, tm, opts, want_grad=False)
        gps = (fp - fm) / 2
        assert abs(g[i] - gps) <= 2e-5 + 1e-4 * abs(gps), (int(i), float(g[i]), gps)


@gpu
def test_pruned_sweep_matches_full_sweep():
    """Compiled stages of a 12-qubit bsz-3 circuit with an untrained input
    layer and fixed gates ahead of a trainable random part: the reverse sweep
    without the input-state gradient (stopped at the first trainable step,
    un-applies of ops on untouched wires left out, nothing stored by its last
    stage) gives the gradients of the full sweep, |dg| <= 1e-5 (1 + |g|)."""
    from circuits import random_ops
    from torchquantum_amd import engine as E, gates as G
    rng = np.random.default_rng(11)
    n, bsz = 12, 3
    ops = [dict(name="ry", wires=[w], params=rng.uniform(0, np.pi, size=(bsz, 1)), inverse=False) for w in range(n)]
    ops += [dict(name="hadamard", wires=[3], params=None, inverse=False),
            dict(name="cnot", wires=[7, 9], params=None, inverse=False)]
    train = random_ops(n, 40, bsz, rng, names=["rx", "ry", "rz", "u3", "cnot", "cz", "crx", "rzz", "rot", "swap",
                                                 "toffoli", "phaseshift"])
    tape = []
    for i, op in enumerate(ops + train):
        p = op["params"]
        if p is not None:
            p = torch.tensor(p, dtype=torch.float32, device=DEV, requires_grad=i >= len(ops))
        tape.append(E.TapeOp(G.canonical(op["name"]), tuple(op["wires"]), p, op["inverse"]))
    c = E.Compiled(tape, n, bsz, E.EngineOptions(mode="fused", stage_bits=10, jit="on"), DEV)
    stop, skip = c.prune_info()
    assert skip and c.jit_kernels("a") and c.jit_kernels("p")
    pool = c.pool(tape, True, True)
    out = E.forward(c, None, pool, init_zero=True)
    g = torch.Generator(device=DEV).manual_seed(5)
    lam = torch.randn(bsz, 1 << n, dtype=torch.complex64, device=DEV, generator=g)
    full, lam_out = E.adjoint(c, out, lam, pool, need_input_grad=True)
    pruned, none = E.adjoint(c, out, lam, pool, need_input_grad=False)
    assert none is None and lam_out is not None
    full, pruned = full.double().cpu().numpy(), pruned.double().cpu().numpy()
    assert np.all(np.abs(pruned - full) <= 1e-5 * (1 + np.abs(full))), float(np.abs(pruned - full).max())


@gpu
@pytest.mark.parametrize("label", ["pergate", "fused-interpreting", "fused-compiled"])
def test_gradients_bitwise_reproducible(label):
    """The same gradient evaluated twice is bitwise equal: every adjoint
    kernel writes per-workgroup partials that a second launch adds up in a
    fixed order (no float atomics).  12 qubits, bsz 4, the gradient ansatz."""
    import torchquantum_amd as tq
    opts = {"pergate": tq.EngineOptions(mode="pergate"),
            "fused-interpreting": tq.EngineOptions(mode="fused", jit="off"),
            "fused-compiled": tq.EngineOptions(mode="fused", jit="on")}[label]
    n, bsz = 12, 4
    ops, obs, coef, phi, theta0 = _ansatz_inputs(n, bsz)
    runs = [_engine_loss_and_grad(n, bsz, ops, obs, coef, phi, theta0, opts) for _ in range(3)]
    for loss, g in runs[1:]:
        assert loss == runs[0][0]
        assert np.array_equal(g, runs[0][1]), float(np.abs(g - runs[0][1]).max())


@gpu
def test_adjoint_rejects_short_partial_buffer():
    """Adjoint entry points check the partial buffer against the launch
    before launching: a short buffer is an error, not an out-of-bounds
    write."""
    from torchquantum_amd import _native as N, engine as E, gates as G
    lib = N.lib()
    n, bsz = 12, 2
    psi = torch.zeros(bsz, 1 << n, dtype=torch.complex64, device=DEV)
    lam = torch.zeros_like(psi)
    grad = torch.zeros(bsz, 8, dtype=torch.float32, device=DEV)
    mat = torch.zeros(8, dtype=torch.complex64, device=DEV)
    work = torch.zeros(4, dtype=torch.float32, device=DEV)
    need = lib.tqsv_adjoint_work_floats(bsz, n)
    assert need > 4
    s = N.stream_of(psi)
    rc = lib.tqsv_adjoint_1q(s, psi.data_ptr(), lam.data_ptr(), bsz, n, 3, 0, 0, mat.data_ptr(), 0, 1,
                             N.fptr(grad), 8, 0, work.data_ptr(), 4)
    assert rc == -1 and b"partial buffer" in lib.tqsv_last_error()
    tape = [E.TapeOp("rx", (w,), torch.full((1, 1), 0.3, device=DEV, requires_grad=True), False) for w in range(n)]
    c = E.Compiled(tape, n, bsz, E.EngineOptions(mode="fused", stage_bits=10, jit="off"), DEV)
    base, st = next(p for k, p in c.steps if k == "stage")
    hp = c.prog_host[base:]
    need_st = lib.tqsv_stage_adjoint_work_floats(hp.ctypes.data_as(N.c_int_p), n, bsz, 0)
    assert need_st > 4
    rc = lib.tqsv_stage_adjoint(s, psi.data_ptr(), lam.data_ptr(), bsz, n, hp.ctypes.data_as(N.c_int_p),
                                c.prog_dev.data_ptr() + 4 * base, mat.data_ptr(), N.fptr(grad), 8,
                                work.data_ptr(), 4, 0)
    assert rc == -1 and b"partial buffer" in lib.tqsv_last_error()
    torch.cuda.synchronize()
