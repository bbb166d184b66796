"""GPU parity: the HIP engine (through the C ABI) against the complex128 oracle.

Bars (written in each test): amplitudes and expectations within 1e-5
max-abs (complex64 engine vs complex128 oracle, circuits of <= 160 gates);
gradients against central finite differences of the float64 oracle.
"""
import numpy as np
import pytest
import torch

from circuits import ALL_GATES, apply_ops_tq, oracle_ops, random_ops, config1_ops, config3_observables
from oracle import statevector as O

gpu = pytest.mark.gpu

DEV = "cuda"


def _qdev(n, bsz, mode, **kw):
    import torchquantum_amd as tq
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device=DEV)
    q.engine_options = tq.EngineOptions(mode=mode, **kw)
    return q


def _run(n, bsz, ops, mode, init=None, **kw):
    q = _qdev(n, bsz, mode, **kw)
    if init is not None:
        q.set_states(torch.tensor(init, dtype=torch.complex64, device=DEV))
    apply_ops_tq(q, ops, DEV)
    return q.get_states_1d().detach().cpu().numpy()


@gpu
@pytest.mark.parametrize("mode", ["pergate", "fused"])
@pytest.mark.parametrize("name", ALL_GATES)
def test_each_gate(name, mode):
    """Every gate, random wires, params, inverse flag, per-row params; 1e-5.
    Seeded by the gate's position in ALL_GATES and the mode, so a failure
    reproduces from its test id."""
    rng = np.random.default_rng(1000 + 2 * ALL_GATES.index(name) + (mode == "fused"))
    n, bsz = 6, 3
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    ops = random_ops(n, 4, bsz, rng, names=[name])
    got = _run(n, bsz, ops, mode, init)
    ref = O.run_ops(oracle_ops(ops, bsz), n, bsz, state=init)
    assert np.abs(got - ref).max() < 1e-5, (name, mode)


@gpu
@pytest.mark.parametrize("name", ["cz", "cu1", "crz", "ccz", "cs"])
def test_low_bit_controlled_diagonals(name):
    """Controlled diagonals whose controls include amplitude bit 1 or 0 (the
    whole-line diagonal path) and bit pairs above it, shared and per-row
    parameters, per-gate kernels, from a random state; 1e-5."""
    from torchquantum_amd import gates as G
    rng = np.random.default_rng(7)
    n, bsz = 6, 3
    k = G.spec(name).num_wires
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    ops = []
    for top in range(n - k + 1):
        for rows in (1, bsz):
            wires = list(range(top, top + k))[::-1] if rows == 1 else list(range(top, top + k))
            npar = G.spec(name).num_params
            p = rng.uniform(-np.pi, np.pi, size=(rows, npar)) if npar else None
            ops.append(dict(name=name, wires=wires, params=p, inverse=bool(rows > 1)))
    got = _run(n, bsz, ops, "pergate", init)
    ref = O.run_ops(oracle_ops(ops, bsz), n, bsz, state=init)
    assert np.abs(got - ref).max() < 1e-5, name


@gpu
@pytest.mark.parametrize("mode,stage_bits", [("pergate", 0), ("fused", 0), ("fused", 4), ("fused", 6),
                                             ("fused", 9), ("fused", 13)])
@pytest.mark.parametrize("n", [4, 7, 11, 14])
def test_random_circuits(n, mode, stage_bits):
    """Random mixed circuits (all gate kinds), several stage widths; 1e-5."""
    rng = np.random.default_rng(1000 * n + stage_bits)
    bsz = 2
    ops = random_ops(n, 60, bsz, rng)
    got = _run(n, bsz, ops, mode, stage_bits=stage_bits)
    ref = O.run_ops(oracle_ops(ops, bsz), n, bsz)
    assert np.abs(got - ref).max() < 1e-5


@gpu
@pytest.mark.parametrize("mode", ["pergate", "fused"])
def test_config1_parity(mode):
    """10 qubits, bsz 32, 8 x (U3 all + CNOT ring), <Z_i>: amplitudes and expvals 1e-5."""
    import torchquantum_amd as tq
    n, bsz = 10, 32
    ops = config1_ops(seed=0)
    q = _qdev(n, bsz, mode)
    apply_ops_tq(q, ops, DEV)
    ev = tq.MeasureAll(tq.PauliZ)(q).cpu().numpy()
    ref = O.run_ops(oracle_ops(ops, 1), n, 1)
    assert np.abs(q.get_states_1d().cpu().numpy() - ref).max() < 1e-5
    assert np.abs(ev - O.expval_z_all(ref, n)).max() < 1e-5


@gpu
def test_expvals():
    """MeasureAll(X/Y/Z/H/I), expval_joint_analytical (incl. Y strings) vs oracle, 1e-5."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(7)
    n, bsz = 9, 3
    ops = random_ops(n, 40, bsz, rng)
    for obs_cls in (tq.PauliZ, tq.PauliX, tq.PauliY, tq.Hadamard, tq.I):
        q = _qdev(n, bsz, "fused")
        apply_ops_tq(q, ops, DEV)
        ev = tq.MeasureAll(obs_cls)(q).cpu().numpy()
        st = O.run_ops(oracle_ops(ops, bsz), n, bsz)
        ref = np.zeros((bsz, n))
        for w in range(n):
            ch = {"PauliZ": "Z", "PauliX": "X", "PauliY": "Y", "I": "I"}.get(obs_cls.__name__)
            if ch is None:   # Hadamard observable: (X + Z)/sqrt(2)
                sx = ["I"] * n
                sx[w] = "X"
                sz = ["I"] * n
                sz[w] = "Z"
                ref[:, w] = (O.expval_pauli(st, "".join(sx)) + O.expval_pauli(st, "".join(sz))) / np.sqrt(2)
            else:
                s = ["I"] * n
                s[w] = ch
                ref[:, w] = O.expval_pauli(st, "".join(s))
        assert np.abs(ev - ref).max() < 1e-5, obs_cls.__name__
    q = _qdev(n, bsz, "fused")
    apply_ops_tq(q, ops, DEV)
    st = O.run_ops(oracle_ops(ops, bsz), n, bsz)
    for _ in range(12):
        obs = "".join(rng.choice(list("IXYZ"), n))
        got = tq.expval_joint_analytical(q, obs).cpu().numpy()
        assert np.abs(got - O.expval_pauli(st, obs)).max() < 1e-5, obs


@gpu
@pytest.mark.parametrize("mode", ["pergate", "fused"])
def test_adjoint_gradients_fd(mode):
    """d(sum ZZ + 0.5 sum X)/dtheta through random trainable gates vs central FD
    of the float64 oracle (h = 1e-4); tolerance 2e-4 abs + 1e-3 rel."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(11)
    n, bsz = 8, 2
    names = ["rx", "ry", "rz", "u3", "rot", "cnot", "cz", "crx", "cry", "crz", "rzz", "rxx", "ryy", "rzx", "cu3",
             "u1", "phaseshift", "cu1", "h", "swap", "toffoli", "multirz", "xxplusyy", "singleexcitation", "cu",
             "u2", "r", "crot"]
    ops = random_ops(n, 50, bsz, rng, names=names)
    obs, coef = config3_observables(n)
    q = _qdev(n, bsz, mode)
    ts = apply_ops_tq(q, ops, DEV, requires_grad=True)
    loss = tq.expval_pauli_sum(q, obs, coef).sum()
    loss.backward()

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
    got = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in ts])
    assert np.all(np.abs(got - fd) <= 2e-4 + 1e-3 * np.abs(fd)), np.abs(got - fd).max()


@gpu
def test_state_gradient():
    """Gradient w.r.t. an input state set with set_states (encoder path) vs FD."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(3)
    n, bsz = 6, 2
    ops = random_ops(n, 20, bsz, rng, names=["rx", "cnot", "u3", "cz", "swap"])
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    x = torch.tensor(init, dtype=torch.complex64, device=DEV, requires_grad=True)
    q = _qdev(n, bsz, "fused")
    q.set_states(x)
    apply_ops_tq(q, ops, DEV)
    obs = "ZXIYZI"
    tq.expval_joint_analytical(q, obs).sum().backward()
    g = x.grad.cpu().numpy()

    def f(v):
        st = O.run_ops(oracle_ops(ops, bsz), n, bsz, state=v)
        return float(O.expval_pauli(st, obs).sum())

    h = 1e-4
    for (b, j) in [(0, 3), (1, 17), (0, 40)]:
        for comp in (1.0, 1j):
            vp, vm = init.copy(), init.copy()
            vp[b, j] += h * comp
            vm[b, j] -= h * comp
            d = (f(vp) - f(vm)) / (2 * h)
            gc = g[b, j].real if comp == 1.0 else g[b, j].imag
            assert abs(gc - d) < 1e-3 * (1 + abs(d))


@gpu
def test_fused_matches_pergate_large():
    """n = 18, bsz 2, 200 random gates: fused stages vs per-gate kernels, 1e-5."""
    rng = np.random.default_rng(21)
    n, bsz = 18, 2
    names = ["h", "rx", "ry", "rz", "u3", "cnot", "cz", "swap", "crx", "rzz", "toffoli", "cswap", "rxx", "s"]
    ops = random_ops(n, 200, bsz, rng, names=names)
    a = _run(n, bsz, ops, "fused")
    b = _run(n, bsz, ops, "pergate")
    assert np.abs(a - b).max() < 1e-5


@gpu
def test_config5_family_parity():
    """The 30-qubit workload's circuit family (RY + CZ brickwork, Toffoli/CSWAP
    every 5 layers) at 14 qubits, bsz 2: amplitudes and chain <ZZ> vs the oracle
    (1e-5) for fused and per-gate plans; 12 parameter gradients vs FD."""
    import torchquantum_amd as tq
    from circuits import config5_ops, chain_zz
    n, bsz = 14, 2
    ops = config5_ops(seed=3, n=n)
    obs = chain_zz(n)
    st = O.run_ops(oracle_ops(ops, 1), n, 1)
    ref_zz = np.array([O.expval_pauli(st, o)[0] for o in obs])
    for mode in ("fused", "pergate"):
        q = _qdev(n, bsz, mode)
        ts = apply_ops_tq(q, ops, DEV, requires_grad=True)
        assert np.abs(q.get_states_1d().detach().cpu().numpy() - st).max() < 1e-5
        loss = tq.expval_pauli_sum(q, obs, [1.0] * len(obs))
        assert np.abs(loss.detach().cpu().numpy() - ref_zz.sum()).max() < 1e-4
        loss.sum().backward()
        got = np.concatenate([t.grad.detach().cpu().numpy().reshape(-1) for t in ts]) / bsz
        flat = np.concatenate([op["params"].reshape(-1) for op in ops if op["params"] is not None])

        def f(v):
            k = 0
            o2 = []
            for op in ops:
                p = op["params"]
                if p is not None:
                    p = v[k:k + p.size].reshape(p.shape)
                    k += p.size
                o2.append(dict(op, params=p))
            s = O.run_ops(oracle_ops(o2, 1), n, 1)
            return float(sum(O.expval_pauli(s, o)[0] for o in obs))

        idx = np.random.default_rng(0).choice(flat.size, 12, replace=False)
        fd = O.central_fd(f, flat.copy(), h=1e-4, idx=idx)[idx]
        assert np.all(np.abs(got[idx] - fd) <= 2e-4 + 1e-3 * np.abs(fd)), (mode, np.abs(got[idx] - fd).max())


@gpu
def test_parameters_updated_between_flushes():
    """A GPU parameter tensor changed in place (optimizer-style), through a view
    or by .data assignment between runs is read afresh every time (the host
    copies of parameter values are reused only while a tensor is unchanged)."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    n = 3
    base = torch.tensor([0.3, 1.1], device=DEV)
    w = torch.tensor([[0.7]], device=DEV)

    def run():
        q = tq.QuantumDevice(n_wires=n, bsz=1, device=DEV)
        tqf.ry(q, wires=0, params=base[0:1])          # 1-D view: normalised to [1, 1]
        tqf.rx(q, wires=1, params=base[1:2])
        tqf.ry(q, wires=2, params=w)
        tqf.cnot(q, wires=[0, 2])
        return q.get_states_1d().cpu().numpy()

    def ref():
        b = base.cpu().numpy().astype(np.float64)
        ops = [dict(name="ry", wires=[0], params=b[0:1], inverse=False),
               dict(name="rx", wires=[1], params=b[1:2], inverse=False),
               dict(name="ry", wires=[2], params=w.cpu().numpy().reshape(-1).astype(np.float64), inverse=False),
               dict(name="cnot", wires=[0, 2], params=None, inverse=False)]
        return O.run_ops(ops, n, 1)

    for step in range(4):
        assert np.abs(run() - ref()).max() < 1e-5, step
        with torch.no_grad():
            if step == 0:
                base.add_(0.25)
            elif step == 1:
                w.mul_(-2.0)
            else:
                w.data = torch.tensor([[2.9]], device=DEV)


def _dense_apply_np(state, U, bits, cmask, cval, n):
    """numpy complex128: U on bit positions ``bits`` (bits[0] = matrix MSB)
    where (index & cmask) == cval, for every row of ``state``."""
    k = len(bits)
    tm = sum(1 << b for b in bits)
    j = np.arange(1 << n)
    base = j[((j & tm) == 0) & ((j & cmask) == cval)]
    offs = np.array([sum(1 << bits[q] for q in range(k) if (r >> (k - 1 - q)) & 1) for r in range(1 << k)])
    idx = base[:, None] + offs[None, :]
    out = state.copy()
    for b in range(state.shape[0]):
        out[b][idx] = state[b][idx] @ U.T
    return out


@gpu
@pytest.mark.parametrize("k", [5, 6])
@pytest.mark.parametrize("nctrl", [0, 1, 2])
def test_dense_kq_matrix_cores(k, nctrl):
    """k = 5, 6 dense gates with one matrix for all rows and >= 64 groups per
    row take the matrix-core kernel: unsorted target bits (low bits
    included), 0-2 controls, through the C ABI; vs numpy complex128, 1e-5."""
    from torchquantum_amd import _native as N
    rng = np.random.default_rng(100 * k + nctrl)
    n, bsz = 14, 2
    for trial in range(3):
        bits = [int(b) for b in rng.permutation(n)[:k]]
        if trial == 0:
            bits = [int(b) for b in rng.permutation(k)]          # the k lowest bits, shuffled
        rest = [b for b in range(n) if b not in bits]
        cb = [int(b) for b in rng.choice(rest, nctrl, replace=False)]
        cmask = sum(1 << b for b in cb)
        cval = sum(1 << b for b in cb if rng.random() < 0.5)
        z = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
        U, _ = np.linalg.qr(z)
        init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
        init /= np.linalg.norm(init, axis=1, keepdims=True)
        st = torch.tensor(init, dtype=torch.complex64, device=DEV)
        m = torch.tensor(U, dtype=torch.complex64, device=DEV)
        N.call("tqsv_apply_kq", N.stream_of(st), st.data_ptr(), bsz, n, k, N.int_array(bits), cmask, cval,
               m.data_ptr(), 0)
        ref = _dense_apply_np(init, U, bits, cmask, cval, n)
        assert np.abs(st.cpu().numpy() - ref).max() < 1e-5, (k, nctrl, trial, bits)


@gpu
@pytest.mark.parametrize("width", [5, 6])
def test_wide_fusion_blocks(width):
    """Fusion width 5 / 6: the 5- and 6-wire blocks run on the matrix-core
    kernel between stages (shared parameters) or the VALU kernel (per-row
    parameters); states vs the oracle, 1e-5, and a gradient through the
    trainable (<= 2-wire) blocks vs FD."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(40 + width)
    n, bsz = 12, 2
    names = ["h", "rx", "ry", "rz", "u3", "cnot", "cz", "swap", "crx", "rzz", "toffoli", "s", "t", "sx", "cu3"]
    for batched_frac in (0.0, 0.5):
        ops = random_ops(n, 160, bsz, rng, names=names, batched_frac=batched_frac)
        got = _run(n, bsz, ops, "fused", fusion_width=width)
        ref = O.run_ops(oracle_ops(ops, bsz), n, bsz)
        assert np.abs(got - ref).max() < 1e-5, (width, batched_frac)
        from torchquantum_amd import engine as E
        q = _qdev(n, bsz, "fused", fusion_width=width)
        apply_ops_tq(q, ops, DEV)
        c = E.Compiled(q._tape, n, bsz, q.engine_options, "cpu")       # the plan, built on the host only
        wide = [k for k in c.kops if k.kind == "matk" and len(k.bits) > 4]
        assert wide and (batched_frac > 0 or any(not k.batched for k in wide))


@gpu
@pytest.mark.parametrize("mode", ["pergate", "fused"])
def test_tiny_and_wide_batches(mode):
    """Edge shapes: no gates at all, 1 and 2 wires, bsz 1, and a wide batch
    (bsz 1024) with per-row parameters; vs the oracle, 1e-5."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    q = _qdev(3, 2, mode)
    st = q.get_states_1d().cpu().numpy()
    assert np.abs(st - O.zero_state(3, 2)).max() == 0.0
    rng = np.random.default_rng(5)
    for n, bsz in ((1, 1), (1, 3), (2, 1), (2, 4)):
        names = ["h", "rx", "ry", "rz", "u3", "s", "t", "sx"] + (["cnot", "cz", "swap", "crx", "rzz"] if n > 1 else [])
        ops = random_ops(n, 12, bsz, rng, names=names)
        got = _run(n, bsz, ops, mode)
        assert np.abs(got - O.run_ops(oracle_ops(ops, bsz), n, bsz)).max() < 1e-5, (n, bsz)
    n, bsz = 8, 1024
    ops = random_ops(n, 30, bsz, rng, names=["rx", "ry", "u3", "cnot", "crx", "rzz"], batched_frac=0.7)
    got = _run(n, bsz, ops, mode)
    assert np.abs(got - O.run_ops(oracle_ops(ops, bsz), n, bsz)).max() < 1e-5


def _apply_pauli_np(state, obs):
    """P psi for a Pauli string (char k <-> wire k), P|j> = i^nY (-1)^popc(j&z) |j^x>."""
    n = len(obs)
    x = z = ny = 0
    for w, ch in enumerate(obs):
        bit = 1 << (n - 1 - w)
        x |= bit if ch in "XY" else 0
        z |= bit if ch in "ZY" else 0
        ny += ch == "Y"
    idx = np.arange(2 ** n)
    par = np.array([bin(v).count("1") & 1 for v in (idx & z)])
    out = np.empty_like(state)
    out[:, idx ^ x] = state * ((1j ** ny) * (1 - 2 * par))
    return out


@gpu
@pytest.mark.parametrize("n", [8, 14])
def test_pauli_sums_grouped(n):
    """Weighted sums of many Pauli strings (X/Y on every wire, ZZ pairs, mixed
    strings, strings flipping more high wires than one local set holds):
    the value per row vs the oracle (1e-5) and its gradient w.r.t. the state,
    2 sum_t c_t P_t psi, vs the oracle (1e-5)."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(40 + n)
    bsz = 3
    obs = []
    for w in range(n):
        for ch in "XY":
            s = ["I"] * n
            s[w] = ch
            obs.append("".join(s))
    for w in range(n - 1):
        s = ["I"] * n
        s[w] = s[w + 1] = "Z"
        obs.append("".join(s))
    for _ in range(10):
        obs.append("".join(rng.choice(list("IXYZ"), n)))
    obs.append("X" * n)
    obs.append("Y" * (n // 2) + "Z" * (n - n // 2))
    coef = rng.normal(size=len(obs))
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    x = torch.tensor(init, dtype=torch.complex64, device=DEV, requires_grad=True)
    q = _qdev(n, bsz, "fused")
    q.set_states(x)
    val = tq.expval_pauli_sum(q, obs, coef)
    val.sum().backward()
    ref = sum(c * O.expval_pauli(init, o) for o, c in zip(obs, coef))
    assert np.abs(val.detach().cpu().numpy() - ref).max() < 1e-5
    gref = 2 * sum(c * _apply_pauli_np(init, o) for o, c in zip(obs, coef))
    assert np.abs(x.grad.cpu().numpy() - gref).max() < 1e-5 * max(1.0, np.abs(gref).max())


@gpu
@pytest.mark.parametrize("n", [5, 13])
def test_pauli_sum_value_and_state_gradient(n):
    """A differentiated Pauli-sum loss (tqsv_pauli_sum_grad: value and
    2 sum_t c_t P_t psi from the same passes; 13 qubits: every term in local
    sets, the value from the last set pass; 5 qubits: per-term passes and a
    separate dot) with per-row loss weights, vs the oracle: values (1e-5) and
    d loss / d conj(psi) = 2 w_row sum_t c_t P_t psi (1e-5); the value
    without a gradient (read-only expectation kernels) agrees too."""
    import torchquantum_amd as tq
    rng = np.random.default_rng(70 + n)
    bsz = 4
    obs = []
    for w in range(n):
        s = ["I"] * n
        s[w] = "XYX"[w % 3]
        obs.append("".join(s))
    for w in range(n - 1):
        s = ["I"] * n
        s[w] = s[w + 1] = "Z"
        obs.append("".join(s))
    s = ["I"] * n
    s[0], s[n - 1] = "Y", "X"
    obs.append("".join(s))
    coef = rng.normal(size=len(obs))
    wrow = rng.normal(size=bsz)
    init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
    init /= np.linalg.norm(init, axis=1, keepdims=True)
    x = torch.tensor(init, dtype=torch.complex64, device=DEV, requires_grad=True)
    q = _qdev(n, bsz, "fused")
    q.set_states(x)
    val = tq.expval_pauli_sum(q, obs, coef)
    (val * torch.tensor(wrow, dtype=torch.float32, device=DEV)).sum().backward()
    ref = sum(c * O.expval_pauli(init, o) for o, c in zip(obs, coef))
    assert np.abs(val.detach().cpu().numpy() - ref).max() < 1e-5
    gref = 2 * wrow[:, None] * sum(c * _apply_pauli_np(init, o) for o, c in zip(obs, coef))
    assert np.abs(x.grad.cpu().numpy() - gref).max() < 1e-5 * max(1.0, np.abs(gref).max())
    with torch.no_grad():
        q2 = _qdev(n, bsz, "fused")
        q2.set_states(x.detach())
        v2 = tq.expval_pauli_sum(q2, obs, coef)
    assert np.abs(v2.cpu().numpy() - ref).max() < 1e-5


@gpu
@pytest.mark.parametrize("k", [3, 4, 5, 6])
@pytest.mark.parametrize("batched", [False, True])
def test_dense_kq_lds_staged(k, batched):
    """Dense k-qubit gates (3 <= k <= 6) on the LDS-staged kernel: one matrix
    or one per batch row, unsorted target bits, 0-2 controls, and a shape
    whose rows hold fewer groups than a workgroup pass (the global-read
    kernel); through the C ABI, vs numpy complex128, 1e-5."""
    from torchquantum_amd import _native as N
    rng = np.random.default_rng(300 + 10 * k + batched)
    for n, bsz, nctrl in ((11, 3, 0), (12, 2, 2), (k + 2, 3, 1)):
        bits = [int(b) for b in rng.permutation(n)[:k]]
        rest = [b for b in range(n) if b not in bits]
        cb = [int(b) for b in rng.choice(rest, min(nctrl, len(rest)), replace=False)]
        cmask = sum(1 << b for b in cb)
        cval = sum(1 << b for b in cb if rng.random() < 0.5)
        mats = []
        for _ in range(bsz if batched else 1):
            z = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
            mats.append(np.linalg.qr(z)[0])
        init = rng.normal(size=(bsz, 2 ** n)) + 1j * rng.normal(size=(bsz, 2 ** n))
        init /= np.linalg.norm(init, axis=1, keepdims=True)
        st = torch.tensor(init, dtype=torch.complex64, device=DEV)
        m = torch.tensor(np.stack(mats), dtype=torch.complex64, device=DEV).contiguous()
        N.call("tqsv_apply_kq", N.stream_of(st), st.data_ptr(), bsz, n, k, N.int_array(bits), cmask, cval,
               m.data_ptr(), (4 ** k) if batched else 0)
        ref = np.stack([_dense_apply_np(init[b:b + 1], mats[b if batched else 0], bits, cmask, cval, n)[0]
                        for b in range(bsz)])
        assert np.abs(st.cpu().numpy() - ref).max() < 1e-5, (k, batched, n, bsz, nctrl)
