"""Density-matrix device (NoiseDevice): rho as a 2n-wire register on the
statevector engine, gates as U (x) conj(U), channels as superoperators.

CPU part: the channel helpers and the recorded register tape; the whole
path run through the CPU emulator of the engine's kernels against the
density oracle (oracle/density.py), itself pinned here against textbook
closed forms and the statevector oracle.  GPU part: the same on the HIP
kernels, expectations Tr(rho P), gradients, sampling.
Bars: 1e-6 (emulator, complex128), 1e-5 (GPU, complex64).
"""
import numpy as np
import pytest
import torch

import emulator as EM
from circuits import apply_ops_tq, oracle_ops, random_ops
from oracle import density as OD
from oracle import statevector as O

gpu = pytest.mark.gpu
DEV = "cuda"

GATES = ["hadamard", "rx", "ry", "rz", "u3", "cnot", "cz", "crx", "cry", "rzz", "rxx", "swap", "toffoli", "s",
         "t", "sx", "u1", "cu3", "iswap", "ecr"]


# ---------------------------------------------------------------- the oracle
def test_density_oracle_pure_states():
    """Unitary evolution of |0><0| is |psi><psi| of the statevector oracle."""
    rng = np.random.default_rng(0)
    n, bsz = 4, 2
    ops = oracle_ops(random_ops(n, 25, bsz, rng, names=GATES), bsz)
    rho = OD.run_density_ops(ops, n, bsz)
    psi = O.run_ops(ops, n, bsz)
    assert np.abs(rho - np.einsum("bi,bj->bij", psi, np.conj(psi))).max() < 1e-12


def test_density_oracle_channels_closed_forms():
    from torchquantum_amd import noisedevice as ND
    g, lam, p = 0.3, 0.4, 0.25
    one = np.array([[[0, 0], [0, 1]]], dtype=complex)
    assert np.allclose(OD.apply_kraus(one, ND.amplitude_damping_kraus(g), [0], 1)[0], np.diag([g, 1 - g]))
    plus = np.full((1, 2, 2), 0.5, dtype=complex)
    r = OD.apply_kraus(plus, ND.phase_damping_kraus(lam), [0], 1)[0]
    assert np.allclose(r, [[0.5, 0.5 * np.sqrt(1 - lam)], [0.5 * np.sqrt(1 - lam), 0.5]])
    zero = np.array([[[1, 0], [0, 0]]], dtype=complex)
    assert np.allclose(OD.apply_kraus(zero, ND.depolarizing_kraus(p), [0], 1)[0], np.diag([1 - 2 * p / 3, 2 * p / 3]))
    assert np.allclose(OD.apply_kraus(zero, ND.bit_flip_kraus(p), [0], 1)[0], np.diag([1 - p, p]))
    assert np.allclose(OD.apply_kraus(plus, ND.phase_flip_kraus(p), [0], 1)[0],
                       [[0.5, 0.5 * (1 - 2 * p)], [0.5 * (1 - 2 * p), 0.5]])


def test_superoperator_is_the_kraus_map():
    """sum_k K_k (x) conj(K_k) applied to vec(rho) (row-major) equals the
    Kraus map; every helper channel preserves the trace."""
    from torchquantum_amd import noisedevice as ND
    rng = np.random.default_rng(1)
    a = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    rho = a @ np.conj(a.T)
    rho /= np.trace(rho)
    ks = [rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)) for _ in range(3)]
    got = (ND.superoperator(ks) @ rho.reshape(-1)).reshape(4, 4)
    assert np.allclose(got, sum(k @ rho @ np.conj(k.T) for k in ks))
    for ch in (ND.amplitude_damping_kraus(0.3), ND.phase_damping_kraus(0.2), ND.depolarizing_kraus(0.4),
               ND.bit_flip_kraus(0.1), ND.phase_flip_kraus(0.7)):
        assert np.allclose(sum(np.conj(k.T) @ k for k in ch), np.eye(2))


# ------------------------------------------------------------- the recording
def test_density_device_records_u_and_conj_u():
    import torchquantum_amd as tq
    d = tq.NoiseDevice(3, bsz=2, device="cpu")
    d.rx(0, params=torch.tensor([[0.3]]))
    d.cnot([0, 2])
    d.depolarizing(1, 0.1)
    t = d.reg._tape
    assert [(o.name, o.wires, o.conj, o.nonunitary) for o in t] == [
        ("rx", (0,), False, False), ("rx", (3,), True, False), ("cnot", (0, 2), False, False),
        ("cnot", (3, 5), True, False), ("qubitunitaryfast", (1, 4), False, True)]
    with pytest.raises(ValueError):
        d.rx(3, params=torch.tensor([[0.3]]))          # wire range is the logical n


def _emulated(n, bsz, ops, channels, mode):
    """Run a NoiseDevice tape through the CPU emulator of the engine."""
    import torchquantum_amd as tq
    from torchquantum_amd import engine as E
    d = tq.NoiseDevice(n, bsz=bsz, device="cpu")
    for i, op in enumerate(ops):
        apply_ops_tq(d, [op], "cpu")
        if i in channels:
            name, w, val = channels[i]
            getattr(d, name)(w, val)
    tape = d.reg._tape
    c = E.Compiled(tape, 2 * n, bsz, E.EngineOptions(mode=mode), "cpu")
    pool = c.pool_host(tape, False, True).astype(np.complex128)
    v = EM.run_compiled(c, pool, None, init_zero=True)
    return v.reshape(bsz, 2 ** n, 2 ** n)


@pytest.mark.parametrize("mode", ["pergate", "fused"])
def test_density_emulated_vs_oracle(mode):
    """Random gates (per-row parameters, inverses) and channels, emulated."""
    from torchquantum_amd import noisedevice as ND
    rng = np.random.default_rng(2)
    n, bsz = 3, 2
    ops = random_ops(n, 20, bsz, rng, names=GATES)
    channels = {3: ("amplitude_damping", 0, 0.3), 8: ("depolarizing", 2, 0.2), 12: ("phase_damping", 1, 0.4),
                16: ("bit_flip", 0, 0.1)}
    got = _emulated(n, bsz, ops, channels, mode)
    kr = {"amplitude_damping": ND.amplitude_damping_kraus, "depolarizing": ND.depolarizing_kraus,
          "phase_damping": ND.phase_damping_kraus, "bit_flip": ND.bit_flip_kraus}
    seq = []
    for i, o in enumerate(oracle_ops(ops, bsz)):
        seq.append(o)
        if i in channels:
            name, w, val = channels[i]
            seq.append({"kraus": kr[name](val), "wires": [w]})
    ref = OD.run_density_ops(seq, n, bsz)
    assert np.abs(got - ref).max() < 1e-6
    assert np.allclose(np.trace(got, axis1=1, axis2=2), 1.0)


# -------------------------------------------------------------------- GPU
def _gpu_density(n, bsz, ops, channels):
    import torchquantum_amd as tq
    d = tq.NoiseDevice(n, bsz=bsz, device=DEV)
    for i, op in enumerate(ops):
        apply_ops_tq(d, [op], DEV)
        if i in channels:
            name, w, val = channels[i]
            getattr(d, name)(w, val)
    return d


@gpu
def test_density_gpu_vs_oracle():
    """Densities, probabilities, trace, purity and Tr(rho P) on the GPU."""
    import torchquantum_amd as tq
    from torchquantum_amd import noisedevice as ND
    rng = np.random.default_rng(3)
    n, bsz = 4, 3
    ops = random_ops(n, 30, bsz, rng, names=GATES)
    channels = {5: ("amplitude_damping", 1, 0.3), 15: ("depolarizing", 3, 0.2), 22: ("phase_flip", 0, 0.3)}
    d = _gpu_density(n, bsz, ops, channels)
    kr = {"amplitude_damping": ND.amplitude_damping_kraus, "depolarizing": ND.depolarizing_kraus,
          "phase_flip": ND.phase_flip_kraus}
    seq = []
    for i, o in enumerate(oracle_ops(ops, bsz)):
        seq.append(o)
        if i in channels:
            name, w, val = channels[i]
            seq.append({"kraus": kr[name](val), "wires": [w]})
    ref = OD.run_density_ops(seq, n, bsz)
    got = d.densities.reshape(bsz, 2 ** n, 2 ** n).cpu().numpy()
    assert np.abs(got - ref).max() < 1e-5
    assert np.abs(d.get_probs_1d().cpu().numpy() - np.diagonal(ref, axis1=1, axis2=2)).max() < 1e-5
    assert np.abs(d.trace().cpu().numpy() - 1).max() < 1e-5
    assert np.abs(d.purity().cpu().numpy() - np.real(np.einsum("bij,bji->b", ref, ref))).max() < 1e-5
    for _ in range(8):
        obs = "".join(rng.choice(list("IXYZ"), n))
        got_e = tq.expval_joint_analytical(d, obs).cpu().numpy()
        assert np.abs(got_e - OD.expval_pauli(ref, obs)).max() < 1e-5, obs
    zs = tq.MeasureAll(tq.PauliZ)(d).cpu().numpy()
    for w in range(n):
        s = ["I"] * n
        s[w] = "Z"
        assert np.abs(zs[:, w] - OD.expval_pauli(ref, "".join(s))).max() < 1e-5


@gpu
def test_density_gradients_match_statevector():
    """A unitary circuit on the density device gives the statevector's
    expectations and parameter gradients (1e-5)."""
    import torchquantum_amd as tq
    from circuits import config3_observables
    rng = np.random.default_rng(4)
    n, bsz = 4, 2
    names = ["rx", "ry", "rz", "u3", "cnot", "crx", "rzz", "cz", "h"]
    ops = random_ops(n, 24, bsz, rng, names=names)
    obs, coef = config3_observables(n)
    res = []
    for dev_cls in (tq.QuantumDevice, tq.NoiseDevice):
        d = dev_cls(n_wires=n, bsz=bsz, device=DEV)
        ts = apply_ops_tq(d, ops, DEV, requires_grad=True)
        loss = tq.expval_pauli_sum(d, obs, coef).sum()
        loss.backward()
        res.append((float(loss.detach()), np.concatenate([t.grad.cpu().numpy().reshape(-1) for t in ts])))
    assert abs(res[0][0] - res[1][0]) < 1e-5
    assert np.abs(res[0][1] - res[1][1]).max() < 1e-5


@gpu
def test_density_channel_gradient_raises_and_sampling():
    import torchquantum_amd as tq
    d = tq.NoiseDevice(3, bsz=2, device=DEV)
    th = torch.tensor([[0.4]], device=DEV, requires_grad=True)
    d.ry(0, params=th)
    d.amplitude_damping(0, 0.5)
    with pytest.raises(NotImplementedError):
        tq.expval_joint_analytical(d, "ZII").sum().backward()
    # sampling follows rho_ii: here p(wire 0 = 1) = sin^2(0.2) * 0.5
    d2 = tq.NoiseDevice(3, bsz=2, device=DEV)
    d2.ry(0, params=torch.tensor([[0.4]], device=DEV))
    d2.amplitude_damping(0, 0.5)
    idx = tq.sample_indices(d2, n_shots=40000, seed=5).cpu().numpy()
    p1 = np.sin(0.2) ** 2 * 0.5
    frac = (idx >= 4).mean(axis=1)
    assert np.all(np.abs(frac - p1) < 5 * np.sqrt(p1 * (1 - p1) / 40000))
    assert np.all(idx % 4 == 0)      # wires 1, 2 stay |0>


@gpu
def test_density_compiled_stages():
    """A 6-wire density (12-wire register, one compiled stage width) with
    channels: compiled and interpreting stage kernels agree (1e-6) and match
    the oracle (1e-5)."""
    import torchquantum_amd as tq
    from torchquantum_amd import engine as E, noisedevice as ND
    E._plan_cache.clear()
    rng = np.random.default_rng(6)
    n, bsz = 6, 2
    ops = random_ops(n, 30, bsz, rng, names=GATES)
    channels = {4: ("depolarizing", 2, 0.1), 17: ("amplitude_damping", 5, 0.25)}
    out = {}
    for jit in ("on", "off"):
        d = tq.NoiseDevice(n, bsz=bsz, device=DEV)
        d.engine_options = tq.EngineOptions(jit=jit)
        for i, op in enumerate(ops):
            apply_ops_tq(d, [op], DEV)
            if i in channels:
                name, w, val = channels[i]
                getattr(d, name)(w, val)
        out[jit] = d.densities.reshape(bsz, 2 ** n, 2 ** n).cpu().numpy()
    assert any(c._jit.get("f") for c in E._plan_cache.values() if c.opts.jit == "on")
    kr = {"depolarizing": ND.depolarizing_kraus, "amplitude_damping": ND.amplitude_damping_kraus}
    seq = []
    for i, o in enumerate(oracle_ops(ops, bsz)):
        seq.append(o)
        if i in channels:
            name, w, val = channels[i]
            seq.append({"kraus": kr[name](val), "wires": [w]})
    ref = OD.run_density_ops(seq, n, bsz)
    assert np.abs(out["on"] - out["off"]).max() < 1e-6
    assert np.abs(out["on"] - ref).max() < 1e-5
