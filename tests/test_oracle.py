"""Pin the complex128 oracle before trusting it (CPU only).

Each function of oracle/statevector.py is checked against something that is
not a retyping of its own formula:
  * golden vectors and known-answer values from the reference's tests
    (tests/golden/reference_vectors.json);
  * textbook identities: rotation gates are matrix exponentials of Pauli
    generators (scipy.linalg.expm), decompositions (Rot = RZ RY RZ, SWAP =
    3 CNOTs, ...), QFT = normalised inverse DFT;
  * brute force: gate application against an explicitly built 2^n x 2^n
    operator; Pauli expectations against the reference's own dense
    Kronecker-Hamiltonian algorithm;
  * finite differences against closed-form derivatives.
A planted mistake (transposed matrix, wrong wire order) must fail the checks.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import statevector as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2)


def _ctrl(u, n_ctrl=1):
    """|0><0| (x) I + |1><1| (x) U, nested (independent of the oracle's helper)."""
    out = u
    for _ in range(n_ctrl):
        d = out.shape[0]
        out = np.kron(np.diag([1, 0]), np.eye(d)) + np.kron(np.diag([0, 1]), out)
    return out


def _cx(golden_re, golden_im=None):
    re = np.array(golden_re, dtype=float)
    im = np.zeros_like(re) if golden_im is None else np.array(golden_im, dtype=float)
    return re + 1j * im


# ------------------------------------------------------------------ golden
def test_golden_init():
    for case in GOLD["init_states"]["cases"]:
        s = O.zero_state(case["n_wires"], case["bsz"])
        np.testing.assert_array_equal(s, _cx(case["expect_re"], case["expect_im"]))


def test_golden_controlled():
    tol = GOLD["tolerance"]
    for case in GOLD["controlled_unitary"]["cases"]:
        s = O.run_ops(case["ops"], case["n_wires"], 1)
        assert np.allclose(s, _cx(case["expect_re"], case["expect_im"]), rtol=tol["rtol"], atol=1e-7)


def test_golden_unsorted_wires():
    g = GOLD["qubitunitary_unsorted_wires"]
    n = g["n_wires"]
    s = O.run_ops([dict(name="paulix", wires=[w]) for w in g["x_on"]], n, 1)
    assert np.allclose(s, _cx(g["expect_after_x_re"]))
    u = np.array(g["matrix_re"], dtype=complex)
    s = O.apply(s, u, g["wires"], n)
    assert np.allclose(s, _cx(g["expect_final_re"]))


def test_golden_expval_kat():
    g = GOLD["expval_joint_kat"]
    s = O.run_ops(g["ops"], g["n_wires"], 1)
    for obs, val in g["expect"].items():
        assert abs(O.expval_pauli(s, obs)[0] - val) < 1e-12
        assert abs(O.expval_pauli_dense(s, obs)[0] - val) < 1e-12


# ------------------------------------------------------------------ matrices
def test_unitarity():
    rng = np.random.default_rng(0)
    for name in list(O.FIXED) + list(O.PARAM):
        m = O.gate_matrix(name, rng.uniform(-3, 3, 4))
        if name == "cu2":   # the reference's cu2 block is not normalised (functional/u2.py)
            assert not np.allclose(m @ m.conj().T, np.eye(m.shape[0]))
            continue
        assert np.allclose(m @ m.conj().T, np.eye(m.shape[0])), name


@pytest.mark.parametrize("t", [0.0, 0.37, -1.9, 3.1])
def test_rotations_are_exponentials(t):
    kr = np.kron
    assert np.allclose(O.gate_matrix("rx", [t]), expm(-0.5j * t * X))
    assert np.allclose(O.gate_matrix("ry", [t]), expm(-0.5j * t * Y))
    assert np.allclose(O.gate_matrix("rz", [t]), expm(-0.5j * t * Z))
    assert np.allclose(O.gate_matrix("rxx", [t]), expm(-0.5j * t * kr(X, X)))
    assert np.allclose(O.gate_matrix("ryy", [t]), expm(-0.5j * t * kr(Y, Y)))
    assert np.allclose(O.gate_matrix("rzz", [t]), expm(-0.5j * t * kr(Z, Z)))
    assert np.allclose(O.gate_matrix("rzx", [t]), expm(-0.5j * t * kr(Z, X)))
    for k in (1, 2, 3, 4):
        zk = Z
        for _ in range(k - 1):
            zk = kr(zk, Z)
        assert np.allclose(O.gate_matrix("multirz", [t], k), expm(-0.5j * t * zk))
    assert np.allclose(O.gate_matrix("phaseshift", [t]), expm(1j * t * (I2 - Z) / 2))
    assert np.allclose(O.gate_matrix("crx", [t]), _ctrl(expm(-0.5j * t * X)))
    assert np.allclose(O.gate_matrix("cry", [t]), _ctrl(expm(-0.5j * t * Y)))
    assert np.allclose(O.gate_matrix("crz", [t]), _ctrl(expm(-0.5j * t * Z)))
    # single excitation: Givens rotation in the {|01>, |10>} subspace, U = exp(-i t (YX - XY) / 4)
    assert np.allclose(O.gate_matrix("singleexcitation", [t]), expm(-0.25j * t * (kr(Y, X) - kr(X, Y))))
    # XX+YY with beta = 0: exp(-i t/4 (XX + YY))
    assert np.allclose(O.gate_matrix("xxplusyy", [t, 0.0]), expm(-0.25j * t * (kr(X, X) + kr(Y, Y))))
    assert np.allclose(O.gate_matrix("xxminyy", [t, 0.0]), expm(-0.25j * t * (kr(X, X) - kr(Y, Y))))


def test_decompositions():
    rng = np.random.default_rng(1)
    a, b, c, d = rng.uniform(-3, 3, 4)
    rz, ry = (lambda t: O.gate_matrix("rz", [t])), (lambda t: O.gate_matrix("ry", [t]))
    assert np.allclose(O.gate_matrix("rot", [a, b, c]), rz(c) @ ry(b) @ rz(a))
    assert np.allclose(O.gate_matrix("u3", [b, -math.pi / 2, math.pi / 2]), O.gate_matrix("rx", [b]))
    assert np.allclose(O.gate_matrix("u3", [b, 0, 0]), ry(b))
    # U3 = e^{i(phi+lam)/2} RZ(phi) RY(theta) RZ(lam)
    assert np.allclose(O.gate_matrix("u3", [b, c, d]), np.exp(0.5j * (c + d)) * rz(c) @ ry(b) @ rz(d))
    assert np.allclose(O.gate_matrix("u2", [c, d]), O.gate_matrix("u3", [math.pi / 2, c, d]))
    assert np.allclose(O.gate_matrix("u1", [c]), O.gate_matrix("u3", [0, 0, c]))
    assert np.allclose(O.gate_matrix("r", [b, 0.0]), O.gate_matrix("rx", [b]))
    assert np.allclose(O.gate_matrix("r", [b, math.pi / 2]), ry(b))
    H = O.gate_matrix("hadamard")
    assert np.allclose(H, (X + Z) / math.sqrt(2))
    S, T = O.gate_matrix("s"), O.gate_matrix("t")
    assert np.allclose(S @ S, Z) and np.allclose(T @ T, S)
    assert np.allclose(O.gate_matrix("sdg"), S.conj().T) and np.allclose(O.gate_matrix("tdg"), T.conj().T)
    SX = O.gate_matrix("sx")
    assert np.allclose(SX @ SX, X) and np.allclose(O.gate_matrix("sxdg"), SX.conj().T)
    assert np.allclose(O.gate_matrix("pauliy"), 1j * X @ Z)
    cnot = _ctrl(X)
    assert np.allclose(O.gate_matrix("cnot"), cnot)
    assert np.allclose(O.gate_matrix("cz"), _ctrl(Z))
    assert np.allclose(O.gate_matrix("cy"), _ctrl(Y))
    assert np.allclose(O.gate_matrix("chadamard"), _ctrl(H))
    assert np.allclose(O.gate_matrix("csx"), _ctrl(SX))
    rev = np.kron(I2, np.diag([1, 0])) + np.kron(X, np.diag([0, 1]))  # CNOT controlled by wire 1
    assert np.allclose(O.gate_matrix("swap"), cnot @ rev @ cnot)
    assert np.allclose(O.gate_matrix("dcx"), rev @ cnot)
    assert np.allclose(O.gate_matrix("iswap"), expm(0.25j * math.pi * (np.kron(X, X) + np.kron(Y, Y))))
    assert np.allclose(O.gate_matrix("sswap") @ O.gate_matrix("sswap"), O.gate_matrix("swap"))
    assert np.allclose(O.gate_matrix("toffoli"), _ctrl(X, 2))
    assert np.allclose(O.gate_matrix("ccz"), _ctrl(Z, 2))
    assert np.allclose(O.gate_matrix("c3x"), _ctrl(X, 3)) and np.allclose(O.gate_matrix("c4x"), _ctrl(X, 4))
    assert np.allclose(O.gate_matrix("c3sx"), _ctrl(SX, 3))
    assert np.allclose(O.gate_matrix("cswap"), _ctrl(O.gate_matrix("swap")))
    assert np.allclose(O.gate_matrix("cu3", [b, c, d]), _ctrl(O.gate_matrix("u3", [b, c, d])))
    assert np.allclose(O.gate_matrix("cu", [b, c, d, a]), _ctrl(np.exp(1j * a) * O.gate_matrix("u3", [b, c, d])))
    assert np.allclose(O.gate_matrix("crot", [a, b, c]), _ctrl(O.gate_matrix("rot", [a, b, c])))
    assert np.allclose(O.gate_matrix("cu1", [a]), _ctrl(O.gate_matrix("u1", [a])))
    assert np.allclose(O.gate_matrix("multicnot", None, 3), _ctrl(X, 2))
    # multixcnot flips the last wire when all controls are 0
    mx = O.gate_matrix("multixcnot", None, 3)
    xm = np.kron(np.kron(X, X), I2) @ _ctrl(X, 2) @ np.kron(np.kron(X, X), I2)
    assert np.allclose(mx, xm)
    # ECR = (IX - XY)/sqrt(2) in Qiskit's little-endian order = (XI - YX)/sqrt(2) with wire 0 first
    assert np.allclose(O.gate_matrix("ecr"), (np.kron(X, I2) - np.kron(Y, X)) / math.sqrt(2))
    # RCCX (Margolus): a relative-phase Toffoli -- equals Toffoli up to a diagonal
    r = O.gate_matrix("rccx")
    dg = r @ _ctrl(X, 2).conj().T
    assert np.allclose(dg, np.diag(np.diag(dg))) and np.allclose(np.abs(np.diag(dg)), 1)
    r3 = O.gate_matrix("rc3x")
    dg = r3 @ _ctrl(X, 3).conj().T
    assert np.allclose(dg, np.diag(np.diag(dg))) and np.allclose(np.abs(np.diag(dg)), 1)


def _circuit_matrix(seq, n):
    """Product of a gate sequence [(name, wires, params)], first gate applied first,
    each embedded by the brute-force Kronecker builder."""
    m = np.eye(2 ** n, dtype=complex)
    for name, wires, p in seq:
        m = O.full_unitary(O.gate_matrix(name, p), wires, n) @ m
    return m


def test_relative_phase_toffolis_exact():
    """RCCX and RC3X with their exact phases, from the published Clifford+T
    circuits (Margolus' 3-qubit gate; the 4-qubit relative-phase Toffoli of
    Maslov, Phys. Rev. A 93, 022311 (2016), as written in Qiskit's RCCXGate /
    RC3XGate definitions), target = last wire."""
    pi = math.pi
    t, td = ("u1", [pi / 4]), ("u1", [-pi / 4])
    seq3 = [("hadamard", [2], None), (t[0], [2], t[1]), ("cnot", [1, 2], None), (td[0], [2], td[1]),
            ("cnot", [0, 2], None), (t[0], [2], t[1]), ("cnot", [1, 2], None), (td[0], [2], td[1]),
            ("hadamard", [2], None)]
    assert np.allclose(_circuit_matrix(seq3, 3), O.gate_matrix("rccx"), atol=1e-12)
    seq4 = [("hadamard", [3], None), (t[0], [3], t[1]), ("cnot", [2, 3], None), (td[0], [3], td[1]),
            ("hadamard", [3], None), ("cnot", [0, 3], None), (t[0], [3], t[1]), ("cnot", [1, 3], None),
            (td[0], [3], td[1]), ("cnot", [0, 3], None), (t[0], [3], t[1]), ("cnot", [1, 3], None),
            (td[0], [3], td[1]), ("hadamard", [3], None), (t[0], [3], t[1]), ("cnot", [2, 3], None),
            (td[0], [3], td[1]), ("hadamard", [3], None)]
    assert np.allclose(_circuit_matrix(seq4, 4), O.gate_matrix("rc3x"), atol=1e-12)
    # a phase flipped on one entry must fail
    bad = O.gate_matrix("rccx").copy()
    bad[6, 7] *= -1
    assert not np.allclose(_circuit_matrix(seq3, 3), bad)


def test_shadamard_is_ry_quarter_pi():
    """SHadamard = RY(pi/4) = exp(-i pi/8 Y) (hadamard.py:16-22)."""
    sh = O.gate_matrix("shadamard")
    assert np.allclose(sh, expm(-0.5j * (math.pi / 4) * Y))
    assert np.allclose(sh, O.gate_matrix("ry", [math.pi / 4]))
    # and it is the basis change that maps Z to H: sh^dagger Z sh = H (up to sign convention)
    assert np.allclose(sh @ Z @ sh.conj().T, O.gate_matrix("hadamard"))


@pytest.mark.parametrize("t,beta", [(0.37, 0.9), (-1.9, -2.4), (3.1, 1.3), (0.8, math.pi / 2)])
def test_xx_pm_yy_with_phase(t, beta):
    """XX+YY(theta, beta) = exp(-i theta/4 D (XX+YY) D^dagger) with D = diag(1, e^{-i beta})
    on wire 0, and XX-YY(theta, beta) = exp(-i theta/4 D' (XX-YY) D'^dagger) with
    D' = diag(1, e^{i beta}) on wire 0: the beta-rotated generators (xx_plus_yy.py,
    xx_min_yy.py).  A sign flip of beta must fail."""
    kr = np.kron
    gp, gm = kr(X, X) + kr(Y, Y), kr(X, X) - kr(Y, Y)
    d = kr(np.diag([1, np.exp(-1j * beta)]), I2)
    dm = kr(np.diag([1, np.exp(1j * beta)]), I2)
    assert np.allclose(O.gate_matrix("xxplusyy", [t, beta]), expm(-0.25j * t * (d @ gp @ d.conj().T)))
    assert np.allclose(O.gate_matrix("xxminyy", [t, beta]), expm(-0.25j * t * (dm @ gm @ dm.conj().T)))
    assert not np.allclose(O.gate_matrix("xxplusyy", [t, -beta]), expm(-0.25j * t * (d @ gp @ d.conj().T)))
    assert not np.allclose(O.gate_matrix("xxminyy", [t, -beta]), expm(-0.25j * t * (dm @ gm @ dm.conj().T)))


def test_qft_is_inverse_dft():
    for k in (1, 2, 3, 4):
        d = 2 ** k
        v = np.random.default_rng(k).normal(size=d) + 0j
        assert np.allclose(O.gate_matrix("qft", None, k) @ v, np.fft.ifft(v) * math.sqrt(d))


# ------------------------------------------------------------------ application
@pytest.mark.parametrize("seed", range(6))
def test_apply_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 6))
    k = int(rng.integers(1, min(n, 3) + 1))
    wires = [int(w) for w in rng.choice(n, k, replace=False)]   # unsorted on purpose
    a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    s = rng.normal(size=(2, 2 ** n)) + 1j * rng.normal(size=(2, 2 ** n))
    full = O.full_unitary(a, wires, n)
    assert np.allclose(O.apply(s, a, wires, n), s @ full.T)
    # per-row matrices
    b = np.stack([a, a.conj()])
    out = O.apply(s, b, wires, n)
    assert np.allclose(out[1], s[1] @ O.full_unitary(a.conj(), wires, n).T)


def test_full_unitary_kron_convention():
    """wire 0 is the most significant bit: U on wire 0 of 2 = kron(U, I)."""
    a = np.random.default_rng(9).normal(size=(2, 2)) + 0j
    assert np.allclose(O.full_unitary(a, [0], 2), np.kron(a, I2))
    assert np.allclose(O.full_unitary(a, [1], 2), np.kron(I2, a))
    b = np.random.default_rng(10).normal(size=(4, 4)) + 0j
    assert np.allclose(O.full_unitary(b, [0, 1], 2), b)
    swap = O.gate_matrix("swap")
    assert np.allclose(O.full_unitary(b, [1, 0], 2), swap @ b @ swap)


def test_planted_mistakes_are_caught():
    """A transposed matrix or reversed wire order must fail the brute-force check."""
    rng = np.random.default_rng(3)
    n, wires = 4, [2, 0]
    a = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    s = rng.normal(size=(1, 16)) + 0j
    ref = s @ O.full_unitary(a, wires, n).T
    assert not np.allclose(O.apply(s, a.T, wires, n), ref)
    assert not np.allclose(O.apply(s, a, wires[::-1], n), ref)


# ------------------------------------------------------------------ expectations
@pytest.mark.parametrize("seed", range(5))
def test_expval_pauli_vs_dense_hamiltonian(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 6))
    s = rng.normal(size=(3, 2 ** n)) + 1j * rng.normal(size=(3, 2 ** n))
    for _ in range(8):
        obs = "".join(rng.choice(list("IXYZ"), n))
        assert np.allclose(O.expval_pauli(s, obs), O.expval_pauli_dense(s, obs))


def test_expval_z_all_vs_strings():
    rng = np.random.default_rng(4)
    n = 5
    s = rng.normal(size=(2, 2 ** n)) + 1j * rng.normal(size=(2, 2 ** n))
    z = O.expval_z_all(s, n)
    for w in range(n):
        obs = "".join("Z" if i == w else "I" for i in range(n))
        assert np.allclose(z[:, w], O.expval_pauli_dense(s, obs))


def test_fd_vs_closed_form():
    """<Z> after RX(t)|0> = cos t, so d/dt = -sin t; <Y> after RX(t)|0> = -sin t."""
    for t in (0.3, 1.2, -2.0):
        f = lambda v: O.expval_z_all(O.run_ops([dict(name="rx", wires=[0], params=[v[0]])], 1, 1), 1)[0, 0]
        assert abs(f(np.array([t])) - math.cos(t)) < 1e-12
        assert abs(O.central_fd(f, np.array([t]), h=1e-4)[0] + math.sin(t)) < 1e-7
        s = O.run_ops([dict(name="rx", wires=[0], params=[t])], 1, 1)
        assert abs(O.expval_pauli(s, "Y")[0] + math.sin(t)) < 1e-12


def test_parameter_shift_identity():
    """For Pauli rotations the central difference with h = pi/2 is exact
    (parameter-shift rule, util/utils.py:1017-1075)."""
    rng = np.random.default_rng(6)
    ops = [dict(name="ry", wires=[0], params=[0.4]), dict(name="cnot", wires=[0, 1]),
           dict(name="rx", wires=[1], params=[1.3]), dict(name="rzz", wires=[0, 1], params=[0.8])]

    def f(v):
        o2 = [dict(op, params=[v[i]]) if op.get("params") else op for i, op in enumerate(ops)]
        s = O.run_ops(o2, 2, 1)
        return O.expval_pauli(s, "ZZ")[0] + 0.3 * O.expval_pauli(s, "XI")[0]

    v = np.array([0.4, 0.0, 1.3, 0.8])
    ps = np.array([(f(v + math.pi / 2 * np.eye(4)[i]) - f(v - math.pi / 2 * np.eye(4)[i])) / 2 for i in range(4)])
    fd = O.central_fd(f, v, h=1e-5)
    assert np.allclose(ps, fd, atol=1e-8)


# ------------------------------------------------------------------ restated reference path
def test_refpath_autograd_matches_oracle():
    """oracle/refpath.py's differentiable restatement of the reference's
    per-gate torch path: states and Pauli expectations equal the oracle's
    (complex128), its autograd gradient equals central FD of the oracle, and
    the complex64 run (the reference's dtype) is within float32 rounding."""
    import torch
    from oracle import refpath as R
    rng = np.random.default_rng(13)
    n, bsz = 4, 2
    names = ["rx", "ry", "rz", "u3", "cnot", "hadamard", "cz"]
    ops, base = [], []
    for _ in range(24):
        nm = names[rng.integers(len(names))]
        k = 2 if nm in ("cnot", "cz") else 1
        w = [int(x) for x in rng.choice(n, k, replace=False)]
        npar = {"rx": 1, "ry": 1, "rz": 1, "u3": 3}.get(nm, 0)
        rows = bsz if (npar and rng.random() < 0.3) else 1
        p = rng.uniform(-3, 3, size=(rows, npar)) if npar else None
        ops.append(dict(name=nm, wires=w, inverse=bool(rng.random() < 0.2)))
        base.append(p)
    obs = ["ZXIY", "XXZI", "IYYZ"]

    def oracle_f(flat):
        k, seq = 0, []
        for op, p in zip(ops, base):
            if p is not None:
                q = flat[k:k + p.size].reshape(p.shape)
                k += p.size
                seq.append(dict(op, params=q if q.shape[0] > 1 else q[0]))
            else:
                seq.append(dict(op, params=None))
        st = O.run_ops(seq, n, bsz)
        return st, float(sum(O.expval_pauli(st, o).sum() for o in obs))

    flat0 = np.concatenate([p.reshape(-1) for p in base if p is not None])
    st_ref, _ = oracle_f(flat0)
    for cdtype, tol in ((torch.complex128, 1e-12), (torch.complex64, 1e-5)):
        tp = [None if p is None else torch.tensor(p, requires_grad=True) for p in base]
        st = R.run_autograd(ops, tp, n, bsz, cdtype)
        assert np.abs(st.reshape(bsz, -1).detach().numpy() - st_ref).max() < tol
        loss = sum(R.expval_pauli(st, o).sum() for o in obs)
        loss.backward()
        g = np.concatenate([t.grad.numpy().reshape(-1) for t in tp if t is not None])
        fd = O.central_fd(lambda v: oracle_f(v)[1], flat0.copy(), h=1e-5)
        assert np.abs(g - fd).max() < (1e-7 if cdtype == torch.complex128 else 1e-4)
