"""complex128 CPU oracle for the batched statevector path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``torchquantum_amd`` imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

It restates, in plain numpy and from the mathematics, what the reference
computes on its hot path:

* state layout ``[bsz] + [2]*n`` with wire 0 the most significant bit of the
  flat index (torchquantum/device/devices.py:64-71);
* gate application ``psi' = (U on wires) psi`` where ``wires[0]`` is the most
  significant bit of the gate-matrix index
  (torchquantum/functional/gate_wrapper.py:96-140);
* the closed-form gate matrices of torchquantum/functional/*.py;
* ``expval`` / ``MeasureAll`` (torchquantum/measurement/measurements.py:282-343)
  and ``expval_joint_analytical`` (measurements.py:232-279);
* central finite differences for gradients.

Everything is complex128 / float64.  How each function is pinned (golden
vectors from the reference's tests, textbook identities, brute force) is in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

C = np.complex128
INV_SQRT2 = 1.0 / math.sqrt(2.0)


# ---------------------------------------------------------------------------
# gate matrices (closed forms)
# ---------------------------------------------------------------------------

def _c(rows) -> np.ndarray:
    return np.array(rows, dtype=C)


def _controlled(u: np.ndarray, n_ctrl: int) -> np.ndarray:
    """Identity except the bottom-right block (all controls = 1) which is ``u``."""
    d = u.shape[0] * (2 ** n_ctrl)
    m = np.eye(d, dtype=C)
    m[d - u.shape[0]:, d - u.shape[0]:] = u
    return m


_X = _c([[0, 1], [1, 0]])
_Y = _c([[0, -1j], [1j, 0]])
_Z = _c([[1, 0], [0, -1]])
_I2 = np.eye(2, dtype=C)
_H = INV_SQRT2 * _c([[1, 1], [1, -1]])
_SX = 0.5 * _c([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])


def _rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[c, -1j * s], [-1j * s, c]])          # functional/rx.py:19-46


def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[c, -s], [s, c]])                     # functional/ry.py:50-67


def _rz(t):
    e = np.exp(-0.5j * t)
    return _c([[e, 0], [0, np.conj(e)]])             # functional/rz.py:150-171


def _u3(t, p, l):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[c, -np.exp(1j * l) * s],
               [np.exp(1j * p) * s, np.exp(1j * (p + l)) * c]])  # functional/u3.py:52-77


def _rot(p, t, w):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[np.exp(-0.5j * (p + w)) * c, -np.exp(0.5j * (p - w)) * s],
               [np.exp(-0.5j * (p - w)) * s, np.exp(0.5j * (p + w)) * c]])  # functional/rot.py


def _phase(p):
    return _c([[1, 0], [0, np.exp(1j * p)]])         # functional/phase_shift.py, u1.py


def _u2(p, l):
    return INV_SQRT2 * _c([[1, -np.exp(1j * l)], [np.exp(1j * p), np.exp(1j * (p + l))]])  # functional/u2.py


def _r(t, p):
    c, s = math.cos(t / 2), math.sin(t / 2)
    e = np.exp(-1j * p)
    return _c([[c, e * (-1j * s)], [np.conj(e) * (-1j * s), c]])  # functional/r.py


def _rxx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    m = np.eye(4, dtype=C) * c
    for a, b in ((0, 3), (1, 2), (2, 1), (3, 0)):
        m[a, b] = -1j * s
    return m                                          # functional/rx.py rxx_matrix


def _ryy(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    m = np.eye(4, dtype=C) * c
    m[0, 3] = 1j * s
    m[1, 2] = -1j * s
    m[2, 1] = -1j * s
    m[3, 0] = 1j * s
    return m                                          # functional/ry.py ryy_matrix


def _rzz(t):
    e = np.exp(-0.5j * t)
    return np.diag([e, np.conj(e), np.conj(e), e]).astype(C)  # functional/rz.py rzz_matrix


def _rzx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[c, -1j * s, 0, 0], [-1j * s, c, 0, 0],
               [0, 0, c, 1j * s], [0, 0, 1j * s, c]])  # functional/rz.py rzx_matrix


def _xxminyy(t, b):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[c, 0, 0, -1j * s * np.exp(-1j * b)], [0, 1, 0, 0], [0, 0, 1, 0],
               [-1j * s * np.exp(1j * b), 0, 0, c]])  # functional/xx_min_yy.py


def _xxplusyy(t, b):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[1, 0, 0, 0], [0, c, -1j * s * np.exp(1j * b), 0],
               [0, -1j * s * np.exp(-1j * b), c, 0], [0, 0, 0, 1]])  # functional/xx_plus_yy.py


def _single_exc(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return _c([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])  # functional/single_excitation.py


def _cu(t, p, l, g):
    c, s = math.cos(t / 2), math.sin(t / 2)
    u = _c([[c * np.exp(1j * g), -s * np.exp(1j * (l + g))],
            [s * np.exp(1j * (p + g)), c * np.exp(1j * (p + l + g))]])
    return _controlled(u, 1)                          # functional/u3.py cu_matrix


def _multirz(t, k):
    # eigenvalue of Z^{(x)k} on basis state j is (-1)^{popcount(j)}  (util/utils.py:82-96)
    sign = np.array([1 - 2 * (bin(j).count("1") & 1) for j in range(2 ** k)], dtype=np.float64)
    return np.diag(np.exp(-0.5j * t * sign)).astype(C)  # functional/rz.py:19-48


def _qft(k):
    d = 2 ** k
    w = np.exp(2j * np.pi / d)
    j = np.arange(d)
    return (w ** np.outer(j, j) / math.sqrt(d)).astype(C)  # functional/qft.py


FIXED = {
    "hadamard": _H,
    "shadamard": _c([[math.cos(math.pi / 8), -math.sin(math.pi / 8)],
                     [math.sin(math.pi / 8), math.cos(math.pi / 8)]]),
    "paulix": _X, "pauliy": _Y, "pauliz": _Z, "i": _I2,
    "s": _c([[1, 0], [0, 1j]]), "sdg": _c([[1, 0], [0, -1j]]),
    "t": _c([[1, 0], [0, np.exp(1j * math.pi / 4)]]),
    "tdg": _c([[1, 0], [0, np.exp(-1j * math.pi / 4)]]),
    "sx": _SX, "sxdg": _SX.conj().T,
    "cnot": _controlled(_X, 1), "cy": _controlled(_Y, 1), "cz": _controlled(_Z, 1),
    "chadamard": _controlled(_H, 1),
    "cs": _controlled(_c([[1, 0], [0, 1j]]), 1),
    "csdg": _controlled(_c([[1, 0], [0, -1j]]), 1),
    "csx": _controlled(_SX, 1),
    "swap": _c([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
    "sswap": _c([[1, 0, 0, 0], [0, (1 + 1j) / 2, (1 - 1j) / 2, 0],
                 [0, (1 - 1j) / 2, (1 + 1j) / 2, 0], [0, 0, 0, 1]]),
    "iswap": _c([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
    "dcx": _c([[1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [0, 1, 0, 0]]),
    "ecr": INV_SQRT2 * _c([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]]),
    "toffoli": _controlled(_X, 2), "ccz": _controlled(_Z, 2),
    "cswap": _controlled(_c([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]), 1),
    "c3x": _controlled(_X, 3), "c4x": _controlled(_X, 4), "c3sx": _controlled(_SX, 3),
}
# rccx (functional/paulix.py:116-127): diag(1,1,1,1,1,-1) then a [[0,-i],[i,0]] tail
_rccx = np.eye(8, dtype=C)
_rccx[5, 5] = -1
_rccx[6:, 6:] = _c([[0, -1j], [1j, 0]])
FIXED["rccx"] = _rccx
# rc3x (functional/paulix.py:129-150): identity on 0..11, then a 4x4 tail block
_rc3x = np.eye(16, dtype=C)
_rc3x[12:, 12:] = _c([[1j, 0, 0, 0], [0, -1j, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]])
FIXED["rc3x"] = _rc3x

PARAM = {
    # name: (n_params, builder)
    "rx": (1, _rx), "ry": (1, _ry), "rz": (1, _rz),
    "phaseshift": (1, _phase), "u1": (1, _phase),
    "u2": (2, _u2), "u3": (3, _u3), "rot": (3, _rot), "r": (2, _r),
    "rxx": (1, _rxx), "ryy": (1, _ryy), "rzz": (1, _rzz), "rzx": (1, _rzx),
    "crx": (1, lambda t: _controlled(_rx(t), 1)),
    "cry": (1, lambda t: _controlled(_ry(t), 1)),
    "crz": (1, lambda t: _controlled(_rz(t), 1)),
    "crot": (3, lambda p, t, w: _controlled(_rot(p, t, w), 1)),
    "cu1": (1, lambda p: _controlled(_phase(p), 1)),
    "cu3": (3, lambda t, p, l: _controlled(_u3(t, p, l), 1)),
    "cu": (4, _cu),
    "singleexcitation": (1, _single_exc),
    "xxminyy": (2, _xxminyy), "xxplusyy": (2, _xxplusyy),
}


def _cu2(p, l):
    # functional/u2.py cu2_matrix: NOT sqrt(2)-normalised in the reference
    m = np.eye(4, dtype=C)
    m[2, 3] = -np.exp(1j * l)
    m[3, 2] = np.exp(1j * p)
    m[3, 3] = np.exp(1j * (p + l))
    return m


PARAM["cu2"] = (2, _cu2)

ALIASES = {
    "h": "hadamard", "sh": "shadamard", "x": "paulix", "y": "pauliy", "z": "pauliz",
    "cx": "cnot", "ccnot": "toffoli", "ccx": "toffoli", "ch": "chadamard",
    "xx": "rxx", "yy": "ryy", "zz": "rzz", "zx": "rzx", "p": "phaseshift",
    "u": "u3", "cp": "cu1", "cr": "cu1", "cphase": "cu1",
    "echoedcrossresonance": "ecr",
}


def gate_matrix(name: str, params: Sequence[float] | None = None, n_wires: int | None = None) -> np.ndarray:
    """Matrix of gate ``name`` (complex128), for one parameter row."""
    name = ALIASES.get(name, name)
    if name in FIXED:
        return FIXED[name].copy()
    if name in PARAM:
        k, fn = PARAM[name]
        p = [float(v) for v in np.asarray(params, dtype=np.float64).reshape(-1)[:k]]
        return fn(*p)
    if name == "multirz":
        return _multirz(float(np.asarray(params).reshape(-1)[0]), n_wires)
    if name == "multicnot":                          # functional/paulix.py:19-35
        return _controlled(_X, n_wires - 1)
    if name == "multixcnot":                         # functional/paulix.py:38-55
        m = np.eye(2 ** n_wires, dtype=C)
        m[[0, 1], [0, 1]] = 0
        m[0, 1] = m[1, 0] = 1
        return m
    if name == "qft":
        return _qft(n_wires)
    if name == "globalphase":
        return np.array([[np.exp(1j * float(np.asarray(params).reshape(-1)[0]))]], dtype=C)
    if name in ("qubitunitary", "qubitunitaryfast"):
        return np.asarray(params, dtype=C).reshape(2 ** n_wires, 2 ** n_wires)
    raise KeyError(name)


# ---------------------------------------------------------------------------
# state and gate application
# ---------------------------------------------------------------------------

def zero_state(n: int, bsz: int = 1) -> np.ndarray:
    s = np.zeros((bsz, 2 ** n), dtype=C)
    s[:, 0] = 1.0
    return s


def apply(state: np.ndarray, mat: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    """Return (U on ``wires``) ``state``; state is ``[bsz, 2**n]``.

    ``mat`` is ``[2^k, 2^k]`` (shared) or ``[bsz, 2^k, 2^k]`` (one per row);
    ``wires[0]`` is the most significant bit of the matrix index, wire 0 the
    most significant bit of the state index (gate_wrapper.py:96-140).
    """
    wires = list(wires)
    k = len(wires)
    bsz = state.shape[0]
    t = state.reshape([bsz] + [2] * n)
    # move the gate axes to the front (after batch): [b, w0, w1, ..., rest]
    axes = [w + 1 for w in wires]
    rest = [a for a in range(1, n + 1) if a not in axes]
    t = np.transpose(t, [0] + axes + rest).reshape(bsz, 2 ** k, -1)
    if mat.ndim == 2:
        out = np.einsum("ij,bjr->bir", mat, t)
    else:
        out = np.einsum("bij,bjr->bir", mat, t)
    out = out.reshape([bsz] + [2] * n)
    inv = np.argsort([0] + axes + rest)
    return np.ascontiguousarray(np.transpose(out, inv)).reshape(bsz, 2 ** n)


def full_unitary(mat: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    """Brute force: the 2^n x 2^n operator of ``mat`` on ``wires``, built by
    explicit basis-index bookkeeping (independent of :func:`apply`)."""
    wires = list(wires)
    k = len(wires)
    d = 2 ** n
    big = np.zeros((d, d), dtype=C)
    for col in range(d):
        bits = [(col >> (n - 1 - w)) & 1 for w in range(n)]
        sub_in = 0
        for w in wires:
            sub_in = (sub_in << 1) | bits[w]
        for sub_out in range(2 ** k):
            amp = mat[sub_out, sub_in]
            if amp == 0:
                continue
            ob = list(bits)
            for i, w in enumerate(wires):
                ob[w] = (sub_out >> (k - 1 - i)) & 1
            row = 0
            for b in ob:
                row = (row << 1) | b
            big[row, col] += amp
    return big


def run_ops(ops: Iterable[dict], n: int, bsz: int = 1, state: np.ndarray | None = None) -> np.ndarray:
    """Run an op list ``[{name, wires, params, inverse}, ...]`` (the reference's
    op-history dict format, gate_wrapper.py:380-391) from |0..0> or ``state``."""
    s = zero_state(n, bsz) if state is None else np.array(state, dtype=C)
    for op in ops:
        wires = op["wires"]
        wires = [wires] if isinstance(wires, (int, np.integer)) else list(wires)
        p = op.get("params")
        is_mat = op["name"] in ("qubitunitary", "qubitunitaryfast")
        pa = None if p is None else np.asarray(p, dtype=C if is_mat else np.float64)
        if pa is not None and pa.ndim == 2 and pa.shape[0] == bsz and bsz > 1 and not is_mat:
            mat = np.stack([gate_matrix(op["name"], pa[b], len(wires)) for b in range(bsz)])
        else:
            mat = gate_matrix(op["name"], pa, len(wires))
        if op.get("inverse", False):
            mat = np.conj(np.swapaxes(mat, -1, -2))
        s = apply(s, mat, wires, n)
    return s


# ---------------------------------------------------------------------------
# expectations
# ---------------------------------------------------------------------------

def expval_z_all(state: np.ndarray, n: int) -> np.ndarray:
    """<Z_w> for every wire, ``[bsz, n]`` (measurements.py:282-312 with PauliZ)."""
    p = np.abs(state) ** 2
    idx = np.arange(2 ** n)
    out = np.empty((state.shape[0], n))
    for w in range(n):
        sign = 1 - 2 * ((idx >> (n - 1 - w)) & 1)
        out[:, w] = p @ sign
    return out


def pauli_string_matrix(obs: str) -> np.ndarray:
    """Dense Kronecker Hamiltonian of a Pauli string, char k <-> wire k
    (measurements.py:265-271).  Exponential in size: tiny n only."""
    table = {"I": _I2, "X": _X, "Y": _Y, "Z": _Z}
    m = np.array([[1.0]], dtype=C)
    for ch in obs.upper():
        m = np.kron(m, table[ch])
    return m


def expval_pauli(state: np.ndarray, obs: str) -> np.ndarray:
    """Re <psi|P|psi> per batch row, via flip/phase masks (one pass, any n).

    P = i^{n_Y} X^{x} Z^{z}: P|j> = i^{n_Y} (-1)^{popcount(j & z)} |j ^ x>.
    """
    n = len(obs)
    obs = obs.upper()
    x = z = 0
    ny = 0
    for w, ch in enumerate(obs):
        bit = 1 << (n - 1 - w)
        if ch in "XY":
            x |= bit
        if ch in "ZY":
            z |= bit
        if ch == "Y":
            ny += 1
    idx = np.arange(2 ** n)
    par = np.array([bin(v).count("1") & 1 for v in (idx & z)]) if z else np.zeros(2 ** n, dtype=int)
    phase = (1j ** ny) * (1 - 2 * par)
    # (P psi)_{j ^ x} = phase_j psi_j
    p_psi = np.empty_like(state)
    p_psi[:, idx ^ x] = state * phase
    return np.real(np.sum(np.conj(state) * p_psi, axis=1))


def expval_pauli_dense(state: np.ndarray, obs: str) -> np.ndarray:
    """The reference's own algorithm (kron Hamiltonian then psi^* . H psi),
    measurements.py:265-279 -- used to pin :func:`expval_pauli` on tiny n."""
    h = pauli_string_matrix(obs)
    return np.real(np.sum(np.conj(state) * (h @ state.T).T, axis=1))


# ---------------------------------------------------------------------------
# gradients
# ---------------------------------------------------------------------------

def central_fd(f, theta: np.ndarray, h: float = 1e-3, idx: Iterable[int] | None = None) -> np.ndarray:
    """Central finite differences of scalar ``f`` at ``theta`` (float64)."""
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    flat = theta.reshape(-1)
    gf = g.reshape(-1)
    for i in (range(flat.size) if idx is None else idx):
        old = flat[i]
        flat[i] = old + h
        fp = f(theta)
        flat[i] = old - h
        fm = f(theta)
        flat[i] = old
        gf[i] = (fp - fm) / (2 * h)
    return g
