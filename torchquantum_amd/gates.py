"""Gate registry: closed-form matrices and the structure the kernels exploit.

Matrix conventions follow the reference exactly (same phases, same wire
order: ``wires[0]`` is the most significant bit of the matrix index,
torchquantum/functional/gate_wrapper.py:120-127).  Builders are written as
pure torch functions of a ``[G, P]`` parameter tensor returning ``[G, d, d]``
complex matrices, so they serve three callers:

* the public ``*_matrix`` functions and ``Operator.matrix`` (torch, any device);
* the host-side plan compiler, which evaluates them in float64 on the CPU to
  fill the kernels' matrix pool;
* forward-mode AD (``torch.func.jacfwd``) for the derivative matrices the
  adjoint kernels need.

Structure ("kind") tells the dispatcher which specialised kernel applies:

* ``diag``  -- diagonal; applied as  amp *= phi[parity(bits & pmask)]  on the
  amplitudes whose control bits match (``ctrl``: wires that must be 1,
  ``parity``: wires whose parity selects phi_even / phi_odd);
* ``ctrl1`` -- a 2x2 block on the last wire, conditioned on the others
  (``cval``: required control values);
* ``swap``  -- a (controlled) exchange of the last two wires;
* ``dense`` -- anything else, by width (1, 2 or k wires).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Optional, Sequence

import numpy as np
import torch

INV_SQRT2 = 1.0 / math.sqrt(2.0)


def _cdtype(t: torch.Tensor) -> torch.dtype:
    return torch.complex128 if t.dtype == torch.float64 else torch.complex64


def _cx(re: torch.Tensor, im: Optional[torch.Tensor] = None) -> torch.Tensor:
    if im is None:
        im = torch.zeros_like(re)
    return torch.complex(re, im)


def _expi(x: torch.Tensor) -> torch.Tensor:
    return torch.complex(torch.cos(x), torch.sin(x))


def _mat(rows) -> torch.Tensor:
    """Stack a nested list (d x d) of [G] complex tensors into [G, d, d]."""
    return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)


def _const(m: np.ndarray) -> Callable[[torch.Tensor], torch.Tensor]:
    arr = np.asarray(m, dtype=np.complex128)

    def build(p: torch.Tensor) -> torch.Tensor:
        g = p.shape[0]
        t = torch.as_tensor(arr, dtype=_cdtype(p) if p.is_floating_point() else torch.complex64,
                            device=p.device)
        return t.unsqueeze(0).expand(g, *arr.shape).clone()

    return build


def _controlled_np(u: np.ndarray, n_ctrl: int) -> np.ndarray:
    d = u.shape[0] << n_ctrl
    m = np.eye(d, dtype=np.complex128)
    m[d - u.shape[0]:, d - u.shape[0]:] = u
    return m


def _embed_ctrl(u: torch.Tensor, n_ctrl: int) -> torch.Tensor:
    """[G,2,2] -> [G, 2^(n+1), 2^(n+1)] with identity except the last block."""
    g = u.shape[0]
    d = 2 << n_ctrl
    m = torch.eye(d, dtype=u.dtype, device=u.device).unsqueeze(0).expand(g, d, d)
    top = m[:, : d - 2, :]
    bot = torch.cat([torch.zeros(g, 2, d - 2, dtype=u.dtype, device=u.device), u], dim=-1)
    return torch.cat([top, bot], dim=-2)


# ----------------------------------------------------------------------------
# parameterised builders ([G, P] real -> [G, d, d] complex)
# ----------------------------------------------------------------------------

def _rx(p):
    t = p[:, 0]
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    z = torch.zeros_like(c)
    return _mat([[_cx(c), _cx(z, -s)], [_cx(z, -s), _cx(c)]])          # functional/rx.py:19-46


def _ry(p):
    t = p[:, 0]
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    return _mat([[_cx(c), _cx(-s)], [_cx(s), _cx(c)]])                  # functional/ry.py:50-67


def _rz(p):
    t = p[:, 0]
    e = _expi(-t / 2)
    z = torch.zeros_like(e)
    return _mat([[e, z], [z, e.conj()]])                                # functional/rz.py:150-171


def _phase(p):
    t = p[:, 0]
    e = _expi(t)
    one, z = torch.ones_like(e), torch.zeros_like(e)
    return _mat([[one, z], [z, e]])                                     # phase_shift.py, u1.py


def _u2(p):
    ph, lam = p[:, 0], p[:, 1]
    one = _cx(torch.ones_like(ph))
    return INV_SQRT2 * _mat([[one, -_expi(lam)], [_expi(ph), _expi(ph + lam)]])  # functional/u2.py


def _u3(p):
    t, ph, lam = p[:, 0], p[:, 1], p[:, 2]
    c, s = _cx(torch.cos(t / 2)), _cx(torch.sin(t / 2))
    return _mat([[c, -s * _expi(lam)], [s * _expi(ph), c * _expi(ph + lam)]])  # functional/u3.py:52-77


def _rot(p):
    ph, t, om = p[:, 0], p[:, 1], p[:, 2]
    c, s = _cx(torch.cos(t / 2)), _cx(torch.sin(t / 2))
    return _mat([[_expi(-(ph + om) / 2) * c, -_expi((ph - om) / 2) * s],
                 [_expi(-(ph - om) / 2) * s, _expi((ph + om) / 2) * c]])  # functional/rot.py


def _r(p):
    t, ph = p[:, 0], p[:, 1]
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    e = _expi(-ph)
    js = _cx(torch.zeros_like(s), -s)
    return _mat([[_cx(c), e * js], [e.conj() * js, _cx(c)]])           # functional/r.py


def _z4(p):
    return torch.zeros(p.shape[0], dtype=_cdtype(p), device=p.device)


def _rxx(p):
    t = p[:, 0]
    c, s = _cx(torch.cos(t / 2)), torch.sin(t / 2)
    z = _z4(p)
    ms = _cx(torch.zeros_like(s), -s)
    return _mat([[c, z, z, ms], [z, c, ms, z], [z, ms, c, z], [ms, z, z, c]])  # functional/rx.py rxx


def _ryy(p):
    t = p[:, 0]
    c, s = _cx(torch.cos(t / 2)), torch.sin(t / 2)
    z = _z4(p)
    js = _cx(torch.zeros_like(s), s)
    return _mat([[c, z, z, js], [z, c, -js, z], [z, -js, c, z], [js, z, z, c]])  # functional/ry.py ryy


def _rzz(p):
    t = p[:, 0]
    e = _expi(-t / 2)
    z = _z4(p)
    ec = e.conj()
    return _mat([[e, z, z, z], [z, ec, z, z], [z, z, ec, z], [z, z, z, e]])  # functional/rz.py rzz


def _rzx(p):
    t = p[:, 0]
    c, s = _cx(torch.cos(t / 2)), torch.sin(t / 2)
    z = _z4(p)
    js = _cx(torch.zeros_like(s), s)
    return _mat([[c, -js, z, z], [-js, c, z, z], [z, z, c, js], [z, z, js, c]])  # functional/rz.py rzx


def _xxminyy(p):
    t, b = p[:, 0], p[:, 1]
    c, s = _cx(torch.cos(t / 2)), torch.sin(t / 2)
    z = _z4(p)
    one = _cx(torch.ones_like(t))
    mis = _cx(torch.zeros_like(s), -s)
    return _mat([[c, z, z, mis * _expi(-b)], [z, one, z, z], [z, z, one, z],
                 [mis * _expi(b), z, z, c]])                            # functional/xx_min_yy.py


def _xxplusyy(p):
    t, b = p[:, 0], p[:, 1]
    c, s = _cx(torch.cos(t / 2)), torch.sin(t / 2)
    z = _z4(p)
    one = _cx(torch.ones_like(t))
    mis = _cx(torch.zeros_like(s), -s)
    return _mat([[one, z, z, z], [z, c, mis * _expi(b), z], [z, mis * _expi(-b), c, z],
                 [z, z, z, one]])                                       # functional/xx_plus_yy.py


def _single_exc(p):
    t = p[:, 0]
    c, s = _cx(torch.cos(t / 2)), _cx(torch.sin(t / 2))
    z = _z4(p)
    one = _cx(torch.ones_like(t))
    return _mat([[one, z, z, z], [z, c, -s, z], [z, s, c, z], [z, z, z, one]])  # single_excitation.py


def _cu(p):
    t, ph, lam, g = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
    c, s = _cx(torch.cos(t / 2)), _cx(torch.sin(t / 2))
    u = _mat([[c * _expi(g), -s * _expi(lam + g)], [s * _expi(ph + g), c * _expi(ph + lam + g)]])
    return _embed_ctrl(u, 1)                                            # functional/u3.py cu_matrix


def _cu2(p):
    ph, lam = p[:, 0], p[:, 1]
    one = _cx(torch.ones_like(ph))
    # functional/u2.py cu2_matrix: the reference does not normalise this block
    u = _mat([[one, -_expi(lam)], [_expi(ph), _expi(ph + lam)]])
    return _embed_ctrl(u, 1)


def _ctrl_of(b1, n_ctrl=1):
    return lambda p: _embed_ctrl(b1(p), n_ctrl)


def _multirz_builder(k: int):
    sign = np.array([1 - 2 * (bin(j).count("1") & 1) for j in range(2 ** k)], dtype=np.float64)

    def build(p):
        t = p[:, 0]
        sg = torch.as_tensor(sign, dtype=p.dtype, device=p.device)
        e = _expi(-0.5 * t[:, None] * sg[None, :])                        # functional/rz.py:19-48
        return torch.diag_embed(e)

    return build


def _globalphase(p):
    return _expi(p[:, 0]).reshape(-1, 1, 1)                             # functional/global_phase.py


# ----------------------------------------------------------------------------
# fixed matrices (numpy complex128)
# ----------------------------------------------------------------------------
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_H = INV_SQRT2 * np.array([[1, 1], [1, -1]], dtype=np.complex128)
_SX = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], dtype=np.complex128)
_S = np.array([[1, 0], [0, 1j]], dtype=np.complex128)
_SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def _rccx_np():
    m = np.eye(8, dtype=np.complex128)
    m[5, 5] = -1
    m[6:, 6:] = [[0, -1j], [1j, 0]]
    return m                                                            # functional/paulix.py:116-127


def _rc3x_np():
    m = np.eye(16, dtype=np.complex128)
    m[12:, 12:] = [[1j, 0, 0, 0], [0, -1j, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]]
    return m                                                            # functional/paulix.py:129-150


FIXED_NP: Dict[str, np.ndarray] = {
    "hadamard": _H,
    "shadamard": np.array([[math.cos(math.pi / 8), -math.sin(math.pi / 8)],
                           [math.sin(math.pi / 8), math.cos(math.pi / 8)]], dtype=np.complex128),
    "paulix": _X, "pauliy": _Y, "pauliz": _Z, "i": np.eye(2, dtype=np.complex128),
    "s": _S, "sdg": _S.conj(),
    "t": np.array([[1, 0], [0, np.exp(1j * math.pi / 4)]], dtype=np.complex128),
    "tdg": np.array([[1, 0], [0, np.exp(-1j * math.pi / 4)]], dtype=np.complex128),
    "sx": _SX, "sxdg": _SX.conj().T,
    "cnot": _controlled_np(_X, 1), "cy": _controlled_np(_Y, 1), "cz": _controlled_np(_Z, 1),
    "chadamard": _controlled_np(_H, 1), "cs": _controlled_np(_S, 1), "csdg": _controlled_np(_S.conj(), 1),
    "csx": _controlled_np(_SX, 1),
    "swap": _SWAP,
    "sswap": np.array([[1, 0, 0, 0], [0, (1 + 1j) / 2, (1 - 1j) / 2, 0],
                       [0, (1 - 1j) / 2, (1 + 1j) / 2, 0], [0, 0, 0, 1]], dtype=np.complex128),
    "iswap": np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], dtype=np.complex128),
    "dcx": np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [0, 1, 0, 0]], dtype=np.complex128),
    "ecr": INV_SQRT2 * np.array([[0, 0, 1, 1j], [0, 0, 1j, 1], [1, -1j, 0, 0], [-1j, 1, 0, 0]],
                                dtype=np.complex128),
    "toffoli": _controlled_np(_X, 2), "ccz": _controlled_np(_Z, 2),
    "cswap": _controlled_np(_SWAP, 1),
    "rccx": _rccx_np(), "rc3x": _rc3x_np(),
    "c3x": _controlled_np(_X, 3), "c4x": _controlled_np(_X, 4), "c3sx": _controlled_np(_SX, 3),
}


# ----------------------------------------------------------------------------
# registry
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class GateSpec:
    name: str
    num_wires: Optional[int]           # None: any (given by n_wires / len(wires))
    num_params: int                    # 0 for fixed gates; -1: a matrix parameter
    kind: str                          # diag | ctrl1 | swap | dense
    builder: Optional[Callable] = None  # [G,P] -> [G,d,d] for fixed-width gates
    # diag: control wires (value 1) and parity wires, as indices into `wires`
    ctrl_idx: tuple = ()
    parity_idx: tuple = ()
    cval0: bool = False                 # ctrl1 with controls required to be 0 (multixcnot)


def _spec(name, nw, npar, kind, builder, **kw):
    return GateSpec(name, nw, npar, kind, builder, **kw)


REGISTRY: Dict[str, GateSpec] = {}


def _add(s: GateSpec):
    REGISTRY[s.name] = s


# 1-wire fixed
for _n in ("hadamard", "shadamard", "paulix", "pauliy", "sx", "sxdg"):
    _add(_spec(_n, 1, 0, "dense", _const(FIXED_NP[_n])))
for _n in ("pauliz", "s", "sdg", "t", "tdg"):
    _add(_spec(_n, 1, 0, "diag", _const(FIXED_NP[_n]), ctrl_idx=(0,)))
_add(_spec("i", 1, 0, "diag", _const(FIXED_NP["i"]), ctrl_idx=(0,)))
# 1-wire parameterised
_add(_spec("rx", 1, 1, "dense", _rx))
_add(_spec("ry", 1, 1, "dense", _ry))
_add(_spec("rz", 1, 1, "diag", _rz, parity_idx=(0,)))
_add(_spec("phaseshift", 1, 1, "diag", _phase, ctrl_idx=(0,)))
_add(_spec("u1", 1, 1, "diag", _phase, ctrl_idx=(0,)))
_add(_spec("u2", 1, 2, "dense", _u2))
_add(_spec("u3", 1, 3, "dense", _u3))
_add(_spec("rot", 1, 3, "dense", _rot))
_add(_spec("r", 1, 2, "dense", _r))
# 2-wire fixed
for _n in ("cnot", "cy", "chadamard", "csx"):
    _add(_spec(_n, 2, 0, "ctrl1", _const(FIXED_NP[_n])))
for _n in ("cz", "cs", "csdg"):
    _add(_spec(_n, 2, 0, "diag", _const(FIXED_NP[_n]), ctrl_idx=(0, 1)))
_add(_spec("swap", 2, 0, "swap", _const(FIXED_NP["swap"])))
for _n in ("sswap", "iswap", "dcx", "ecr"):
    _add(_spec(_n, 2, 0, "dense", _const(FIXED_NP[_n])))
# 2-wire parameterised
_add(_spec("rxx", 2, 1, "dense", _rxx))
_add(_spec("ryy", 2, 1, "dense", _ryy))
_add(_spec("rzz", 2, 1, "diag", _rzz, parity_idx=(0, 1)))
_add(_spec("rzx", 2, 1, "dense", _rzx))
_add(_spec("crx", 2, 1, "ctrl1", _ctrl_of(_rx)))
_add(_spec("cry", 2, 1, "ctrl1", _ctrl_of(_ry)))
_add(_spec("crz", 2, 1, "diag", _ctrl_of(_rz), ctrl_idx=(0,), parity_idx=(1,)))
_add(_spec("crot", 2, 3, "ctrl1", _ctrl_of(_rot)))
_add(_spec("cu1", 2, 1, "diag", _ctrl_of(_phase), ctrl_idx=(0, 1)))
_add(_spec("cu2", 2, 2, "ctrl1", _cu2))
_add(_spec("cu3", 2, 3, "ctrl1", _ctrl_of(_u3)))
_add(_spec("cu", 2, 4, "ctrl1", _cu))
_add(_spec("singleexcitation", 2, 1, "dense", _single_exc))
_add(_spec("xxminyy", 2, 2, "dense", _xxminyy))
_add(_spec("xxplusyy", 2, 2, "dense", _xxplusyy))
# 3+ wire fixed
_add(_spec("toffoli", 3, 0, "ctrl1", _const(FIXED_NP["toffoli"])))
_add(_spec("ccz", 3, 0, "diag", _const(FIXED_NP["ccz"]), ctrl_idx=(0, 1, 2)))
_add(_spec("cswap", 3, 0, "swap", _const(FIXED_NP["cswap"])))
_add(_spec("rccx", 3, 0, "dense", _const(FIXED_NP["rccx"])))
_add(_spec("rc3x", 4, 0, "dense", _const(FIXED_NP["rc3x"])))
_add(_spec("c3x", 4, 0, "ctrl1", _const(FIXED_NP["c3x"])))
_add(_spec("c3sx", 4, 0, "ctrl1", _const(FIXED_NP["c3sx"])))
_add(_spec("c4x", 5, 0, "ctrl1", _const(FIXED_NP["c4x"])))
# variable width
_add(_spec("multicnot", None, 0, "ctrl1", None))
_add(_spec("multixcnot", None, 0, "ctrl1", None, cval0=True))
_add(_spec("multirz", None, 1, "diag", None))
_add(_spec("qft", None, 0, "dense", None))
_add(_spec("globalphase", 0, 1, "diag", _globalphase))
_add(_spec("qubitunitary", None, -1, "dense", None))
_add(_spec("qubitunitaryfast", None, -1, "dense", None))
_add(_spec("qubitunitarystrict", None, -1, "dense", None))

ALIASES = {
    "h": "hadamard", "sh": "shadamard", "ch": "chadamard", "x": "paulix", "y": "pauliy",
    "z": "pauliz", "cx": "cnot", "ccnot": "toffoli", "ccx": "toffoli",
    "xx": "rxx", "yy": "ryy", "zz": "rzz", "zx": "rzx", "p": "phaseshift",
    "cp": "cu1", "cr": "cu1", "cphase": "cu1", "u": "u3", "echoedcrossresonance": "ecr",
}


def canonical(name: str) -> str:
    name = name.lower()
    return ALIASES.get(name, name)


def spec(name: str) -> GateSpec:
    try:
        return REGISTRY[canonical(name)]
    except KeyError:
        raise ValueError(f"unknown gate '{name}'") from None


def width_builder(name: str, k: int) -> Callable:
    """Builder for a variable-width gate at width ``k``."""
    name = canonical(name)
    if name == "multicnot":                                           # functional/paulix.py:19-35
        return _const(_controlled_np(_X, k - 1))
    if name == "multixcnot":                                          # functional/paulix.py:38-55
        m = np.eye(2 ** k, dtype=np.complex128)
        m[[0, 1], [0, 1]] = 0
        m[0, 1] = m[1, 0] = 1
        return _const(m)
    if name == "multirz":
        return _multirz_builder(k)
    if name == "qft":                                                 # functional/qft.py
        d = 2 ** k
        j = np.arange(d)
        m = np.exp(2j * np.pi * np.outer(j, j) / d) / math.sqrt(d)
        return _const(m)
    s = spec(name)
    if s.builder is None:
        raise ValueError(f"gate '{name}' has no fixed builder")
    return s.builder


def build_matrices(name: str, params: Optional[torch.Tensor], k: int, rows: int = 1,
                   dtype=torch.float64, device="cpu") -> torch.Tensor:
    """[rows, 2^k, 2^k] matrices of gate ``name`` for a [rows, P] parameter tensor."""
    b = width_builder(name, k)
    if params is None:
        p = torch.zeros(rows, 1, dtype=dtype, device=device)
    else:
        p = params.to(dtype=dtype, device=device).reshape(params.shape[0] if params.dim() > 1 else 1, -1)
    return b(p)


def ctrl_values(s: GateSpec, k: int):
    """(control wire indices, required values, target index) for ctrl1 gates."""
    if s.cval0:
        return list(range(k - 1)), [0] * (k - 1), k - 1
    return list(range(k - 1)), [1] * (k - 1), k - 1


def diag_form(s: GateSpec, k: int):
    """(ctrl wire indices, parity wire indices) for diag gates of width k."""
    if s.name == "multirz":
        return (), tuple(range(k))
    return s.ctrl_idx, s.parity_idx


def diag_phase_indices(ctrl_idx: Sequence[int], parity_idx: Sequence[int], k: int):
    """Matrix-diagonal indices holding phi_even and phi_odd."""
    base = 0
    for i in ctrl_idx:
        base |= 1 << (k - 1 - i)
    odd = base
    if parity_idx:
        odd |= 1 << (k - 1 - parity_idx[0])
    return base, odd
