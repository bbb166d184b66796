"""Functional gate API (torchquantum/functional/__init__.py:97-218).

``tqf.<gate>(q_device, wires, params=None, n_wires=None, static=False,
parent_graph=None, inverse=False, comp_method="bmm")`` for every gate name of
the reference, plus the ``*_matrix`` builders, ``mat_dict`` and
``func_name_dict``.  A gate call validates its arguments and records the gate
on the device's tape (``gate_wrapper``); the HIP engine applies it.
``comp_method`` is accepted for compatibility: every method runs the same
kernels (the reference silently skipped a gate for an unknown method,
gate_wrapper.py:441-446; here an unknown method raises).
"""
from __future__ import annotations

from typing import Callable, Dict

import numpy as np
import torch

from .. import gates as G
from ..engine import TapeOp
from ..macro import C_DTYPE, F_DTYPE

_MATRIX_GATES = ("qubitunitary", "qubitunitaryfast", "qubitunitarystrict")


def _normalize_params(name: str, params):
    """Parameter normalisation of gate_wrapper.py:361-378."""
    if params is None:
        return None
    if name in _MATRIX_GATES:
        if not isinstance(params, torch.Tensor):
            params = torch.tensor(np.asarray(params), dtype=C_DTYPE)
        if not params.is_complex():
            params = params.to(C_DTYPE)
        if params.dim() == 2:
            params = params.unsqueeze(0)
        return params
    if not isinstance(params, torch.Tensor):
        params = torch.tensor(params, dtype=F_DTYPE)
    if params.is_complex():
        raise TypeError(f"gate '{name}' takes real parameters")
    if params.dim() == 1:
        params = params.unsqueeze(-1)
    elif params.dim() == 0:
        params = params.unsqueeze(-1).unsqueeze(-1)
    return params


def _wire_list(wires):
    if isinstance(wires, (int, np.integer)):
        return [int(wires)]
    if isinstance(wires, (list, tuple)) and all(type(w) is int for w in wires):
        return list(wires)
    return [int(w) for w in np.array(wires).reshape(-1)]


def gate_wrapper(name, q_device, wires, params=None, n_wires=None, static=False, parent_graph=None,
                 inverse=False, comp_method="bmm"):
    """Record one gate on ``q_device`` (the reference's choke point,
    torchquantum/functional/gate_wrapper.py:324-446)."""
    cname = G.canonical(name)
    _record(name, cname, G.spec(cname), q_device, wires, params, inverse, comp_method)


def _record(name, cname, s, q_device, wires, params, inverse, comp_method):
    if comp_method not in ("bmm", "einsum"):
        raise ValueError(f"unknown comp_method '{comp_method}'")
    params = _normalize_params(cname, params)
    wires = _wire_list(wires)
    if getattr(q_device, "record_op", False):
        q_device.op_history.append({
            "name": name,
            "wires": np.array(wires).squeeze().tolist(),
            "params": params.squeeze().detach().cpu().numpy().tolist() if params is not None else None,
            "inverse": inverse,
            "trainable": params.requires_grad if params is not None else False,
        })
    # static mode (gate_wrapper.py:393-403 hands the gate to a QuantumGraph):
    # here every gate goes on the device tape; the graph-top module's
    # static_support flush compiles the tape with its wires_per_block.
    n = q_device.n_wires
    if len(set(wires)) != len(wires):
        raise ValueError(f"gate '{name}' applied to repeated wires {wires}")
    for w in wires:
        if not 0 <= w < n:
            raise ValueError(f"gate '{name}': wire {w} out of range for {n} wires")
    k = len(wires)
    if s.num_wires is not None and s.num_wires != k:
        raise ValueError(f"gate '{name}' acts on {s.num_wires} wires, got {k}")
    if cname in ("multicnot", "multixcnot") and k < 2:
        raise ValueError(f"gate '{name}' needs at least 2 wires")
    if s.num_params > 0:
        if params is None:
            raise ValueError(f"gate '{name}' needs {s.num_params} parameter(s)")
        if params.shape[-1] != s.num_params:
            if params.dim() == 2 and params.shape[-1] == 1 and params.shape[0] == s.num_params:
                params = params.reshape(1, s.num_params)
            else:
                raise ValueError(f"gate '{name}' takes {s.num_params} parameters per row, got shape "
                                 f"{tuple(params.shape)}")
    elif s.num_params == 0:
        params = None
    else:
        if params is None:
            raise ValueError(f"gate '{name}' needs a matrix parameter")
        if params.shape[-1] != 1 << k or params.shape[-2] != 1 << k:
            raise ValueError(f"gate '{name}': matrix {tuple(params.shape[-2:])} does not match {k} wires")
        if cname == "qubitunitary":
            u = params.detach().cpu().to(torch.complex128).numpy()
            eye = np.eye(u.shape[-1])
            if not np.allclose(u @ np.conj(np.swapaxes(u, -1, -2)), eye, atol=1e-5):
                raise ValueError("Operator must be unitary.")                 # qubit_unitary.py:27-55
        if cname == "qubitunitarystrict":
            uu, _, vh = torch.linalg.svd(params.to(torch.complex128))
            params = (uu @ vh).to(C_DTYPE)                                    # qubit_unitary.py:72-86
        cname = "qubitunitaryfast"
    q_device._push(TapeOp(cname, tuple(wires), params, bool(inverse),
                          params._version if params is not None else 0))


def _make_gate(name: str):
    cname = G.canonical(name)
    spec = G.spec(cname)
    nw, npar = spec.num_wires, spec.num_params
    plain = cname not in _MATRIX_GATES and cname not in ("multicnot", "multixcnot") and nw is not None

    def gate(q_device, wires, params=None, n_wires=None, static=False, parent_graph=None, inverse=False,
             comp_method="bmm"):
        # fast path, same checks as _record: a fixed-width gate on int wires
        # with no parameter or a real [rows, P] tensor, no op history
        if plain and comp_method == "bmm" and not getattr(q_device, "record_op", False):
            tw = type(wires)
            ws = None
            if tw is int:
                ws = (wires,)
            elif tw is list or tw is tuple:
                ws = tuple(wires)
                for w in ws:
                    if type(w) is not int:
                        ws = None
                        break
            if ws is not None and len(ws) == nw and (nw == 1 or len(set(ws)) == nw):
                n = q_device.n_wires
                ok = True
                for w in ws:
                    if w < 0 or w >= n:
                        ok = False
                        break
                if ok:
                    if npar == 0:
                        q_device._push(TapeOp(cname, ws, None, bool(inverse), 0))
                        return
                    if (type(params) is torch.Tensor or isinstance(params, torch.Tensor)) and params.dim() == 2 \
                            and params.shape[-1] == npar and not params.is_complex():
                        q_device._push(TapeOp(cname, ws, params, bool(inverse), params._version))
                        return
        _record(name, cname, spec, q_device, wires, params, inverse, comp_method)

    gate.__name__ = name
    gate.__qualname__ = name
    gate.__doc__ = f"Apply the {name} gate (torchquantum.functional.{name} semantics)."
    return gate


# ---------------------------------------------------------------- matrices
def _matrix_fn(name: str, nparams: int):
    builder = G.width_builder(name, G.spec(name).num_wires or 1)

    def fn(params):
        p = params if isinstance(params, torch.Tensor) else torch.tensor(params, dtype=F_DTYPE)
        if p.dim() == 0:
            p = p.reshape(1, 1)
        elif p.dim() == 1:
            p = p.reshape(1, -1) if nparams > 1 else p.unsqueeze(-1)
        return builder(p).squeeze(0)

    fn.__name__ = f"{name}_matrix"
    return fn


def _width_matrix_fn(name: str):
    def fn(*args):
        if name == "multirz":
            params, k = args
            p = params if isinstance(params, torch.Tensor) else torch.tensor(params, dtype=F_DTYPE)
            p = p.reshape(-1, 1)
            return G.width_builder(name, k)(p).squeeze(0)
        (k,) = args
        return G.width_builder(name, k)(torch.zeros(1, 1, dtype=F_DTYPE)).squeeze(0)

    fn.__name__ = f"{name}_matrix"
    return fn


mat_dict: Dict[str, object] = {}
for _name, _s in G.REGISTRY.items():
    if _s.num_params == 0 and _s.num_wires is not None:
        mat_dict[_name] = torch.tensor(G.FIXED_NP[_name] if _name in G.FIXED_NP else
                                       G.width_builder(_name, _s.num_wires)(torch.zeros(1, 1)).squeeze(0).numpy(),
                                       dtype=C_DTYPE)
    elif _s.num_params > 0 and _s.num_wires is not None:
        mat_dict[_name] = _matrix_fn(_name, _s.num_params)
    elif _name in ("multicnot", "multixcnot", "multirz", "qft"):
        mat_dict[_name] = _width_matrix_fn(_name)
mat_dict["qubitunitary"] = lambda params: torch.as_tensor(params).squeeze(0)
mat_dict["qubitunitaryfast"] = lambda params: torch.as_tensor(params).squeeze(0)


def _qus(params):
    u, _, vh = torch.linalg.svd(torch.as_tensor(params))
    return u @ vh


mat_dict["qubitunitarystrict"] = _qus

# module-level *_matrix names (rx_matrix, u3_matrix, ...)
for _name, _v in list(mat_dict.items()):
    if callable(_v):
        globals()[f"{_name}_matrix"] = _v

FUNC_NAMES = [
    "hadamard", "h", "sh", "paulix", "pauliy", "pauliz", "i", "s", "t", "sx", "cnot", "cz", "cy", "rx", "ry",
    "rz", "rxx", "xx", "ryy", "yy", "rzz", "zz", "rzx", "zx", "swap", "sswap", "cswap", "toffoli",
    "phaseshift", "p", "cp", "rot", "multirz", "crx", "cry", "crz", "crot", "u1", "u2", "u3", "u", "cu1",
    "cphase", "cr", "cu2", "cu3", "cu", "qubitunitary", "qubitunitaryfast", "qubitunitarystrict", "multicnot",
    "multixcnot", "x", "y", "z", "cx", "ccnot", "ccx", "singleexcitation", "ecr", "echoedcrossresonance", "qft",
    "sdg", "tdg", "sxdg", "ch", "ccz", "iswap", "cs", "csdg", "csx", "chadamard", "dcx", "xxminyy", "xxplusyy",
    "c3x", "r", "globalphase", "c3sx", "rccx", "rc3x", "c4x", "shadamard",
]

func_name_dict: Dict[str, Callable] = {}
for _name in FUNC_NAMES:
    _f = _make_gate(_name)
    func_name_dict[_name] = _f
    globals()[_name] = _f


def reset(q_device, wires, inverse=False):
    """Project the given wires onto |0> and renormalise (functional/reset.py)."""
    import torch as _t
    wires = [wires] if isinstance(wires, int) else list(wires)
    st = q_device.get_states_1d()
    n = q_device.n_wires
    idx = _t.arange(1 << n, device=st.device)
    keep = _t.ones(1 << n, dtype=_t.bool, device=st.device)
    for w in wires:
        keep &= ((idx >> (n - 1 - w)) & 1) == 0
    st = st * keep.to(st.dtype)
    nrm = _t.linalg.vector_norm(st, dim=-1, keepdim=True)
    q_device.states = st / nrm


func_name_dict["reset"] = reset


def controlled_unitary(qdev, c_wires, t_wires, params):
    """Controlled-U over sets of target wires (functional/func_controlled_unitary.py:31-133):
    builds the block matrix and records it as one qubitunitaryfast gate."""
    if isinstance(params, np.ndarray):
        params = torch.from_numpy(params)
    params = torch.as_tensor(params).to(C_DTYPE)
    c = np.array(c_wires)
    t = np.array(t_wires)
    wires = [int(c)] if c.ndim == 0 else [int(x) for x in c]
    n_c = len(wires)
    k_u = int(np.log2(params.shape[-1]))
    tw = [int(t)] if t.ndim == 0 else [int(x) for x in t.reshape(-1)]
    n_t = len(tw)
    n_sets = n_t // k_u
    wires += tw
    nw = n_c + n_t
    cu = params
    for _ in range(n_sets - 1):
        cu = torch.kron(cu, params)
    d = 1 << nw
    dc = cu.shape[-1]
    u = torch.eye(d, dtype=C_DTYPE)
    u[d - dc:, d - dc:] = cu
    gate_wrapper("qubitunitaryfast", qdev, wires, u.to(qdev.device if hasattr(qdev, "device") else "cpu"))


def matrix_exp(qdev, wires, params):
    """Apply exp(params) (functional/func_mat_exp.py:34-72)."""
    m = torch.as_tensor(params).to(torch.complex128)
    u = torch.linalg.matrix_exp(m).to(C_DTYPE)
    gate_wrapper("qubitunitaryfast", qdev, wires, u)


func_name_dict_collect = {"matrix_exp": matrix_exp, "controlled_unitary": controlled_unitary}

__all__ = ["gate_wrapper", "mat_dict", "func_name_dict", "func_name_dict_collect", "reset",
           "controlled_unitary", "matrix_exp"] + FUNC_NAMES
