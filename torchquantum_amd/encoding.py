"""Encoders: classical inputs -> gates or amplitudes (torchquantum/encoding/encodings.py).

``GeneralEncoder`` feeds columns of a [bsz, F] input as per-row parameters of
the gates in its ``func_list`` (encodings.py:50-102).  On this engine those
gates' matrices are built on the GPU from the input tensor itself
(``EngineOptions.device_gates``): no device-to-host copy of the inputs and no
host work per row, so an encoder + ansatz training step stays asynchronous.
"""
from __future__ import annotations

from typing import Iterable, List

import torch

from .functional import func_name_dict
from .macro import C_DTYPE
from .module import QuantumModule, static_support

__all__ = ["Encoder", "GeneralEncoder", "PhaseEncoder", "MultiPhaseEncoder", "StateEncoder", "AmplitudeEncoder",
           "MagnitudeEncoder", "encoder_op_list_name_dict"]


def _n_params(func: str) -> int:
    from .operator import op_name_dict
    return op_name_dict[func].num_params


class Encoder(QuantumModule):
    """Base class: ``forward(qdev, x)`` encodes ``x`` into ``qdev`` (encodings.py:35-47)."""

    def forward(self, qdev, x):
        raise NotImplementedError


class GeneralEncoder(Encoder):
    """Gates from a list of {"func", "wires", "input_idx"} dicts; gate k takes
    ``x[:, input_idx]`` as its per-row parameters (encodings.py:50-102)."""

    def __init__(self, func_list):
        super().__init__()
        self.func_list = func_list

    @static_support
    def forward(self, qdev, x):
        for info in self.func_list:
            params = x[:, info["input_idx"]] if _n_params(info["func"]) > 0 else None   # advanced index: a copy
            func_name_dict[info["func"]](qdev, wires=info["wires"], params=params, static=self.static_mode,
                                         parent_graph=self.graph)


class PhaseEncoder(Encoder):
    """``func`` on every wire k with parameter ``x[:, k]`` (encodings.py:122-146)."""

    def __init__(self, func):
        super().__init__()
        self.func = func

    @static_support
    def forward(self, qdev, x):
        for k in range(qdev.n_wires):
            self.func(qdev, wires=k, params=x[:, k], static=self.static_mode, parent_graph=self.graph)


_STRIDE = {"rx": 1, "ry": 1, "rz": 1, "u1": 1, "phaseshift": 1, "u2": 2, "u3": 3}


class MultiPhaseEncoder(Encoder):
    """A list of one-wire rotations on ``wires`` (default: every wire, repeated),
    taking consecutive input columns (1, 2 or 3 per gate) (encodings.py:149-205)."""

    def __init__(self, funcs, wires=None):
        super().__init__()
        self.funcs = funcs if isinstance(funcs, Iterable) and not isinstance(funcs, str) else [funcs]
        self.wires = wires

    @static_support
    def forward(self, qdev, x):
        if self.wires is None:
            self.wires = list(range(qdev.n_wires)) * (len(self.funcs) // qdev.n_wires)
        col = 0
        for k, func in enumerate(self.funcs):
            if func not in _STRIDE:
                raise ValueError(func)
            st = _STRIDE[func]
            func_name_dict[func](qdev, wires=self.wires[k], params=x[:, col:col + st], static=self.static_mode,
                                 parent_graph=self.graph)
            col += st


class StateEncoder(Encoder):
    """Amplitude encoding: the normalised input row, zero-padded to 2^n
    amplitudes, becomes the state (encodings.py:208-240)."""

    def forward(self, qdev, x):
        x = x / torch.sqrt((x.abs() ** 2).sum(dim=-1)).unsqueeze(-1)
        pad = (1 << qdev.n_wires) - x.shape[1]
        if pad < 0:
            raise ValueError(f"StateEncoder: {x.shape[1]} inputs exceed 2^{qdev.n_wires} amplitudes")
        state = torch.cat([x, torch.zeros(x.shape[0], pad, device=x.device, dtype=x.dtype)], dim=-1)
        qdev.set_states(state.to(C_DTYPE).view([x.shape[0]] + [2] * qdev.n_wires))


class MagnitudeEncoder(Encoder):
    """Placeholder class of the reference (encodings.py:243-247): no forward."""


AmplitudeEncoder = StateEncoder


# ---------------------------------------------------------------------------
# the reference's named encoder layouts (encodings.py:249-560), generated
# ---------------------------------------------------------------------------
def _cycle(funcs: List[str], n_wires: int, n_inputs: int) -> List[dict]:
    """Input column i drives funcs[(i // n_wires) % len(funcs)] on wire i % n_wires."""
    return [{"input_idx": [i], "func": funcs[(i // n_wires) % len(funcs)], "wires": [i % n_wires]}
            for i in range(n_inputs)]


def _u3_rx(n_wires: int, h_after_first: bool) -> List[dict]:
    out = []
    for w in range(n_wires):
        out.append({"input_idx": [4 * w, 4 * w + 1, 4 * w + 2], "func": "u3", "wires": [w]})
        out.append({"input_idx": [4 * w + 3], "func": "rx", "wires": [w]})
        if h_after_first and w == 0:
            out += [{"func": "h", "wires": [v]} for v in range(n_wires)]
    return out


def _rz_sx(n_wires: int, n_inputs: int) -> List[dict]:
    out = []
    for i in range(n_inputs):
        out.append({"input_idx": [i], "func": "rz", "wires": [i % n_wires]})
        out.append({"input_idx": None, "func": "sx", "wires": [i % n_wires]})
    return out


encoder_op_list_name_dict = {
    "1x1_ry": _cycle(["ry"], 1, 1),
    "2x1_ryry": _cycle(["ry"], 2, 2),
    "2x8_rxryrzrxryrzrxry": _cycle(["rx", "ry", "rz", "rx", "ry", "rz", "rx", "ry"], 2, 16),
    "3x1_ryryry": _cycle(["ry"], 3, 3),
    "3x1_rxrxrx": _cycle(["rx"], 3, 3),
    "4_ry": _cycle(["ry"], 4, 4),
    "4x4_u3rx": _u3_rx(4, False),
    "4x4_u3_h_rx": _u3_rx(4, True),
    "4x4_ryzxy": _cycle(["ry", "rz", "rx", "ry"], 4, 16),
    "8x2_ry": _cycle(["ry"], 8, 16),
    "16_ry": _cycle(["ry"], 16, 16),
    "4x4_rzsx": _rz_sx(4, 16),
    "15_ryrz": _cycle(["ry", "rz"], 15, 16),
    "2x8_ryzxyzxyz": _cycle(["ry", "rz", "rx", "ry", "rz", "rx", "ry", "rz"], 2, 16),
    "10_ryzx": _cycle(["ry", "rz", "rx"], 4, 10),
    "10_ry": _cycle(["ry"], 10, 10),
    "25_ry": _cycle(["ry"], 25, 25),
    "25_ryrz": _cycle(["ry", "rz"], 21, 25),
    "6x6_ryzxy": _cycle(["ry", "rz", "rx", "ry"], 10, 36),
    "6x6_ryrz": _cycle(["ry", "rz"], 21, 36),
    "6x6_ryrzrx": _cycle(["ry", "rz", "rx"], 16, 36),
    "8x8_ryzxyzxy": _cycle(["ry", "rz", "rx"] * 3, 10, 64),
    "10x10_ryzxyzxyzxy": _cycle(["ry", "rz", "rx"] * 4, 10, 100),
    "8x2_ryz": _cycle(["ry", "rz"], 8, 16),
    "16x1_ry": _cycle(["ry"], 16, 16),
}
