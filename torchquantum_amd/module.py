"""QuantumModule (torchquantum/module/modules.py:38-383): nn.Module with the
reference's static-mode attributes, op-history replay and ``get_unitary``."""
from __future__ import annotations

import functools
from typing import Iterable

import torch
import torch.nn as nn

__all__ = ["QuantumModule", "QuantumModuleList", "QuantumModuleDict", "static_support"]

# widest dense block the fused planner builds
MAX_STATIC_BLOCK = 6


def _find_device(module, args, kwargs):
    from .device import QuantumDevice
    for a in list(args) + list(kwargs.values()):
        if isinstance(a, QuantumDevice):
            return a
    q = getattr(module, "q_device", None)
    return q if isinstance(q, QuantumDevice) else None


def static_support(f):
    """Decorator for ``QuantumModule.forward`` (graph/graphs.py:60-81).

    In the reference, static mode records the gates of the forward into a
    QuantumGraph, packs them into blocks of ``wires_per_block`` wires and
    applies one unitary per block when the graph-top module's forward ends.
    Here the gates are already recorded on the device tape; when a graph-top
    module in static mode finishes its forward, the tape is compiled by the
    fused planner with fusion width ``wires_per_block`` (blocks wider than
    MAX_STATIC_BLOCK wires are built as MAX_STATIC_BLOCK-wire blocks, which
    gives the same state) and run.  Blocks of 5 and 6 wires run on the
    matrix-core kernel."""

    @functools.wraps(f)
    def forward_static(self, *args, **kwargs):
        res = f(self, *args, **kwargs)
        if getattr(self, "static_mode", False) and getattr(self, "is_graph_top", False):
            self.static_forward(_find_device(self, args, kwargs))
        return res

    return forward_static


class QuantumModule(nn.Module):
    def __init__(self) -> None:
        super().__init__()
        self.static_mode = False
        self.graph = None
        self.parent_graph = None
        self.is_graph_top = False
        self.unitary = None
        self.wires = None
        self.n_wires = None
        self.q_device = None
        self.device = None
        self.wires_per_block = None
        self.qiskit_processor = None
        self.noise_model_tq = None
        self.Operator_list = None

    def load_op_history(self, op_history):
        """Build operators from an op-history dict list (modules.py:92-133)."""
        from .operator import op_name_dict
        ops = []
        for op in op_history:
            Oper = op_name_dict[op["name"]]
            trainable = op.get("trainable", False)
            has_params = (op.get("params", None) is not None) or trainable
            wires = op["wires"]
            n_wires = len(wires) if isinstance(wires, Iterable) else 1
            ops.append(Oper(has_params=has_params, trainable=trainable, init_params=op.get("params", None),
                            n_wires=n_wires, wires=wires, inverse=op.get("inverse", False)))
        self.Operator_list = QuantumModuleList(ops)

    @classmethod
    def from_op_history(cls, op_history):
        qm = cls()
        qm.load_op_history(op_history)
        return qm

    def set_qiskit_processor(self, processor):
        self.qiskit_processor = processor

    def set_noise_model_tq(self, noise_model_tq):
        self.noise_model_tq = noise_model_tq

    def set_wires_per_block(self, wires_per_block):
        self.wires_per_block = wires_per_block

    def static_on(self, is_graph_top=True, wires_per_block=3):
        """Static mode (modules.py:230-261): this module and its quantum
        sub-modules record into one graph; the top module's forward (under
        ``static_support``) ends by running the recorded gates fused into
        blocks of ``wires_per_block`` wires."""
        if int(wires_per_block) < 1:
            raise ValueError(f"wires_per_block must be >= 1 (got {wires_per_block})")
        self.wires_per_block = int(wires_per_block)
        self.static_mode = True
        self.is_graph_top = is_graph_top
        for m in self.children():
            if isinstance(m, QuantumModule) and not _is_device(m):
                m.static_on(is_graph_top=False, wires_per_block=wires_per_block)

    def static_off(self):
        """modules.py:263-277"""
        self.static_mode = False
        self.graph = None
        for m in self.children():
            if isinstance(m, QuantumModule) and not _is_device(m):
                m.static_off()

    def static_options(self):
        """Engine options of a static-mode flush."""
        from .engine import EngineOptions
        width = max(1, min(int(self.wires_per_block or 1), MAX_STATIC_BLOCK))
        return EngineOptions(mode="fused", fusion_width=width)

    def static_forward(self, q_device):
        """Run the gates recorded on ``q_device`` as fused blocks of
        ``wires_per_block`` wires (modules.py:296-316)."""
        if q_device is None:
            raise RuntimeError("static_forward: no QuantumDevice among the forward arguments")
        self.q_device = q_device
        q_device.flush(self.static_options())

    def forward(self, q_device, *args, **kwargs):
        if self.Operator_list is not None:
            for op in self.Operator_list:
                op(q_device)
            return
        raise NotImplementedError

    def get_unitary(self, x=None):
        """Unitary of the module via the identity-state trick (modules.py:326-354)."""
        from .device import QuantumDevice
        assert self.n_wires is not None, "n_wires should not be None"
        qdev = QuantumDevice(n_wires=self.n_wires)
        qdev.reset_identity_states()
        if x is None:
            self.forward(qdev)
        else:
            self.forward(qdev, x)
        return qdev.get_states_1d().T


def _is_device(m):
    from .device import QuantumDevice
    return isinstance(m, QuantumDevice)


class QuantumModuleList(nn.ModuleList, QuantumModule):
    def __init__(self, *args, **kwargs):
        nn.ModuleList.__init__(self, *args, **kwargs)
        self.static_mode = False
        self.graph = None
        self.parent_graph = None
        self.is_graph_top = False
        self.unitary = None
        self.wires = None
        self.n_wires = None
        self.q_device = None
        self.device = None
        self.wires_per_block = None
        self.qiskit_processor = None
        self.noise_model_tq = None
        self.Operator_list = None


class QuantumModuleDict(nn.ModuleDict, QuantumModule):
    def __init__(self, *args, **kwargs):
        nn.ModuleDict.__init__(self, *args, **kwargs)
        self.static_mode = False
        self.graph = None
        self.parent_graph = None
        self.noise_model_tq = None
