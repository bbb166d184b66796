"""QuantumModule (torchquantum/module/modules.py:38-383): nn.Module with the
reference's static-mode attributes, op-history replay and ``get_unitary``."""
from __future__ import annotations

from typing import Iterable

import torch
import torch.nn as nn

__all__ = ["QuantumModule", "QuantumModuleList", "QuantumModuleDict"]


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
        """Static mode (modules.py:230-261).  The engine always compiles the
        recorded gates into fused stages, so static mode only records the
        block width, used as the fused-stage hint."""
        self.wires_per_block = wires_per_block
        self.static_mode = False
        self.is_graph_top = is_graph_top

    def static_off(self):
        self.static_mode = False
        self.graph = None

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
