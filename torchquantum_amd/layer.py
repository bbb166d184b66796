"""Circuit templates that define the benchmark circuits
(torchquantum/layer/layers/layers.py, layer/entanglement/*.py)."""
from __future__ import annotations

import numpy as np
import torch

from . import operator as op
from .module import QuantumModule, QuantumModuleList, static_support

__all__ = ["Op1QAllLayer", "Op2QAllLayer", "EntanglePairwise", "LayerTemplate0", "RXYZCXLayer0",
           "RandomLayer"]


class Op1QAllLayer(QuantumModule):
    """Same single-qubit op on every wire (layers.py:71-96)."""

    def __init__(self, op, n_wires: int, has_params=False, trainable=False):
        super().__init__()
        self.n_wires = n_wires
        self.op = op
        self.ops_all = QuantumModuleList()
        for _ in range(n_wires):
            self.ops_all.append(op(has_params=has_params, trainable=trainable))

    @static_support
    def forward(self, q_device):
        for k in range(self.n_wires):
            self.ops_all[k](q_device, wires=k)


class Op2QAllLayer(QuantumModule):
    """Two-qubit op on [k, (k + jump) % n] (op2_layer.py:43-110)."""

    def __init__(self, op, n_wires: int, has_params=False, trainable=False, wire_reverse=False, jump=1,
                 circular=False):
        super().__init__()
        self.n_wires = n_wires
        self.jump = jump
        self.circular = circular
        self.op = op
        self.wire_reverse = wire_reverse
        self.ops_all = QuantumModuleList()
        n_ops = n_wires if circular else n_wires - jump
        for _ in range(n_ops):
            self.ops_all.append(op(has_params=has_params, trainable=trainable))

    @static_support
    def forward(self, q_device):
        for k in range(len(self.ops_all)):
            wires = [k, (k + self.jump) % self.n_wires]
            if self.wire_reverse:
                wires.reverse()
            self.ops_all[k](q_device, wires=wires)


class EntanglePairwise(QuantumModule):
    """Even pairs (i, i+1) then odd pairs (entanglement.py:178-231)."""

    def __init__(self, op, n_wires: int, has_params=False, trainable=False, wire_reverse=False):
        super().__init__()
        self.n_wires = n_wires
        self.op = op
        self.wire_reverse = wire_reverse
        self.ops_all = QuantumModuleList()
        for _ in range(n_wires - 1):
            self.ops_all.append(op(has_params=has_params, trainable=trainable))

    @static_support
    def forward(self, q_device):
        k = 0
        for start in (0, 1):
            for i in range(start, self.n_wires - 1, 2):
                wires = [i, i + 1]
                if self.wire_reverse:
                    wires.reverse()
                self.ops_all[k](q_device, wires=wires)
                k += 1


class LayerTemplate0(QuantumModule):
    """Base for block templates (layers.py LayerTemplate0): arch = {n_wires, n_blocks, ...}."""

    def __init__(self, arch: dict = None):
        super().__init__()
        self.n_wires = arch["n_wires"]
        self.arch = arch
        self.n_blocks = arch.get("n_blocks", None)
        self.n_layers_per_block = arch.get("n_layers_per_block", None)
        self.layers_all = self.build_layers()

    def build_layers(self):
        raise NotImplementedError

    @static_support
    def forward(self, q_device):
        for layer in self.layers_all:
            layer(q_device)


class RXYZCXLayer0(LayerTemplate0):
    """RX, RY, RZ on all wires then a CNOT ring, per block (layers.py:408-451)."""
    name = "rxyzcx_0"

    def build_layers(self):
        layers_all = QuantumModuleList()
        for _ in range(self.arch["n_blocks"]):
            layers_all.append(Op1QAllLayer(op=op.RX, n_wires=self.n_wires, has_params=True, trainable=True))
            layers_all.append(Op1QAllLayer(op=op.RY, n_wires=self.n_wires, has_params=True, trainable=True))
            layers_all.append(Op1QAllLayer(op=op.RZ, n_wires=self.n_wires, has_params=True, trainable=True))
            layers_all.append(Op2QAllLayer(op=op.CNOT, n_wires=self.n_wires, jump=1, circular=True))
        return layers_all


class RandomLayer(QuantumModule):
    """Random gates drawn from ``op_types`` with distinct random wires
    (random_layers.py:80-230; the op/wire stream of a given seed is this
    build's own, seeded through numpy's Generator)."""

    def __init__(self, wires, n_ops=None, n_params=None, op_ratios=None, op_types=(op.RX, op.RY, op.RZ, op.CNOT),
                 seed=None, qiskit_compatible=False):
        super().__init__()
        self.wires = wires if isinstance(wires, (list, tuple)) else [wires]
        self.n_wires = len(self.wires)
        self.n_ops = n_ops if n_ops is not None else 50
        self.op_types = list(op_types)
        self.op_ratios = op_ratios
        self.rng = np.random.default_rng(seed)
        self.op_list = QuantumModuleList()
        self.op_wires = []
        self.build_random_layer()

    def build_random_layer(self):
        ratios = None if self.op_ratios is None else np.asarray(self.op_ratios) / np.sum(self.op_ratios)
        while len(self.op_list) < self.n_ops:
            cls = self.op_types[self.rng.choice(len(self.op_types), p=ratios)]
            nw = cls.num_wires if isinstance(cls.num_wires, int) and cls.num_wires > 0 else 1
            if nw > self.n_wires:
                continue
            w = [self.wires[i] for i in self.rng.choice(self.n_wires, nw, replace=False)]
            o = cls(has_params=cls.num_params > 0, trainable=cls.num_params > 0)
            if cls.num_params > 0:
                with torch.no_grad():
                    o.params.uniform_(-np.pi, np.pi)
            self.op_list.append(o)
            self.op_wires.append(w)

    @static_support
    def forward(self, q_device):
        for o, w in zip(self.op_list, self.op_wires):
            o(q_device, wires=w)
