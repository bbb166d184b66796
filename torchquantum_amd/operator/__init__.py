"""Operator classes (torchquantum/operator/op_types.py:53-443 and
torchquantum/operator/standard_gates/): parameters are ``nn.Parameter`` of
shape ``[1, num_params]`` initialised U(-pi, pi) (op_types.py:377-405);
``forward(q_device, wires=None, params=None, inverse=None)`` calls the
functional gate, which records it on the engine tape."""
from __future__ import annotations

from abc import ABCMeta
from typing import Iterable

import numpy as np
import torch
import torch.nn as nn

from .. import functional as tqf
from .. import gates as G
from ..macro import C_DTYPE, F_DTYPE
from ..module import QuantumModule

__all__ = ["Operator", "Operation", "Observable", "DiagonalOperation", "op_name_dict", "fixed_ops",
           "parameterized_ops"]


class Operator(QuantumModule):
    fixed_ops = []
    parameterized_ops = []
    num_params = 0
    num_wires = 1
    op_name = None
    func = None

    @property
    def name(self):
        return self._name

    @name.setter
    def name(self, value):
        self._name = value

    def __init__(self, has_params=False, trainable=False, init_params=None, n_wires=None, wires=None,
                 inverse=False):
        super().__init__()
        self.params = None
        self.n_wires = n_wires
        self.wires = wires
        self._name = self.__class__.__name__
        self.static_matrix = None
        self.inverse = inverse
        self.clifford_quantization = False
        if trainable and not has_params:
            has_params = True
        self.has_params = has_params
        self.trainable = trainable
        if self.has_params:
            self.params = self.build_params(trainable=self.trainable)
            self.reset_params(init_params)

    @classmethod
    def _matrix(cls, params):
        raise NotImplementedError

    @property
    def matrix(self):
        return self._matrix(self.params)

    @classmethod
    def _eigvals(cls, params):
        return None

    @property
    def eigvals(self):
        return self._eigvals(self.params)

    def _get_unitary_matrix(self):
        return self.matrix

    def set_wires(self, wires):
        self.wires = [wires] if isinstance(wires, int) else wires

    def build_params(self, trainable):
        p = nn.Parameter(torch.empty([1, max(self.num_params, 0)], dtype=F_DTYPE))
        p.requires_grad = bool(trainable)
        return p

    def reset_params(self, init_params=None):
        if init_params is not None:
            if isinstance(init_params, Iterable):
                for k, v in enumerate(init_params):
                    torch.nn.init.constant_(self.params[:, k], v)
            else:
                torch.nn.init.constant_(self.params, init_params)
        else:
            torch.nn.init.uniform_(self.params, -np.pi, np.pi)

    def forward(self, q_device, wires=None, params=None, inverse=None):
        """op_types.py:182-264"""
        d = self.__dict__        # plain attributes are set directly (nn.Module.__setattr__ costs microseconds)
        if inverse is not None:
            d["inverse"] = inverse
        if params is not None:
            self.params = params
        p = self.params
        if p is not None and isinstance(p, torch.Tensor) and p.dim() == 1:
            self.params = p = p.unsqueeze(-1)
        if wires is not None:
            d["wires"] = [wires] if isinstance(wires, int) else wires
        w, nw, inv = d["wires"], d["n_wires"], d["inverse"]
        if p is None:
            if nw is None:
                self.func(q_device, w, inverse=inv)
            else:
                self.func(q_device, w, n_wires=nw, inverse=inv)
        else:
            if nw is None:
                self.func(q_device, w, params=p, inverse=inv)
            else:
                self.func(q_device, w, params=p, n_wires=nw, inverse=inv)

    def __repr__(self):
        return f" class: {self.name} \n parameters: {self.params} \n wires: {self.wires} \n inverse: {self.inverse}"


class Observable(Operator, metaclass=ABCMeta):
    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.return_type = None

    def diagonalizing_gates(self):
        raise NotImplementedError


class Operation(Operator, metaclass=ABCMeta):
    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        if type(self.num_wires) == int:
            self.n_wires = self.num_wires

    def init_params(self):
        raise NotImplementedError


class DiagonalOperation(Operation, metaclass=ABCMeta):
    @classmethod
    def _matrix(cls, params):
        return torch.diag(cls._eigvals(params))


def _fixed_matrix(name):
    return torch.tensor(G.FIXED_NP[name], dtype=C_DTYPE)


def _make_class(cls_name, op_name, num_params, num_wires, base=Operation, extra=None):
    fn = tqf.func_name_dict[op_name]
    attrs = {"num_params": num_params, "num_wires": num_wires, "op_name": op_name,
             "func": staticmethod(fn), "__doc__": f"Class for {cls_name} gate."}
    cname = G.canonical(op_name)
    if num_params > 0:
        mfn = tqf.mat_dict[cname]
        attrs["_matrix"] = classmethod(lambda cls, params, _m=mfn: _m(params))
    elif cname in G.FIXED_NP:
        m = _fixed_matrix(cname)
        attrs["matrix"] = m
        attrs["_matrix"] = classmethod(lambda cls, params, _m=m: _m)
    if extra:
        attrs.update(extra)
    return type(cls_name, (base,), attrs)


# observables ---------------------------------------------------------------
def _obs(cls_name, op_name, eig, diag_gates):
    m = _fixed_matrix(G.canonical(op_name))
    return type(cls_name, (Observable,), {
        "num_params": 0, "num_wires": 1, "op_name": op_name, "func": staticmethod(tqf.func_name_dict[op_name]),
        "eigvals": torch.tensor(eig, dtype=C_DTYPE), "matrix": m,
        "_matrix": classmethod(lambda cls, params, _m=m: _m),
        "_eigvals": classmethod(lambda cls, params: cls.eigvals),
        "diagonalizing_gates": diag_gates,
        "__init__": Observable.__init__,
    })


PauliX = _obs("PauliX", "paulix", [1, -1], lambda self: [Hadamard()])          # standard_gates/paulix.py:28-29
PauliY = _obs("PauliY", "pauliy", [1, -1], lambda self: [PauliZ(), S(), Hadamard()])  # pauliy.py:28-29
PauliZ = _obs("PauliZ", "pauliz", [1, -1], lambda self: [])                     # pauliz.py:28-29
I = _obs("I", "i", [1, 1], lambda self: [])
Hadamard = _obs("Hadamard", "hadamard", [1, -1],
                lambda self: [RY(has_params=True, trainable=False, init_params=-np.pi / 4)])  # hadamard.py:30-31

# gates -----------------------------------------------------------------------
_TABLE = [
    # class, op_name, params, wires
    ("SHadamard", "shadamard", 0, 1), ("CHadamard", "chadamard", 0, 2),
    ("S", "s", 0, 1), ("SDG", "sdg", 0, 1), ("T", "t", 0, 1), ("TDG", "tdg", 0, 1),
    ("SX", "sx", 0, 1), ("SXDG", "sxdg", 0, 1), ("CNOT", "cnot", 0, 2), ("CY", "cy", 0, 2),
    ("CZ", "cz", 0, 2), ("SWAP", "swap", 0, 2), ("SSWAP", "sswap", 0, 2), ("ISWAP", "iswap", 0, 2),
    ("CSWAP", "cswap", 0, 3), ("Toffoli", "toffoli", 0, 3), ("CCZ", "ccz", 0, 3), ("CS", "cs", 0, 2),
    ("CSDG", "csdg", 0, 2), ("CSX", "csx", 0, 2), ("DCX", "dcx", 0, 2), ("ECR", "ecr", 0, 2),
    ("RCCX", "rccx", 0, 3), ("RC3X", "rc3x", 0, 4), ("C3X", "c3x", 0, 4), ("C4X", "c4x", 0, 5),
    ("C3SX", "c3sx", 0, 4),
    ("RX", "rx", 1, 1), ("RY", "ry", 1, 1), ("RZ", "rz", 1, 1), ("PhaseShift", "phaseshift", 1, 1),
    ("U1", "u1", 1, 1), ("U2", "u2", 2, 1), ("U3", "u3", 3, 1), ("Rot", "rot", 3, 1), ("R", "r", 2, 1),
    ("RXX", "rxx", 1, 2), ("RYY", "ryy", 1, 2), ("RZZ", "rzz", 1, 2), ("RZX", "rzx", 1, 2),
    ("CRX", "crx", 1, 2), ("CRY", "cry", 1, 2), ("CRZ", "crz", 1, 2), ("CRot", "crot", 3, 2),
    ("CU1", "cu1", 1, 2), ("CU2", "cu2", 2, 2), ("CU3", "cu3", 3, 2), ("CU", "cu", 4, 2),
    ("SingleExcitation", "singleexcitation", 1, 2), ("XXMINYY", "xxminyy", 2, 2),
    ("XXPLUSYY", "xxplusyy", 2, 2), ("GlobalPhase", "globalphase", 1, 0),
]
for _c, _o, _p, _w in _TABLE:
    globals()[_c] = _make_class(_c, _o, _p, _w)

EchoedCrossResonance = ECR
H = Hadamard
SH = SHadamard
CH = CHadamard
CCX = Toffoli
U = U3


class MultiRZ(Operation):
    num_params = 1
    num_wires = -1
    op_name = "multirz"
    func = staticmethod(tqf.multirz)

    @classmethod
    def _matrix(cls, params, n_wires=1):
        return tqf.mat_dict["multirz"](params, n_wires)

    @property
    def matrix(self):
        return self._matrix(self.params, self.n_wires)


class MultiCNOT(Operation):
    num_params = 0
    num_wires = -1
    op_name = "multicnot"
    func = staticmethod(tqf.multicnot)

    @property
    def matrix(self):
        return tqf.mat_dict["multicnot"](self.n_wires)


class MultiXCNOT(Operation):
    num_params = 0
    num_wires = -1
    op_name = "multixcnot"
    func = staticmethod(tqf.multixcnot)

    @property
    def matrix(self):
        return tqf.mat_dict["multixcnot"](self.n_wires)


class QFT(Operation):
    num_params = 0
    num_wires = -1
    op_name = "qft"
    func = staticmethod(tqf.qft)

    @property
    def matrix(self):
        return tqf.mat_dict["qft"](self.n_wires)


class QubitUnitary(Operation):
    """Arbitrary unitary given as the parameter (standard_gates/qubit_unitary.py)."""
    num_params = -1
    num_wires = -1
    op_name = "qubitunitary"
    func = staticmethod(tqf.qubitunitary)

    def build_params(self, trainable):
        return None

    def reset_params(self, init_params=None):
        self.params = torch.as_tensor(np.asarray(init_params), dtype=C_DTYPE)
        if self.params.dim() == 2:
            self.params = self.params.unsqueeze(0)
        if self.n_wires is None:
            self.n_wires = int(np.log2(self.params.shape[-1]))

    @classmethod
    def _matrix(cls, params):
        return params.squeeze(0)

    @property
    def matrix(self):
        return self.params.squeeze(0)

    @classmethod
    def from_controlled_operation(cls, op, c_wires, t_wires, trainable=False):
        """Controlled version of ``op`` (standard_gates/qubit_unitary.py)."""
        c = np.array(c_wires)
        t = np.array(t_wires)
        wires = [int(c)] if c.ndim == 0 else [int(x) for x in c]
        u0 = op.matrix.to(C_DTYPE)
        k_u = int(np.log2(u0.shape[-1]))
        tw = [int(t)] if t.ndim == 0 else [int(x) for x in t.reshape(-1)]
        n_sets = len(tw) // k_u
        cu = u0
        for _ in range(n_sets - 1):
            cu = torch.kron(cu, u0)
        nw = len(wires) + len(tw)
        d = 1 << nw
        u = torch.eye(d, dtype=C_DTYPE)
        u[d - cu.shape[-1]:, d - cu.shape[-1]:] = cu.detach()
        return cls(has_params=True, trainable=trainable, init_params=u.numpy(), n_wires=nw, wires=wires + tw)


class QubitUnitaryFast(QubitUnitary):
    op_name = "qubitunitaryfast"
    func = staticmethod(tqf.qubitunitaryfast)


class TrainableUnitary(QubitUnitaryFast):
    pass


class TrainableUnitaryStrict(QubitUnitaryFast):
    op_name = "qubitunitarystrict"
    func = staticmethod(tqf.qubitunitarystrict)


class Reset(Operator):
    num_params = 0
    num_wires = -1
    op_name = "reset"
    func = staticmethod(tqf.reset)


_ALL = [PauliX, PauliY, PauliZ, I, Hadamard] + [globals()[c] for c, *_ in _TABLE] + [
    MultiRZ, MultiCNOT, MultiXCNOT, QFT, QubitUnitary, QubitUnitaryFast, TrainableUnitary, TrainableUnitaryStrict,
    Reset]
op_name_dict = {c.op_name: c for c in _ALL}
op_name_dict.update({"h": H, "sh": SH, "u": U, "x": PauliX, "y": PauliY, "z": PauliZ, "cx": CNOT, "xx": RXX,
                     "yy": RYY, "zz": RZZ, "zx": RZX, "ccx": Toffoli, "p": U1, "cp": CU1, "cr": CU1,
                     "qubitunitary": QubitUnitary, "qubitunitaryfast": QubitUnitaryFast})
fixed_ops = [c.__name__ for c in _ALL if c.num_params == 0]
parameterized_ops = [c.__name__ for c in _ALL if c.num_params > 0]
__all__ += [c.__name__ for c in _ALL] + ["EchoedCrossResonance", "H", "SH", "CH", "CCX", "U"]
