"""QuantumDevice: the statevector container (torchquantum/device/devices.py:37-136).

Same constructor, attributes and methods as the reference.  Differences that
callers cannot observe through the API:

* gate calls append to a tape; the tape is compiled and executed by the HIP
  engine the next time the state is read (``states``, ``get_states_1d``,
  measurements, ``set_states``...).  Every flush produces a new state tensor,
  so a tensor obtained from ``qdev.states`` is never modified afterwards --
  the same aliasing behaviour as the reference's out-of-place gates;
* the |0...0> start state of a fresh (or reset) device is never materialised
  when gates follow: the first fused stage generates it in-kernel;
* ``device`` defaults to the ROCm GPU when one is present.
"""
from __future__ import annotations

from typing import List, Optional, Union

import numpy as np
import torch
import torch.nn as nn

from . import engine as E
from .autograd import run_tape
from .macro import C_DTYPE

__all__ = ["QuantumDevice", "set_engine_options", "get_engine_options"]

_global_options = E.EngineOptions()


def set_engine_options(**kw) -> E.EngineOptions:
    """Change the default engine options (mode='fused'|'pergate', stage_bits, ...)."""
    global _global_options
    d = dict(_global_options.__dict__)
    d.update(kw)
    _global_options = E.EngineOptions(**d)
    return _global_options


def get_engine_options() -> E.EngineOptions:
    return _global_options


def _default_device():
    return "cuda" if torch.cuda.is_available() else "cpu"


class QuantumDevice(nn.Module):
    def __init__(self, n_wires: int, device_name: str = "default", bsz: int = 1,
                 device: Union[torch.device, str, None] = None, record_op: bool = False):
        super().__init__()
        if n_wires < 0:
            raise ValueError("n_wires must be >= 0")
        self.n_wires = n_wires
        self.device_name = device_name
        self.bsz = bsz
        self.device = torch.device(device if device is not None else _default_device())
        self.record_op = record_op
        self.op_history: List[dict] = []
        self.engine_options: Optional[E.EngineOptions] = None
        self._tape: List[E.TapeOp] = []
        self.register_buffer("_sv", None, persistent=False)   # [bsz, 2^n] complex64 or None (= |0..0>)

    # ------------------------------------------------------------------ tape
    def _options(self) -> E.EngineOptions:
        return self.engine_options if self.engine_options is not None else _global_options

    def _sv_device(self) -> torch.device:
        return self._sv.device if self._sv is not None else self.device

    def _push(self, op: E.TapeOp):
        self._tape.append(op)

    def _zero_state(self) -> torch.Tensor:
        s = torch.zeros((self.bsz, 1 << self.n_wires), dtype=C_DTYPE, device=self._sv_device())
        s[:, 0] = 1.0
        return s

    def _flush(self, options: Optional[E.EngineOptions] = None):
        opts = options if options is not None else self._options()
        if not self._tape:
            if self._sv is None:
                self._sv = self._zero_state()
            return
        tape, self._tape = self._tape, []
        for op in tape:
            if op.params is not None and op.params._version != op.version:
                raise RuntimeError(f"parameter of gate '{op.name}' was modified in place after the gate call")
        dev = self._sv_device()
        if self._sv is None:
            dummy = torch.empty(0, dtype=C_DTYPE, device=dev)
            self._sv = run_tape(tape, dummy, self.n_wires, self.bsz, opts, init_zero=True)
        else:
            st = self._sv
            if st.dtype != C_DTYPE:
                st = st.to(C_DTYPE)
            self._sv = run_tape(tape, st.reshape(self.bsz, -1).contiguous(), self.n_wires, self.bsz, opts)

    def flush(self, options: Optional[E.EngineOptions] = None):
        """Execute pending gates now (normally implicit on any state read),
        with ``options`` instead of the device's options if given."""
        self._flush(options)

    # ---------------------------------------------------------------- states
    @property
    def states(self) -> torch.Tensor:
        self._flush()
        return self._sv.reshape([self.bsz] + [2] * self.n_wires)

    @states.setter
    def states(self, value: torch.Tensor):
        self._tape = []
        value = torch.as_tensor(value)
        bsz = value.shape[0]
        self.bsz = bsz
        v = value.reshape(bsz, -1)
        if v.shape[1] != 1 << self.n_wires:
            raise ValueError(f"state has {v.shape[1]} amplitudes per row, expected {1 << self.n_wires}")
        if v.dtype != C_DTYPE:
            v = v.to(C_DTYPE)
        self._sv = v

    @property
    def state(self) -> torch.Tensor:
        """The single |0...0> state, shape [2]*n (devices.py:64-67)."""
        s = torch.zeros(1 << self.n_wires, dtype=C_DTYPE, device=self._sv_device())
        s[0] = 1.0
        return s.reshape([2] * self.n_wires)

    def reset_op_history(self):
        self.op_history = []

    def clone_states(self, existing_states: torch.Tensor):
        self.states = existing_states.clone()

    def set_states(self, states: torch.Tensor):
        bsz = states.shape[0]
        self.states = torch.reshape(states, [bsz] + [2] * self.n_wires)

    def reset_states(self, bsz: int):
        self._tape = []
        self.bsz = bsz
        self._sv = None

    def reset_identity_states(self):
        d = 1 << self.n_wires
        self.states = torch.eye(d, device=self._sv_device(), dtype=C_DTYPE)

    def reset_all_eq_states(self, bsz: int):
        energy = np.sqrt(1 / (2 ** self.n_wires) / 2)
        s = torch.full((bsz, 1 << self.n_wires), complex(energy, energy), dtype=C_DTYPE,
                       device=self._sv_device())
        self.states = s

    def get_states_1d(self) -> torch.Tensor:
        self._flush()
        return self._sv.reshape(self.bsz, 1 << self.n_wires)

    def get_state_1d(self) -> torch.Tensor:
        return self.state.reshape(1 << self.n_wires)

    @property
    def name(self):
        return self.__class__.__name__

    def __repr__(self):
        return (f" class: {self.name} \n device name: {self.device_name} \n number of qubits: {self.n_wires} \n"
                f" batch size: {self.bsz} \n current computing device: {self._sv_device()} \n"
                f" recording op history: {self.record_op} \n"
                f" current states: {repr(self.get_states_1d().cpu().detach().numpy())}")

    def _apply(self, fn, recurse=True):
        # moving the module moves the state; pending gates stay on the tape
        out = super()._apply(fn, recurse)
        if self._sv is not None:
            self.device = self._sv.device
        else:
            probe = fn(torch.empty(0, device=self.device))
            self.device = probe.device
        return out


def _bind_methods():
    from .functional import func_name_dict, func_name_dict_collect
    from .noisedevice import NoiseDevice
    for name, fn in func_name_dict.items():
        setattr(QuantumDevice, name, fn)
    for name, fn in func_name_dict_collect.items():
        setattr(QuantumDevice, name, fn)
    # the density device takes the same gate methods (noisedevices.py:101-105)
    for d in (func_name_dict, func_name_dict_collect):
        for name, fn in d.items():
            if name not in NoiseDevice.__dict__:
                setattr(NoiseDevice, name, fn)
