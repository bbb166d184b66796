"""Record a circuit once, run it many times without per-gate Python.

Every call of a functional gate or an ``Operator`` appends one tape entry
(about a microsecond of Python each), and every flush hashes the tape to find
its compiled plan.  For a 160-gate circuit that host work alone exceeds the
GPU time of the whole circuit.  A loop whose circuit structure stays fixed
(serving, the inner loop of a training run) can record it once instead:

    rep = CircuitReplay(lambda q, x: model(q, x), n_wires=10, bsz=32,
                        example_inputs=(x0,))
    for x in batches:
        q = rep(x)                          # a QuantumDevice holding the output state
        out = tq.MeasureAll(tq.PauliZ)(q)

The recording function runs once on a fresh ``QuantumDevice``.  The replay
keeps its tape and compiled plan, and each call runs that plan again with
the current values of every tensor the recording used: trainable parameters
after an optimizer step, and ``example_inputs``, whose recorded copies are
overwritten by the call's inputs.  Gate matrices of device-resident
parameters are built on the GPU (``tqsv_build_dev_gates``), so a call makes
no host-side matrix work and no device-to-host copy.  Plans run on the
run-time compiled stage kernels by default (``jit="on"``, compiled when the
replay is built), also for states too small for the engine's default.

As with CUDA-graph capture, what the recording decided is fixed: the gate
sequence, wires, which tensors are parameters, shapes and Python control
flow.  Gradients reach the recorded parameters (adjoint method); the inputs
are copied, so no gradient flows to them.  The reference has no counterpart
(its gates always run through ``gate_wrapper``, torchquantum/functional/
gate_wrapper.py:324-446); the gates, their order and their semantics are the
recorded ones.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch

from . import _native as N
from . import engine as E
from .autograd import CircuitFunction, _Payload
from .device import QuantumDevice
from .macro import C_DTYPE

__all__ = ["CircuitReplay"]


class CircuitReplay:
    def __init__(self, record: Callable, n_wires: int, bsz: int = 1, device="cuda",
                 example_inputs: Sequence[torch.Tensor] = (), options: Optional[E.EngineOptions] = None):
        self.n_wires, self.bsz = int(n_wires), int(bsz)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise NotImplementedError("CircuitReplay runs on MI355X (ROCm) devices only; use device='cuda'")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.inputs = [x.detach().clone().to(self.device) for x in example_inputs]
        self.options = options if options is not None else E.EngineOptions(mode="fused", jit="on")
        q = QuantumDevice(self.n_wires, bsz=self.bsz, device=self.device)
        q.engine_options = self.options
        record(q, *self.inputs)
        if q._sv is not None:
            raise ValueError("the recording function read the state (a flush) before its last gate; a replay "
                             "needs the whole circuit as one recorded run")
        if not q._tape:
            raise ValueError("the recording function applied no gate")
        self.tape = list(q._tape)
        q._tape = []
        self.compiled = E.Compiled(self.tape, self.n_wires, self.bsz, self.options, self.device)
        c = self.compiled
        if c.prog_host is not None:          # compile (jit="on") or queue (jit="auto") the stage kernels now
            c.jit_kernels("f")
            if c.any_param_grad:
                c.jit_kernels("a")
                c.jit_kernels("p")
        self._payload = _Payload(c, self.tape, True, c.param_idx)
        self._dummy = torch.empty(0, dtype=C_DTYPE, device=self.device)
        self._qdev = QuantumDevice(self.n_wires, bsz=self.bsz, device=self.device)

    def __call__(self, *inputs: torch.Tensor) -> QuantumDevice:
        """Run the recorded circuit on |0..0> with the current parameter
        values and these inputs (same shapes as ``example_inputs``); returns
        a QuantumDevice holding the output state (reused by every call)."""
        if len(inputs) != len(self.inputs):
            raise ValueError(f"the replay was recorded with {len(self.inputs)} input(s), got {len(inputs)}")
        with torch.no_grad():
            for dst, x in zip(self.inputs, inputs):
                if x.shape != dst.shape:
                    raise ValueError(f"input shape {tuple(x.shape)} differs from the recorded {tuple(dst.shape)}")
                dst.copy_(x)
        c, tape = self.compiled, self.tape
        if torch.is_grad_enabled() and c.any_param_grad:
            params = [tape[i].params for i in c.param_idx]
            out = CircuitFunction.apply(self._payload, self._dummy, *params)
        else:
            with torch.no_grad():
                pool = c.pool(tape, with_grad=False, with_adj=False)
                out = E.forward(c, None, pool, init_zero=True)
        q = self._qdev
        q._tape = []
        q._sv = out
        return q
