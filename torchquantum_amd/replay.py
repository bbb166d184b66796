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
are copied, so no gradient flows to them.

Inputs reach the gates in two ways, both followed per call: as views of an
input (``x[:, k:k+1]``, ``x[:, k]``, reshapes) and as copies of input columns
(``x[:, [k]]``, the advanced indexing of ``GeneralEncoder``).  The recording
function runs a second time with other input values to tell them apart; a
gate parameter that changes with the inputs but is neither of those (an
arithmetic transform of an input) is rejected, since a replay could not
recompute it.  The reference has no counterpart
(its gates always run through ``gate_wrapper``, torchquantum/functional/
gate_wrapper.py:324-446); the gates, their order and their semantics are the
recorded ones.
"""
from __future__ import annotations

from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import engine as E
from .autograd import CircuitFunction, _Payload
from .device import QuantumDevice
from .macro import C_DTYPE

__all__ = ["CircuitReplay"]


def _shares_storage(p: torch.Tensor, xs: Sequence[torch.Tensor]) -> bool:
    ptr = p.untyped_storage().data_ptr()
    return any(x.untyped_storage().data_ptr() == ptr for x in xs)


def bind_inputs(tape_a, tape_b, inputs_a: Sequence[torch.Tensor], inputs_b: Sequence[torch.Tensor],
                bsz: int) -> List[Tuple[int, int, List[int]]]:
    """Which gate parameters of a recording come from its inputs, and how.

    tape_a / tape_b: the tapes of two recordings of the same function, made
    with inputs_a and inputs_b (other values, same shapes).  A parameter that
    is the same tensor or equal in both recordings does not depend on the
    inputs; one that shares storage with an input of its recording is a view
    and follows a copy into that input.  Every other parameter must equal,
    column by column, columns of one input (viewed as [bsz, -1]) in both
    recordings: returned as (tape index, input index, input columns).
    Raises ValueError when the two recordings differ in structure or a
    parameter is computed from the inputs in another way."""
    if len(tape_a) != len(tape_b) or any((a.name, a.wires, a.inverse) != (b.name, b.wires, b.inverse)
                                         for a, b in zip(tape_a, tape_b)):
        raise ValueError("the recorded circuit depends on the input values (gate sequence or wires changed "
                         "between two recordings); a replay needs a fixed structure")
    cache = {}

    def columns(k, dtype):
        """Input k as host columns of `dtype` in recording a, and a map from
        the bytes of each column in recording b to its index (first wins)."""
        key = (k, dtype)
        if key not in cache:
            xa, xb = inputs_a[k], inputs_b[k]
            if xa.dim() == 0 or xa.shape[0] != bsz:
                cache[key] = None
            else:
                fa = xa.detach().reshape(bsz, -1).to(dtype).cpu().numpy()
                fb = xb.detach().reshape(bsz, -1).to(dtype).cpu().numpy()
                index = {}
                for c in range(fb.shape[1]):
                    index.setdefault(np.ascontiguousarray(fb[:, c]).tobytes(), c)
                cache[key] = (fa, index)
        return cache[key]

    gathers = []
    for i, (a, b) in enumerate(zip(tape_a, tape_b)):
        pa, pb = a.params, b.params
        if pa is None or pa is pb:
            continue
        if _shares_storage(pa, inputs_a):
            if not _shares_storage(pb, inputs_b):
                raise ValueError(f"gate {a.name} on wires {a.wires}: parameters are a view of an input in one "
                                 f"recording only")
            continue
        if pa.shape == pb.shape and torch.equal(pa, pb):
            continue
        msg = (f"gate {a.name} on wires {a.wires}: its parameters change with the inputs but are neither a view "
               f"nor a copy of input columns; compute that transform before the call and pass its result as "
               f"the input")
        if pa.is_complex() or pa.dim() == 0 or pa.shape[0] != bsz or pa.shape != pb.shape:
            raise ValueError(msg)
        ca = pa.detach().reshape(bsz, -1).cpu().numpy()
        cb = pb.detach().reshape(bsz, -1).cpu().numpy()
        found = None
        for k in range(len(inputs_a)):
            cols_of = columns(k, pa.dtype)
            if cols_of is None:
                continue
            fa = cols_of[0]
            cols = []
            for j in range(ca.shape[1]):
                hit = cols_of[1].get(cb[:, j].tobytes())
                if hit is None or not np.array_equal(fa[:, hit], ca[:, j]):
                    break
                cols.append(hit)
            if len(cols) == ca.shape[1]:
                found = (i, k, cols)
                break
        if found is None:
            raise ValueError(msg)
        gathers.append(found)
    return gathers


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
        self.tape = self._record(record, self.inputs)
        self._gathers = []       # (input index, columns, column-gather buffer)
        if self.inputs:
            others = [torch.rand_like(x) if x.is_floating_point() else x.clone() for x in self.inputs]
            binds = bind_inputs(self.tape, self._record(record, others), self.inputs, others, self.bsz)
            # parameters copied from input columns become views of one buffer
            # per input, refilled by one gather per call
            for k in sorted({b[1] for b in binds}):
                mine = [b for b in binds if b[1] == k]
                cols = [c for _, _, cs in mine for c in cs]
                buf = torch.empty(self.bsz, len(cols), dtype=self.tape[mine[0][0]].params.dtype,
                                  device=self.device)
                idx = torch.tensor(cols, dtype=torch.long, device=self.device)
                off = 0
                for i, _, cs in mine:
                    p = self.tape[i].params
                    v = buf.narrow(1, off, len(cs))
                    self.tape[i].params = v if p.dim() == 2 else v.reshape(p.shape)
                    off += len(cs)
                self._gathers.append((k, idx, buf))
            self._fill_gathers()
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

    def _record(self, record: Callable, inputs: Sequence[torch.Tensor]):
        q = QuantumDevice(self.n_wires, bsz=self.bsz, device=self.device)
        q.engine_options = self.options
        record(q, *inputs)
        if q._sv is not None:
            raise ValueError("the recording function read the state (a flush) before its last gate; a replay "
                             "needs the whole circuit as one recorded run")
        if not q._tape:
            raise ValueError("the recording function applied no gate")
        tape = list(q._tape)
        q._tape = []
        return tape

    def _fill_gathers(self):
        with torch.no_grad():
            for k, idx, buf in self._gathers:
                torch.index_select(self.inputs[k].reshape(self.bsz, -1).to(buf.dtype), 1, idx, out=buf)

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
        self._fill_gathers()
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
