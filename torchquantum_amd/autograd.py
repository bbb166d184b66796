"""torch.autograd.Functions over the HIP kernels.

* ``CircuitFunction``: a recorded run of gates.  Forward is one execution of
  the plan; backward is the adjoint method (one reverse sweep over psi and
  lam = dL/dpsi with O(1) extra states), replacing the per-gate saved
  tensors of autograd through the reference's bmm chain
  (torchquantum/functional/gate_wrapper.py:127-136).
* ``ExpvalZAll`` / ``ExpvalPauli``: <Z_w> for all wires and Pauli-string
  expectations (torchquantum/measurement/measurements.py:232-312).  Their
  backward is lam = 2 * sum_t g_t P_t psi, torch's gradient convention for a
  real function of a complex tensor (dL/dRe + i dL/dIm).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import engine as E


class _Payload:
    __slots__ = ("c", "tape", "init_zero", "param_idx")

    def __init__(self, c, tape, init_zero, param_idx):
        self.c, self.tape, self.init_zero, self.param_idx = c, tape, init_zero, param_idx


class CircuitFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, payload: _Payload, state_in: torch.Tensor, *params: torch.Tensor):
        c = payload.c
        want_grad = any(ctx.needs_input_grad[2:])
        pool = c.pool(payload.tape, with_grad=want_grad, with_adj=True)
        out = E.forward(c, None if payload.init_zero else state_in, pool, init_zero=payload.init_zero)
        ctx.payload = payload
        ctx.pool = pool
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, gout: torch.Tensor):
        (out,) = ctx.saved_tensors
        payload = ctx.payload
        c = payload.c
        lam_in = gout.to(torch.complex64).reshape(out.shape).contiguous()
        grad, lam = E.adjoint(c, out, lam_in, ctx.pool, need_input_grad=ctx.needs_input_grad[1])
        gsum = None            # batch-row sums for parameters shared by all rows, one reduction
        grads: List = []
        for j, ti in enumerate(payload.param_idx):
            if not ctx.needs_input_grad[2 + j]:
                grads.append(None)
                continue
            op = payload.tape[ti]
            off = c.param_offsets[ti]
            P = op.n_params()
            if op.params.shape[0] == 1:
                if gsum is None:
                    gsum = grad.sum(0, keepdim=True)
                g = gsum[:, off:off + P]
            else:
                g = grad[:, off:off + P]
            grads.append(g.reshape(op.params.shape).to(op.params.dtype))
        gstate = lam if ctx.needs_input_grad[1] else None
        return (None, gstate, *grads)


def run_tape(tape: Sequence[E.TapeOp], state_in: torch.Tensor, n: int, bsz: int, opts: E.EngineOptions,
             init_zero: bool = False) -> torch.Tensor:
    """Execute a tape on ``state_in`` ([bsz, 2^n] complex64); differentiable."""
    device = state_in.device
    N.require_gpu(state_in, "the quantum state")
    c = E.compile_tape(tape, n, bsz, opts, device)
    need = torch.is_grad_enabled() and (state_in.requires_grad or c.any_param_grad)
    if need and any(op.nonunitary for op in tape):
        raise NotImplementedError("gradients through noise channels are not supported (the adjoint sweep "
                                  "un-computes states with U^dagger, which a channel does not have)")
    if need:
        params = [tape[i].params for i in c.param_idx]
        payload = _Payload(c, list(tape), init_zero, c.param_idx)
        return CircuitFunction.apply(payload, state_in, *params)
    with torch.no_grad():
        pool = c.pool(tape, with_grad=False, with_adj=False)
        return E.forward(c, None if init_zero else state_in, pool, init_zero=init_zero)


# ----------------------------------------------------------------------------
# expectations
# ----------------------------------------------------------------------------
def _rows(state: torch.Tensor, bsz: int) -> torch.Tensor:
    """[bsz, 2^n] contiguous and 16-byte aligned (the kernels read float4)."""
    st = state.reshape(bsz, -1).contiguous()
    return st if st.data_ptr() % 16 == 0 else st.clone()


def _workspace(state: torch.Tensor, bsz: int, n: int, nterms: int) -> torch.Tensor:
    nbytes = int(N.lib().tqsv_expval_workspace_bytes(bsz, n, nterms))
    if nbytes < 0:
        raise ValueError(f"expectation kernels support 1 <= n_wires <= 32 (got {n})")
    return torch.empty(max(nbytes // 4, 1), dtype=torch.float32, device=state.device)


def _pauli_apply(state: torch.Tensor, n: int, xs: Sequence[int], zs: Sequence[int], coef: torch.Tensor) -> torch.Tensor:
    bsz = state.shape[0]
    out = torch.empty_like(state)
    coef = coef.to(torch.complex64).contiguous()
    N.call("tqsv_pauli_apply", N.stream_of(state), out.data_ptr(), state.data_ptr(), bsz, n, len(xs),
           N.u64_array(xs), N.u64_array(zs), coef.data_ptr(), 0)
    return out


class ExpvalZAll(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state: torch.Tensor, n: int):
        bsz = state.shape[0]
        st = _rows(state, bsz)
        out = torch.empty((bsz, n), dtype=torch.float32, device=state.device)
        ws = _workspace(st, bsz, n, n)
        N.call("tqsv_expval_z_all", N.stream_of(st), st.data_ptr(), bsz, n, N.fptr(out), ws.data_ptr(),
               ws.numel() * 4)
        ctx.save_for_backward(st)
        ctx.n = n
        ctx.shape = state.shape
        return out

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        (st,) = ctx.saved_tensors
        n = ctx.n
        zs = [1 << (n - 1 - w) for w in range(n)]
        coef = (2.0 * g.to(torch.float32)).to(torch.complex64)
        lam = _pauli_apply(st, n, [0] * n, zs, coef)
        return lam.reshape(ctx.shape), None


def pauli_masks(obs: str, n: int) -> Tuple[int, int, int]:
    """(x mask, z mask, number of Y) of a Pauli string; char k <-> wire k."""
    obs = obs.upper()
    if len(obs) != n:
        raise ValueError(f"observable '{obs}' has {len(obs)} characters for {n} wires")
    x = z = ny = 0
    for w, ch in enumerate(obs):
        b = 1 << (n - 1 - w)
        if ch not in "IXYZ":
            raise ValueError(f"invalid Pauli character '{ch}' in '{obs}'")
        if ch in "XY":
            x |= b
        if ch in "ZY":
            z |= b
        if ch == "Y":
            ny += 1
    return x, z, ny


def _expval_terms(st: torch.Tensor, n: int, xs: tuple, zs: tuple) -> torch.Tensor:
    bsz, T = st.shape[0], len(xs)
    out = torch.empty((bsz, T), dtype=torch.float32, device=st.device)
    ws = _workspace(st, bsz, n, T)
    N.call("tqsv_expval_pauli", N.stream_of(st), st.data_ptr(), bsz, n, T, N.u64_array(xs), N.u64_array(zs),
           N.fptr(out), T, ws.data_ptr(), ws.numel() * 4)
    return out


class ExpvalPauli(torch.autograd.Function):
    """Re<psi|P_t|psi> for a list of Pauli strings -> [bsz, T]."""

    @staticmethod
    def forward(ctx, state: torch.Tensor, n: int, xs: tuple, zs: tuple, nys: tuple):
        bsz = state.shape[0]
        st = _rows(state, bsz)
        out = _expval_terms(st, n, xs, zs)
        ctx.save_for_backward(st)
        ctx.meta = (n, xs, zs, nys, state.shape)
        return out

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        (st,) = ctx.saved_tensors
        n, xs, zs, nys, shape = ctx.meta
        phase = torch.tensor([(1j) ** (k % 4) for k in nys], dtype=torch.complex64, device=st.device)
        coef = 2.0 * g.to(torch.complex64) * phase[None, :]
        lam = _pauli_apply(st, n, xs, zs, coef)
        return lam.reshape(shape), None, None, None, None


class ExpvalPauliSum(torch.autograd.Function):
    """sum_t c_t Re<psi|P_t|psi> per row -> [bsz] (c: real weights).

    When the state needs a gradient, the forward runs the passes that build
    W = 2 sum_t c_t i^{nY_t} X^x_t Z^z_t psi and reads each row's value off
    Re<psi|W> / 2 (tqsv_pauli_sum_grad); the backward is then g_row W_row,
    one elementwise pass, instead of the expectation passes plus the
    backward's own passes over the state.  Without a gradient the read-only
    expectation kernels run."""

    @staticmethod
    def forward(ctx, state: torch.Tensor, n: int, xs: tuple, zs: tuple, nys: tuple, coef: tuple):
        bsz = state.shape[0]
        st = _rows(state, bsz)
        if not ctx.needs_input_grad[0]:
            c = torch.tensor(coef, dtype=torch.float32, device=st.device)
            return _expval_terms(st, n, xs, zs) @ c
        T = len(xs)
        cw = torch.tensor([2.0 * float(c) * (1j) ** (k % 4) for c, k in zip(coef, nys)], dtype=torch.complex64,
                          device=st.device)
        w = torch.empty_like(st)
        dot = torch.empty(bsz, dtype=torch.float32, device=st.device)
        nws = int(N.lib().tqsv_pauli_sum_grad_workspace_floats(bsz, n))
        if nws < 0:
            raise ValueError(f"Pauli sums support 1 <= n_wires <= 32 (got {n})")
        ws = torch.empty(max(nws, 1), dtype=torch.float32, device=st.device)
        N.call("tqsv_pauli_sum_grad", N.stream_of(st), w.data_ptr(), st.data_ptr(), bsz, n, T, N.u64_array(xs),
               N.u64_array(zs), cw.data_ptr(), N.fptr(dot), N.fptr(ws), ws.numel())
        ctx.w = w
        ctx.shape = state.shape
        return dot * 0.5

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        lam = ctx.w * g.to(torch.float32)[:, None]
        return lam.reshape(ctx.shape), None, None, None, None, None


def expval_pauli_sum_state(state: torch.Tensor, n: int, observables: Sequence[str],
                           coeffs: Sequence[float]) -> torch.Tensor:
    """sum_t c_t <P_t> -> [bsz] for a [bsz, 2^n] state."""
    if len(observables) != len(coeffs):
        raise ValueError(f"{len(observables)} observables but {len(coeffs)} coefficients")
    masks = [pauli_masks(o, n) for o in observables]
    return ExpvalPauliSum.apply(state, n, tuple(m[0] for m in masks), tuple(m[1] for m in masks),
                                tuple(m[2] for m in masks), tuple(float(c) for c in coeffs))


def expval_pauli_strings(state: torch.Tensor, n: int, observables: Sequence[str]) -> torch.Tensor:
    masks = [pauli_masks(o, n) for o in observables]
    xs = tuple(m[0] for m in masks)
    zs = tuple(m[1] for m in masks)
    nys = tuple(m[2] for m in masks)
    return ExpvalPauli.apply(state, n, xs, zs, nys)
