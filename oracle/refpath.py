"""Restatement of the reference's per-gate torch path (TEST INFRASTRUCTURE ONLY).

Used as the timed CPU baseline (``bench.py`` cpu_baseline leg, kind "port")
and as a second parity reference in tests.  It follows
torchquantum/functional/gate_wrapper.py:96-140 (``apply_unitary_bmm``): move
the gate's wire dims next to the batch dim, reshape to [B, 2^k, rest], batched
complex64 GEMM with the (expanded) gate matrix, view + permute back -- written
from reading that code, with the reference's complex64 matrices
(torchquantum/functional/*.py closed forms, here taken from the engine's gate
table, which tests pin to the oracle).  Nothing here ships in the product.
"""
from __future__ import annotations

import numpy as np
import torch


def apply_unitary_bmm(state: torch.Tensor, mat: torch.Tensor, wires) -> torch.Tensor:
    mat = mat.to(torch.complex64)
    dims = [w + 1 for w in wires]
    perm = list(range(state.dim()))
    for d in sorted(dims, reverse=True):
        del perm[d]
    perm = perm[:1] + dims + perm[1:]
    back = list(np.argsort(perm))
    shape = state.shape
    permuted = state.permute(perm).reshape([shape[0], mat.shape[-1], -1])
    if mat.dim() > 2:
        out = mat.bmm(permuted)
    else:
        out = mat.expand([permuted.shape[0]] + list(mat.shape)).bmm(permuted)
    return out.view(shape).permute(back)


def run(ops, n: int, bsz: int, device="cpu") -> torch.Tensor:
    """Apply op dicts {name, wires, params, inverse} from |0..0>; [bsz, 2^n]."""
    from torchquantum_amd import gates as G
    st = torch.zeros([bsz, 1 << n], dtype=torch.complex64, device=device)
    st[:, 0] = 1
    st = st.reshape([bsz] + [2] * n)
    for op in ops:
        k = len(op["wires"])
        p = op.get("params")
        pt = None if p is None else torch.as_tensor(np.asarray(p, dtype=np.float32)).reshape(-1, G.spec(op["name"]).num_params)
        m = G.build_matrices(op["name"], pt, k, dtype=torch.float32).to(device)
        m = m[0] if m.shape[0] == 1 else m
        if op.get("inverse"):
            m = m.conj().transpose(-1, -2)
        st = apply_unitary_bmm(st, m, op["wires"])
    return st.reshape(bsz, -1)


# ---------------------------------------------------------------------------
# differentiable restatement: torch autograd through the reference's own path
# ---------------------------------------------------------------------------
def _stack2(a, b, c, d):
    """[[a, b], [c, d]] from [rows, 1] blocks -> [rows, 2, 2]."""
    return torch.stack([torch.cat([a, b], dim=-1), torch.cat([c, d], dim=-1)], dim=-2)


def ref_matrix(name: str, params, cdtype=torch.complex64) -> torch.Tensor:
    """Gate matrices as the reference builds them, differentiable in torch.

    The angle is cast to the complex dtype *before* cos/sin/exp, as the
    reference does (rx.py:29): rx (rx.py:19-46), ry (ry.py:50-67), rz
    (rz.py:150-171), u3 (u3.py:52-77); fixed: hadamard (hadamard.py:12-15),
    cnot (paulix.py:93-95), cz (pauliz.py:20-22).  ``params`` is [rows, P]."""
    if name in ("rx", "ry", "rz", "u3"):
        t = params[:, 0:1].to(cdtype)
        c, s = torch.cos(t / 2), torch.sin(t / 2)
        if name == "rx":
            return _stack2(c, -1j * s, -1j * s, c)
        if name == "ry":
            return _stack2(c, -s, s, c)
        if name == "rz":
            e = torch.exp(-0.5j * t)
            z = torch.zeros_like(e)
            return _stack2(e, z, z, torch.exp(0.5j * t))
        phi, lam = params[:, 1:2].to(cdtype), params[:, 2:3].to(cdtype)
        return _stack2(c, -torch.exp(1j * lam) * s, torch.exp(1j * phi) * s, torch.exp(1j * (phi + lam)) * c)
    fixed = {
        "hadamard": [[1, 1], [1, -1]],
        "cnot": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
        "cz": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]],
    }
    if name not in fixed:
        raise NotImplementedError(f"ref_matrix: {name}")
    m = torch.tensor(fixed[name], dtype=cdtype)
    return m / np.sqrt(2) if name == "hadamard" else m


def apply_bmm_any(state: torch.Tensor, mat: torch.Tensor, wires) -> torch.Tensor:
    """``apply_unitary_bmm`` without the complex64 cast, so the same path can
    also run in complex128 (gate_wrapper.py:96-140)."""
    dims = [w + 1 for w in wires]
    perm = list(range(state.dim()))
    for d in sorted(dims, reverse=True):
        del perm[d]
    perm = perm[:1] + dims + perm[1:]
    back = list(np.argsort(perm))
    shape = state.shape
    permuted = state.permute(perm).reshape([shape[0], mat.shape[-1], -1])
    if mat.dim() > 2 and mat.shape[0] > 1:
        out = mat.bmm(permuted)
    else:
        m = mat[0] if mat.dim() > 2 else mat
        out = m.expand([permuted.shape[0]] + list(m.shape)).bmm(permuted)
    return out.view(shape).permute(back)


_PAULI = {"X": [[0, 1], [1, 0]], "Y": [[0, -1j], [1j, 0]], "Z": [[1, 0], [0, -1]]}


def expval_pauli(state: torch.Tensor, obs: str) -> torch.Tensor:
    """Re <psi|P|psi> per row for a Pauli string (char k <-> wire k), the value
    ``expval_joint_analytical`` returns (measurements.py:232-279), computed by
    applying P wire by wire on the [bsz] + [2]*n state instead of a dense kron."""
    p = state
    for w, ch in enumerate(obs):
        if ch != "I":
            p = apply_bmm_any(p, torch.tensor(_PAULI[ch], dtype=state.dtype), [w])
    b = state.shape[0]
    return (state.reshape(b, -1).conj() * p.reshape(b, -1)).sum(-1).real


def run_autograd(ops, params, n: int, bsz: int, cdtype=torch.complex64) -> torch.Tensor:
    """Apply op dicts from |0..0> with torch autograd: ``params[i]`` is the
    [rows, P] tensor of ops[i] (None for fixed gates).  Returns [bsz] + [2]*n."""
    st = torch.zeros([bsz, 1 << n], dtype=cdtype)
    st[:, 0] = 1
    st = st.reshape([bsz] + [2] * n)
    for op, p in zip(ops, params):
        m = ref_matrix(op["name"], p, cdtype)
        if op.get("inverse"):
            m = m.conj().transpose(-1, -2)
        st = apply_bmm_any(st, m, op["wires"])
    return st
