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
