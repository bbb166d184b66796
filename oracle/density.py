"""complex128 CPU oracle for density matrices (the NoiseDevice path).

TEST INFRASTRUCTURE ONLY: used by ``tests/`` as the checker, never by
``torchquantum_amd``.

Restates, from the mathematics, what torchquantum's density simulation is
meant to compute (torchquantum/device/noisedevices.py:36-105,
torchquantum/functional/gate_wrapper.py:265-321):

* layout: rho of n wires as [bsz, 2^n, 2^n], row index = ket, wire 0 the most
  significant bit of each index (devices.py:64-71);
* a gate: rho -> U rho U^dagger with U the full 2^n x 2^n operator built by
  brute force (:func:`oracle.statevector.full_unitary`);
* a channel: rho -> sum_k K_k rho K_k^dagger (K_k embedded the same way).

Pinned in tests/test_density.py: pure-state evolution equals |psi><psi| from
the statevector oracle; channels preserve the trace; amplitude damping,
phase damping and depolarizing match their textbook closed forms.
"""
from __future__ import annotations

from typing import Iterable, Optional, Sequence

import numpy as np

from . import statevector as S

C = np.complex128


def zero_density(n: int, bsz: int = 1) -> np.ndarray:
    rho = np.zeros((bsz, 2 ** n, 2 ** n), dtype=C)
    rho[:, 0, 0] = 1.0
    return rho


def apply_unitary(rho: np.ndarray, mat: np.ndarray, wires: Sequence[int], n: int) -> np.ndarray:
    """U rho U^dagger; mat [d, d] shared or [bsz, d, d] per row."""
    bsz = rho.shape[0]
    out = np.empty_like(rho)
    for b in range(bsz):
        m = mat[b] if mat.ndim == 3 else mat
        U = S.full_unitary(np.asarray(m, dtype=C), wires, n)
        out[b] = U @ rho[b] @ np.conj(U.T)
    return out


def apply_kraus(rho: np.ndarray, kraus: Sequence[np.ndarray], wires: Sequence[int], n: int) -> np.ndarray:
    """sum_k K_k rho K_k^dagger."""
    Ks = [S.full_unitary(np.asarray(k, dtype=C), wires, n) for k in kraus]
    return np.stack([sum(K @ r @ np.conj(K.T) for K in Ks) for r in rho])


def run_density_ops(ops: Iterable[dict], n: int, bsz: int = 1, rho: Optional[np.ndarray] = None) -> np.ndarray:
    """ops: gate dicts ``{name, wires, params, inverse}`` (statevector oracle
    format, per-row params expanded) or channel dicts ``{kraus, wires}``."""
    r = zero_density(n, bsz) if rho is None else np.array(rho, dtype=C)
    for op in ops:
        wires = op["wires"]
        wires = [wires] if isinstance(wires, (int, np.integer)) else list(wires)
        if "kraus" in op:
            r = apply_kraus(r, op["kraus"], wires, n)
            continue
        p = op.get("params")
        is_mat = op["name"] in ("qubitunitary", "qubitunitaryfast")
        pa = None if p is None else np.asarray(p, dtype=C if is_mat else np.float64)
        if pa is not None and pa.ndim == 2 and pa.shape[0] == bsz and bsz > 1 and not is_mat:
            mat = np.stack([S.gate_matrix(op["name"], pa[b], len(wires)) for b in range(bsz)])
        else:
            mat = S.gate_matrix(op["name"], pa, len(wires))
        if op.get("inverse", False):
            mat = np.conj(np.swapaxes(mat, -1, -2))
        r = apply_unitary(r, mat, wires, n)
    return r


def expval_pauli(rho: np.ndarray, obs: str) -> np.ndarray:
    """Re Tr(rho P) per row, P the dense Kronecker Pauli string."""
    P = S.pauli_string_matrix(obs)
    return np.real(np.einsum("bij,ji->b", rho, P))
