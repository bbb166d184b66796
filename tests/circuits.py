"""Seeded circuit generators shared by the tests and the benchmark checks
(test infrastructure)."""
from __future__ import annotations

import numpy as np
import torch

from torchquantum_amd import gates as G

ALL_GATES = ["hadamard", "shadamard", "paulix", "pauliy", "pauliz", "s", "sdg", "t", "tdg", "sx", "sxdg", "i",
             "rx", "ry", "rz", "phaseshift", "u1", "u2", "u3", "rot", "r",
             "cnot", "cy", "cz", "chadamard", "cs", "csdg", "csx", "swap", "sswap", "iswap", "dcx", "ecr",
             "rxx", "ryy", "rzz", "rzx", "crx", "cry", "crz", "crot", "cu1", "cu2", "cu3", "cu",
             "singleexcitation", "xxminyy", "xxplusyy", "toffoli", "ccz", "cswap", "rccx", "rc3x", "c3x",
             "c3sx", "c4x", "multicnot", "multixcnot", "multirz", "qft", "qubitunitaryfast"]


def random_ops(n, n_gates, bsz, rng, names=ALL_GATES, batched_frac=0.25, inverse_frac=0.2):
    """List of op dicts {name, wires, params (np [rows,P] or matrix), inverse}."""
    ops = []
    while len(ops) < n_gates:
        nm = names[rng.integers(len(names))]
        s = G.spec(nm)
        k = s.num_wires if s.num_wires is not None else int(rng.integers(2, min(n, 5) + 1))
        if k > n or k == 0:
            continue
        w = [int(x) for x in rng.choice(n, k, replace=False)]
        p = None
        if s.num_params > 0:
            rows = bsz if (bsz > 1 and rng.random() < batched_frac) else 1
            p = rng.uniform(-np.pi, np.pi, size=(rows, s.num_params))
        elif s.num_params < 0:
            a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
            p, _ = np.linalg.qr(a)
        ops.append(dict(name=nm, wires=w, params=p, inverse=bool(rng.random() < inverse_frac)))
    return ops


def apply_ops_tq(qdev, ops, device, requires_grad=False):
    """Apply op dicts through the functional API; returns the param tensors."""
    import torchquantum_amd.functional as tqf
    ts = []
    for op in ops:
        p = op["params"]
        if p is not None:
            if np.iscomplexobj(p):
                p = torch.tensor(p, dtype=torch.complex64, device=device)
            else:
                p = torch.tensor(p, dtype=torch.float32, device=device, requires_grad=requires_grad)
                ts.append(p)
        tqf.func_name_dict[op["name"]](qdev, op["wires"], params=p, inverse=op["inverse"])
    return ts


def oracle_ops(ops, bsz, params_override=None):
    """Op dicts for oracle.statevector.run_ops (per-row params expanded)."""
    out = []
    for op in ops:
        p = op["params"]
        out.append(dict(name=op["name"], wires=op["wires"], inverse=op["inverse"],
                        params=None if p is None else (p if np.iscomplexobj(p) else
                                                       (np.broadcast_to(p, (bsz, p.shape[1])) if p.shape[0] > 1
                                                        else p[0]))))
    return out


# ---------------------------------------------------------------- benchmark circuits
def config1_ops(seed=0, n=10, layers=8):
    """8 x [U3 on every wire (angles U[0, 2pi)), CNOT ring i -> (i+1) % n]."""
    g = torch.Generator().manual_seed(seed)
    ops = []
    for _ in range(layers):
        for w in range(n):
            ops.append(dict(name="u3", wires=[w], params=(torch.rand(1, 3, generator=g) * 2 * np.pi).numpy()
                            .astype(np.float64), inverse=False))
        for w in range(n):
            ops.append(dict(name="cnot", wires=[w, (w + 1) % n], params=None, inverse=False))
    return ops


def config2_ops(seed=0, n=20, layers=20):
    """20 x [RY(U[0, 2pi)) on every wire, CZ on even pairs then odd pairs]."""
    g = torch.Generator().manual_seed(seed)
    ops = []
    for _ in range(layers):
        for w in range(n):
            ops.append(dict(name="ry", wires=[w], params=(torch.rand(1, 1, generator=g) * 2 * np.pi).numpy()
                            .astype(np.float64), inverse=False))
        for start in (0, 1):
            for i in range(start, n - 1, 2):
                ops.append(dict(name="cz", wires=[i, i + 1], params=None, inverse=False))
    return ops


def config3_ops(seed=0, n=24, layers=10):
    """10 x [RX, RY, RZ on every wire (angles N(0, 0.1)), CNOT ring] (RXYZCXLayer0)."""
    g = torch.Generator().manual_seed(seed)
    ops = []
    for _ in range(layers):
        for name in ("rx", "ry", "rz"):
            for w in range(n):
                ops.append(dict(name=name, wires=[w], params=(torch.randn(1, 1, generator=g) * 0.1).numpy()
                                .astype(np.float64), inverse=False))
        for w in range(n):
            ops.append(dict(name="cnot", wires=[w, (w + 1) % n], params=None, inverse=False))
    return ops


def config3_observables(n):
    """loss = sum_i <Z_i Z_{i+1}> (ring) + 0.5 sum_i <X_i>."""
    obs, coef = [], []
    for i in range(n):
        s = ["I"] * n
        s[i] = "Z"
        s[(i + 1) % n] = "Z"
        obs.append("".join(s))
        coef.append(1.0)
    for i in range(n):
        s = ["I"] * n
        s[i] = "X"
        obs.append("".join(s))
        coef.append(0.5)
    return obs, coef


def config5_ops(seed=0, n=30, layers=20):
    """20 x [RY(U[0, 2pi)) on every wire, CZ even then odd pairs]; after every
    5th layer one Toffoli or CSWAP on three seeded distinct wires."""
    g = torch.Generator().manual_seed(seed)
    rng = np.random.default_rng(seed)
    ops = []
    for layer in range(layers):
        for w in range(n):
            ops.append(dict(name="ry", wires=[w], params=(torch.rand(1, 1, generator=g) * 2 * np.pi).numpy()
                            .astype(np.float64), inverse=False))
        for start in (0, 1):
            for i in range(start, n - 1, 2):
                ops.append(dict(name="cz", wires=[i, i + 1], params=None, inverse=False))
        if (layer + 1) % 5 == 0:
            w = [int(x) for x in rng.choice(n, 3, replace=False)]
            ops.append(dict(name="toffoli" if rng.random() < 0.5 else "cswap", wires=w, params=None,
                            inverse=False))
    return ops


def chain_zz(n):
    """sum_i Z_i Z_{i+1}, i = 0..n-2 (open chain)."""
    out = []
    for i in range(n - 1):
        s = ["I"] * n
        s[i] = s[i + 1] = "Z"
        out.append("".join(s))
    return out
