"""Expectation values (torchquantum/measurement/measurements.py:232-343) on the
HIP expectation kernels.

* ``expval(qdev, wires, observables)``: applies each observable's
  diagonalizing gates to ``qdev`` in place, exactly as the reference does
  (measurements.py:293-295), then reads <Z> of the requested wires from one
  all-wire <Z> pass;
* ``MeasureAll(obs)``: ``expval`` over every wire;
* ``expval_joint_analytical(qdev, "XZ..")``: one Pauli string -> [bsz]
  (the reference builds a dense 2^n x 2^n Kronecker Hamiltonian; here the
  string is applied through bit masks in one pass);
* ``expval_joint_analytical_batch`` / ``expval_pauli_sum``: many strings
  (and a weighted sum) in one autograd node.
"""
from __future__ import annotations

from typing import List, Sequence, Union

import numpy as np
import torch

from . import operator as op
from .autograd import ExpvalZAll, expval_pauli_strings
from .module import QuantumModule

__all__ = ["expval", "MeasureAll", "expval_joint_analytical", "expval_joint_analytical_batch",
           "expval_pauli_sum", "MeasureMultiPauliSum"]


def _z_all(qdev) -> torch.Tensor:
    sv = qdev.get_states_1d()
    return ExpvalZAll.apply(sv, qdev.n_wires)


def expval(qdev, wires: Union[int, List[int]], observables):
    if isinstance(wires, int):
        wires = [wires]
        observables = [observables]
    for wire, obs in zip(wires, observables):
        for rotation in obs.diagonalizing_gates():
            rotation(qdev, wires=wire)
    z = _z_all(qdev)
    evs = [obs.eigvals.real.to(torch.float32).to(z.device) for obs in observables]
    norm = None
    if any(float(e[0] + e[1]) != 0.0 for e in evs):
        # marginals sum to the squared norm, which need not be 1 (set_states, cu2, reset...)
        norm = expval_pauli_strings(qdev.get_states_1d(), qdev.n_wires, ["I" * qdev.n_wires])[:, 0]
    out = []
    for wire, ev in zip(wires, evs):
        # <O> = p0 * e0 + p1 * e1 with p0 - p1 = <Z_wire>, p0 + p1 = |psi|^2
        zc = z[:, wire]
        val = 0.5 * (ev[0] - ev[1]) * zc
        if norm is not None:
            val = val + 0.5 * (ev[0] + ev[1]) * norm
        out.append(val)
    return torch.stack(out, dim=-1)


class MeasureAll(QuantumModule):
    """Expectation of ``obs`` on every wire -> [bsz, n] (measurements.py:315-343)."""

    def __init__(self, obs, v_c_reg_mapping=None):
        super().__init__()
        self.obs = obs
        self.v_c_reg_mapping = v_c_reg_mapping

    def forward(self, qdev):
        x = expval(qdev, list(range(qdev.n_wires)), [self.obs()] * qdev.n_wires)
        if self.v_c_reg_mapping is not None:
            c2v = self.v_c_reg_mapping["c2v"]
            perm = [c2v[k] for k in range(x.shape[-1]) if k in c2v.keys()]
            x = x[:, perm]
        return x

    def set_v_c_reg_mapping(self, mapping):
        self.v_c_reg_mapping = mapping


def expval_joint_analytical(qdev, observable: str) -> torch.Tensor:
    """<psi|P|psi> for one Pauli string (char k <-> wire k) -> [bsz]."""
    observable = observable.upper()
    assert len(observable) == qdev.n_wires
    sv = qdev.get_states_1d()
    return expval_pauli_strings(sv, qdev.n_wires, [observable])[:, 0]


def expval_joint_analytical_batch(qdev, observables: Sequence[str]) -> torch.Tensor:
    """Many Pauli strings in one node -> [bsz, len(observables)]."""
    sv = qdev.get_states_1d()
    return expval_pauli_strings(sv, qdev.n_wires, [o.upper() for o in observables])


def expval_pauli_sum(qdev, observables: Sequence[str], coeffs: Sequence[float]) -> torch.Tensor:
    """sum_t c_t <P_t> -> [bsz]."""
    ev = expval_joint_analytical_batch(qdev, observables)
    c = torch.as_tensor(np.asarray(coeffs, dtype=np.float32), device=ev.device)
    return ev @ c


class MeasureMultiPauliSum(QuantumModule):
    """Weighted Pauli sums (measurements.py MeasureMultiPauliSum, analytic)."""

    def __init__(self, obs_list):
        super().__init__()
        self.obs_list = obs_list

    def forward(self, qdev):
        outs = []
        for d in self.obs_list:
            wires = d["wires"]
            chars = ["I"] * qdev.n_wires
            for w, o in zip(wires, d["observables"]):
                chars[w] = o.upper()
            outs.append(expval_pauli_sum(qdev, ["".join(chars)], [d.get("coefficient", [1.0])[0]
                                                                  if isinstance(d.get("coefficient", 1.0), list)
                                                                  else d.get("coefficient", 1.0)]))
        return torch.stack(outs, dim=-1)
