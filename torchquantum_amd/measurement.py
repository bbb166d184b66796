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
  (and a weighted sum) in one autograd node;
* ``measure`` / ``expval_joint_sampling`` / ``expval_joint_sampling_grouping``
  (measurements.py:37-229): shots drawn on the GPU by ``tqsv_sample`` from
  |psi|^2 without copying the state to the host.
"""
from __future__ import annotations

from collections import OrderedDict
from typing import Dict, List, Optional, Sequence, Union

import numpy as np
import torch

from . import _native as N
from . import operator as op
from .autograd import ExpvalZAll, expval_pauli_strings, expval_pauli_sum_state
from .macro import F_DTYPE
from .module import QuantumModule
from .noisedevice import NoiseDevice, density_expval_strings

__all__ = ["expval", "MeasureAll", "expval_joint_analytical", "expval_joint_analytical_batch",
           "expval_pauli_sum", "MeasureMultiPauliSum", "gen_bitstrings", "sample_indices", "measure",
           "find_observable_groups", "expval_joint_sampling", "expval_joint_sampling_grouping"]


def _z_all(qdev) -> torch.Tensor:
    if isinstance(qdev, NoiseDevice):     # Tr(rho Z_w)
        n = qdev.n_wires
        return density_expval_strings(qdev, ["I" * w + "Z" + "I" * (n - 1 - w) for w in range(n)])
    sv = qdev.get_states_1d()
    return ExpvalZAll.apply(sv, qdev.n_wires)


def _norm(qdev) -> torch.Tensor:
    """|psi|^2 per row, or tr(rho) for a density device."""
    if isinstance(qdev, NoiseDevice):
        return density_expval_strings(qdev, ["I" * qdev.n_wires])[:, 0]
    return expval_pauli_strings(qdev.get_states_1d(), qdev.n_wires, ["I" * qdev.n_wires])[:, 0]


# (observable eigenvalue tensors, device) -> (coefficients on the device, any half-sum != 0)
_COEF_CACHE: dict = {}


def _expval_coef(observables, device):
    """[2, len] coefficients of <Z> and of |psi|^2 per observable: <O> = p0 e0
    + p1 e1 = (e0 - e1)/2 <Z> + (e0 + e1)/2 |psi|^2.  Cached per observable
    set (class-level eigenvalues) and device."""
    key = (tuple(id(obs.eigvals) for obs in observables), str(device))
    hit = _COEF_CACHE.get(key)
    if hit is not None:
        return hit
    ev = torch.stack([obs.eigvals.real.detach().to("cpu", torch.float32).reshape(-1)[:2] for obs in observables])
    half_diff, half_sum = 0.5 * (ev[:, 0] - ev[:, 1]), 0.5 * (ev[:, 0] + ev[:, 1])
    coef = torch.stack([half_diff, half_sum]).to(device)
    out = (coef, bool((half_sum != 0).any()), bool((half_diff == 1).all() and (half_sum == 0).all()))
    if all(isinstance(getattr(type(obs), "eigvals", None), torch.Tensor) for obs in observables):
        if len(_COEF_CACHE) > 256:
            _COEF_CACHE.clear()
        _COEF_CACHE[key] = out
    return out


def expval(qdev, wires: Union[int, List[int]], observables):
    if isinstance(wires, int):
        wires = [wires]
        observables = [observables]
    for wire, obs in zip(wires, observables):
        for rotation in obs.diagonalizing_gates():
            rotation(qdev, wires=wire)
    z = _z_all(qdev)
    coef, any_sum, pure_z = _expval_coef(observables, z.device)
    all_wires = list(wires) == list(range(qdev.n_wires))
    if pure_z and all_wires:
        return z                       # PauliZ on every wire: <Z_w> itself
    zc = z if all_wires else z[:, torch.as_tensor(list(wires), dtype=torch.long).to(z.device)]
    out = zc * coef[0]
    if any_sum:
        # marginals sum to the squared norm, which need not be 1 (set_states, cu2, reset...)
        out = out + _norm(qdev)[:, None] * coef[1]
    return out


class MeasureAll(QuantumModule):
    """Expectation of ``obs`` on every wire -> [bsz, n] (measurements.py:315-343)."""

    def __init__(self, obs, v_c_reg_mapping=None):
        super().__init__()
        self.obs = obs
        self.v_c_reg_mapping = v_c_reg_mapping

    def forward(self, qdev):
        inst = self.__dict__.get("_obs_inst")
        if inst is None or type(inst) is not self.obs:
            inst = self.obs()                 # one observable instance, reused across calls
            self.__dict__["_obs_inst"] = inst
        x = expval(qdev, list(range(qdev.n_wires)), [inst] * qdev.n_wires)
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
    if isinstance(qdev, NoiseDevice):
        return density_expval_strings(qdev, [observable])[:, 0]
    sv = qdev.get_states_1d()
    return expval_pauli_strings(sv, qdev.n_wires, [observable])[:, 0]


def expval_joint_analytical_batch(qdev, observables: Sequence[str]) -> torch.Tensor:
    """Many Pauli strings in one node -> [bsz, len(observables)]."""
    if isinstance(qdev, NoiseDevice):
        return density_expval_strings(qdev, [o.upper() for o in observables])
    sv = qdev.get_states_1d()
    return expval_pauli_strings(sv, qdev.n_wires, [o.upper() for o in observables])


def expval_pauli_sum(qdev, observables: Sequence[str], coeffs: Sequence[float]) -> torch.Tensor:
    """sum_t c_t <P_t> -> [bsz]."""
    if isinstance(qdev, NoiseDevice):
        ev = expval_joint_analytical_batch(qdev, observables)
        c = torch.as_tensor(np.asarray(coeffs, dtype=np.float32), device=ev.device)
        return ev @ c
    return expval_pauli_sum_state(qdev.get_states_1d(), qdev.n_wires, [o.upper() for o in observables],
                                  [float(c) for c in coeffs])


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


# ---------------------------------------------------------------------------
# shot sampling
# ---------------------------------------------------------------------------
def gen_bitstrings(n_wires: int) -> List[str]:
    """All 2^n bitstrings, wire 0 first (measurements.py:33-34)."""
    return ["{:0{}b}".format(k, n_wires) for k in range(2 ** n_wires)]


def sample_indices(qdev, n_shots: int = 1024, seed: Optional[int] = None) -> torch.Tensor:
    """n_shots computational-basis outcomes per batch row, drawn on the GPU from
    |psi|^2 -> int64 [bsz, n_shots] of in-row indices (wire 0 = most
    significant bit).  ``seed`` None draws one from torch's default generator,
    so ``torch.manual_seed`` makes the shots reproducible."""
    if isinstance(qdev, NoiseDevice):
        # outcome probabilities rho_ii drawn through amplitudes sqrt(rho_ii)
        p = qdev.get_probs_1d().detach().real.clamp_min(0)
        sv = torch.sqrt(p).to(torch.complex64)
    else:
        sv = qdev.get_states_1d().detach()
    if not sv.is_cuda:
        raise NotImplementedError("torchquantum_amd samples on the GPU only (the state is on the CPU)")
    sv = sv.to(torch.complex64).contiguous()
    bsz, n = sv.shape[0], qdev.n_wires
    if seed is None:
        seed = int(torch.randint(0, 2 ** 62, (1,)).item())
    nbytes = int(N.lib().tqsv_sample_workspace_bytes(bsz, n))
    if nbytes < 0:
        raise ValueError(f"sampling supports n_wires <= 32 (got {n})")
    ws = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=sv.device)
    out = torch.empty((bsz, int(n_shots)), dtype=torch.int64, device=sv.device)
    N.call("tqsv_sample", N.stream_of(sv), sv.data_ptr(), bsz, n, int(n_shots), int(seed) & (2 ** 64 - 1),
           out.data_ptr(), ws.data_ptr(), ws.numel())
    return out


def measure(qdev, n_shots: int = 1024, draw_id: Optional[int] = None, seed: Optional[int] = None):
    """Bitstring counts per batch row: a list of OrderedDicts over all 2^n
    bitstrings in sorted order (measurements.py:37-74)."""
    idx = sample_indices(qdev, n_shots, seed).cpu().numpy()
    n = qdev.n_wires
    keys = gen_bitstrings(n)
    out = []
    for row in idx:
        counts = np.bincount(row, minlength=1 << n)
        out.append(OrderedDict((k, int(c)) for k, c in zip(keys, counts)))
    if draw_id is not None:
        import matplotlib.pyplot as plt
        plt.bar(out[draw_id].keys(), out[draw_id].values())
        plt.xticks(rotation="vertical")
        plt.xlabel("bitstring [qubit0, qubit1, ..., qubitN]")
        plt.title("distribution of measured bitstrings")
        plt.show()
    return out


def _rotate_to_basis(qdev, observable: str):
    """A copy of ``qdev`` with each wire's diagonalizing gates applied
    (measurements.py:203-219)."""
    from .device import QuantumDevice
    if isinstance(qdev, NoiseDevice):
        raise NotImplementedError("sampled Pauli expectations of a density device: use "
                                  "expval_joint_analytical (exact Tr(rho P))")
    pauli = {"X": op.PauliX, "Y": op.PauliY, "Z": op.PauliZ, "I": op.I}
    clone = QuantumDevice(n_wires=qdev.n_wires, bsz=qdev.bsz, device=qdev.get_states_1d().device)
    clone.engine_options = qdev.engine_options
    clone.clone_states(qdev.states.detach())
    for wire in range(qdev.n_wires):
        for rotation in pauli[observable[wire]]().diagonalizing_gates():
            rotation(clone, wires=wire)
    return clone


def _parity_expval(idx: torch.Tensor, observable: str, n: int) -> torch.Tensor:
    """mean over shots of (-1)^(number of non-I wires measured 1) -> [bsz]."""
    mask = 0
    for w, ch in enumerate(observable):
        if ch != "I":
            mask |= 1 << (n - 1 - w)
    x = idx.cpu().numpy() & mask
    par = np.zeros_like(x)
    while np.any(x):
        par ^= x & 1
        x >>= 1
    return torch.tensor((1.0 - 2.0 * par).mean(axis=1), dtype=F_DTYPE)


def expval_joint_sampling(qdev, observable: str, n_shots: int = 1024, seed: Optional[int] = None) -> torch.Tensor:
    """<P> estimated from n_shots samples in the observable's basis -> [bsz]
    (measurements.py:168-229)."""
    observable = observable.upper()
    if len(observable) != qdev.n_wires:
        raise ValueError(f"observable '{observable}' has {len(observable)} chars for {qdev.n_wires} wires")
    clone = _rotate_to_basis(qdev, observable)
    return _parity_expval(sample_indices(clone, n_shots, seed), observable, qdev.n_wires)


def find_observable_groups(observables: Sequence[str]) -> Dict[str, List[str]]:
    """Greedy grouping of Pauli strings that can share one measurement basis
    (measurements.py:77-111): an observable joins the first group whose every
    non-I character agrees with its own, and fills the group's I positions."""
    groups: Dict[str, List[str]] = {}
    for obs in observables:
        target = None
        for g in groups:
            if all(o == "I" or g[k] == "I" or g[k] == o for k, o in enumerate(obs)):
                target = g
                break
        if target is None:
            groups[obs] = [obs]
            continue
        merged = "".join(o if target[k] == "I" and o != "I" else target[k] for k, o in enumerate(obs))
        members = groups.pop(target)
        members.append(obs)
        groups[merged] = members
    return groups


def expval_joint_sampling_grouping(qdev, observables: Sequence[str], n_shots_per_group: int = 1024,
                                   seed: Optional[int] = None) -> Dict[str, torch.Tensor]:
    """Sampled <P> for many strings, one set of shots per compatible group
    (measurements.py:114-165) -> {observable: [bsz]}."""
    obs = [o.upper() for o in observables]
    if len(obs) != len(set(obs)):
        raise ValueError("each observable should be unique")
    out: Dict[str, torch.Tensor] = {}
    for gi, (group, members) in enumerate(find_observable_groups(obs).items()):
        clone = _rotate_to_basis(qdev, group)
        idx = sample_indices(clone, n_shots_per_group, None if seed is None else seed + gi)
        for m in members:
            out[m] = _parity_expval(idx, m, qdev.n_wires)
    return out
