"""Density-matrix device: torchquantum's ``NoiseDevice``
(torchquantum/device/noisedevices.py:36-105, gate application
torchquantum/functional/gate_wrapper.py:265-321 and :433-440).

The density matrix rho of n wires is kept as a 2n-wire register on the same
engine as the statevectors: flat index r * 2^n + c holds rho[r, c] (the
reference's layout, ``densities`` of shape [bsz] + [2] * 2n, |r><c|).  Then

    vec(U rho U^dagger) = (U (x) conj(U)) vec(rho),

so a gate U on wires W is the gate U on register wires W (the row index)
followed by conj(U) on register wires W + n (the column index).  Every gate
kernel, the fused stages and the adjoint gradient apply unchanged; gradients
w.r.t. gate parameters flow through both copies.

A channel with Kraus operators K_k on wires W acts on vec(rho) as the
superoperator sum_k K_k (x) conj(K_k): one dense op on register wires
(W, W + n).  Channels are not unitary, so the adjoint sweep cannot pass
through them: a differentiable pass with a channel raises.

Deliberate differences from the reference (both are errors there):
* the reference computes rho U^dagger as ``rho @ conj(U)``
  (gate_wrapper.py:287-305, no transpose), which is wrong for every
  non-symmetric U (RY, U3, ...); here it is rho U^dagger;
* ``measure`` draws from |rho_ii|^2 there (measurements.py:48-58 squares the
  probabilities again); here from rho_ii.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence, Union

import numpy as np
import torch
import torch.nn as nn

from . import engine as E
from .device import QuantumDevice, _default_device

C_DTYPE = torch.complex64

__all__ = ["NoiseDevice", "superoperator", "amplitude_damping_kraus", "phase_damping_kraus",
           "depolarizing_kraus", "bit_flip_kraus", "phase_flip_kraus"]


def superoperator(kraus: Sequence[np.ndarray]) -> np.ndarray:
    """sum_k K_k (x) conj(K_k): the channel rho -> sum_k K_k rho K_k^dagger
    acting on vec(rho) (row-major: index r * d + c)."""
    ks = [np.asarray(k, dtype=np.complex128) for k in kraus]
    d = ks[0].shape[0]
    if any(k.shape != (d, d) for k in ks):
        raise ValueError("Kraus operators must be square and of one size")
    return sum(np.kron(k, np.conj(k)) for k in ks)


def amplitude_damping_kraus(gamma: float) -> List[np.ndarray]:
    g = float(gamma)
    return [np.array([[1, 0], [0, np.sqrt(1 - g)]]), np.array([[0, np.sqrt(g)], [0, 0]])]


def phase_damping_kraus(lam: float) -> List[np.ndarray]:
    l_ = float(lam)
    return [np.array([[1, 0], [0, np.sqrt(1 - l_)]]), np.array([[0, 0], [0, np.sqrt(l_)]])]


def depolarizing_kraus(p: float) -> List[np.ndarray]:
    """rho -> (1 - p) rho + p/3 (X rho X + Y rho Y + Z rho Z)."""
    p = float(p)
    x = np.array([[0, 1], [1, 0]])
    y = np.array([[0, -1j], [1j, 0]])
    z = np.array([[1, 0], [0, -1]])
    return [np.sqrt(1 - p) * np.eye(2), np.sqrt(p / 3) * x, np.sqrt(p / 3) * y, np.sqrt(p / 3) * z]


def bit_flip_kraus(p: float) -> List[np.ndarray]:
    p = float(p)
    return [np.sqrt(1 - p) * np.eye(2), np.sqrt(p) * np.array([[0, 1], [1, 0]])]


def phase_flip_kraus(p: float) -> List[np.ndarray]:
    p = float(p)
    return [np.sqrt(1 - p) * np.eye(2), np.sqrt(p) * np.array([[1, 0], [0, -1]])]


class NoiseDevice(nn.Module):
    """Density matrices of ``bsz`` rows on ``n_wires`` wires, starting in
    |0..0><0..0| (noisedevices.py:36-74).  The functional gates and the gate
    methods apply U rho U^dagger; the channel methods apply Kraus maps."""

    def __init__(self, n_wires: int, device_name: str = "noisedevice", bsz: int = 1,
                 device: Union[torch.device, str, None] = None, record_op: bool = False):
        super().__init__()
        if n_wires < 1:
            raise ValueError("n_wires must be >= 1")
        self.n_wires = n_wires
        self.device_name = device_name
        self.record_op = record_op
        self.op_history: List[dict] = []
        self.reg = QuantumDevice(2 * n_wires, bsz=bsz,
                                 device=device if device is not None else _default_device())

    # ------------------------------------------------------------ plumbing
    @property
    def bsz(self) -> int:
        return self.reg.bsz

    @property
    def device(self) -> torch.device:
        return self.reg.device

    @property
    def engine_options(self):
        return self.reg.engine_options

    @engine_options.setter
    def engine_options(self, v):
        self.reg.engine_options = v

    def _push(self, op: E.TapeOp):
        """A gate U on wires W: U on the row wires, conj(U) on the column wires."""
        n = self.n_wires
        self.reg._push(op)
        self.reg._push(dataclasses.replace(op, wires=tuple(w + n for w in op.wires), conj=not op.conj))

    def flush(self, options: Optional[E.EngineOptions] = None):
        self.reg.flush(options)

    # ------------------------------------------------------------ channels
    def kraus(self, wires: Union[int, Sequence[int]], kraus_ops: Sequence[np.ndarray]):
        """rho -> sum_k K_k rho K_k^dagger on ``wires`` (K_k: 2^k x 2^k, wires[0]
        the most significant bit, as for the gates)."""
        ws = [int(wires)] if isinstance(wires, (int, np.integer)) else [int(w) for w in wires]
        n = self.n_wires
        if len(set(ws)) != len(ws) or any(not 0 <= w < n for w in ws):
            raise ValueError(f"bad wires {ws} for {n} wires")
        S = superoperator(kraus_ops)
        if S.shape[0] != 1 << (2 * len(ws)):
            raise ValueError(f"Kraus operators of size {int(np.sqrt(S.shape[0]))} do not match {len(ws)} wire(s)")
        if len(ws) > 5:
            raise ValueError("channels act on at most 5 wires (a 2^(2k) superoperator, 2k <= 10)")
        mat = torch.tensor(S, dtype=C_DTYPE, device=self.reg._sv_device())
        reg_wires = tuple(ws) + tuple(w + n for w in ws)
        self.reg._push(E.TapeOp("qubitunitaryfast", reg_wires, mat, False, mat._version, nonunitary=True))

    def amplitude_damping(self, wires, gamma: float):
        for w in ([wires] if isinstance(wires, (int, np.integer)) else wires):
            self.kraus(int(w), amplitude_damping_kraus(gamma))

    def phase_damping(self, wires, lam: float):
        for w in ([wires] if isinstance(wires, (int, np.integer)) else wires):
            self.kraus(int(w), phase_damping_kraus(lam))

    def depolarizing(self, wires, p: float):
        for w in ([wires] if isinstance(wires, (int, np.integer)) else wires):
            self.kraus(int(w), depolarizing_kraus(p))

    def bit_flip(self, wires, p: float):
        for w in ([wires] if isinstance(wires, (int, np.integer)) else wires):
            self.kraus(int(w), bit_flip_kraus(p))

    def phase_flip(self, wires, p: float):
        for w in ([wires] if isinstance(wires, (int, np.integer)) else wires):
            self.kraus(int(w), phase_flip_kraus(p))

    # ------------------------------------------------------------ state
    @property
    def densities(self) -> torch.Tensor:
        """[bsz] + [2] * 2n: densities[b, r..., c...] = rho_b[r, c]."""
        return self.reg.states

    @densities.setter
    def densities(self, value: torch.Tensor):
        value = torch.as_tensor(value)
        bsz = value.shape[0]
        v = value.reshape(bsz, -1)
        if v.shape[1] != 1 << (2 * self.n_wires):
            raise ValueError(f"density has {v.shape[1]} entries per row, expected {1 << (2 * self.n_wires)}")
        self.reg.states = v

    @property
    def density(self) -> torch.Tensor:
        """The initial |0..0><0..0| as [2] * 2n (the reference's buffer)."""
        d = torch.zeros(1 << (2 * self.n_wires), dtype=C_DTYPE, device=self.reg._sv_device())
        d[0] = 1.0
        return d.reshape([2] * (2 * self.n_wires))

    def matrices(self) -> torch.Tensor:
        """[bsz, 2^n, 2^n] view of the density matrices."""
        d = 1 << self.n_wires
        return self.reg.get_states_1d().reshape(self.bsz, d, d)

    def get_probs_1d(self) -> torch.Tensor:
        """Diagonals rho_ii, [bsz, 2^n] (noisedevices.py:88-92)."""
        return torch.diagonal(self.matrices(), offset=0, dim1=1, dim2=2)

    def get_prob_1d(self) -> torch.Tensor:
        """Diagonal of the initial density (noisedevices.py:94-97)."""
        d = 1 << self.n_wires
        return torch.diagonal(self.density.reshape(d, d), offset=0, dim1=0, dim2=1)

    def reset_states(self, bsz: int):
        self.reg.reset_states(bsz)

    def trace(self) -> torch.Tensor:
        return self.get_probs_1d().sum(-1)

    def purity(self) -> torch.Tensor:
        """tr(rho^2) = sum |rho_ij|^2 (rho Hermitian), [bsz]."""
        m = self.matrices()
        return (m.real ** 2 + m.imag ** 2).sum((-1, -2))

    @property
    def name(self):
        return self.__class__.__name__

    def __repr__(self):
        return (f" class: {self.name} \n device name: {self.device_name} \n number of qubits: {self.n_wires} \n"
                f" batch size: {self.bsz} \n current computing device: {self.reg._sv_device()} \n"
                f" recording op history: {self.record_op} \n"
                f" current states: {repr(self.get_probs_1d().cpu().detach().numpy())}")

    def _apply(self, fn, recurse=True):
        return super()._apply(fn, recurse)


# ----------------------------------------------------------------------------
# expectations on densities: Tr(rho P) for Pauli strings
# ----------------------------------------------------------------------------
def _pauli_masks(obs: str):
    n = len(obs)
    x = z = ny = 0
    for w, ch in enumerate(obs.upper()):
        bit = 1 << (n - 1 - w)
        if ch in "XY":
            x |= bit
        if ch in "ZY":
            z |= bit
        if ch == "Y":
            ny += 1
        if ch not in "IXYZ":
            raise ValueError(f"unknown Pauli character '{ch}' in '{obs}'")
    return x, z, ny


def density_expval_strings(dev: NoiseDevice, observables: Sequence[str]) -> torch.Tensor:
    """Re Tr(rho P_t) for Pauli strings (char k <-> wire k) -> [bsz, T].

    P|r> = i^{nY} (-1)^{popcount(r & z)} |r ^ x>, so
    Tr(rho P) = i^{nY} sum_r (-1)^{popcount(r & z)} rho[r, r ^ x]:
    a gather of 2^n entries of the 4^n per row (differentiable; the gradient
    reaches the gates through the engine's adjoint sweep)."""
    n = dev.n_wires
    flat = dev.reg.get_states_1d()                         # [bsz, 4^n]
    d = 1 << n
    r = np.arange(d, dtype=np.int64)
    cols = []
    for obs in observables:
        if len(obs) != n:
            raise ValueError(f"observable '{obs}' has {len(obs)} characters for {n} wires")
        x, z, ny = _pauli_masks(obs)
        idx = torch.from_numpy(r * d + (r ^ x)).to(flat.device)
        par = np.array([bin(v).count("1") & 1 for v in (r & z)], dtype=np.float32) if z else np.zeros(d, np.float32)
        sgn = torch.from_numpy(1.0 - 2.0 * par).to(flat.device)
        s = (flat[:, idx] * sgn).sum(-1)                       # sum_r sign rho[r, r^x]
        ph = 1j ** ny                                          # Re(i^ny s)
        cols.append(s.real * ph.real - s.imag * ph.imag)
    return torch.stack(cols, dim=-1)
