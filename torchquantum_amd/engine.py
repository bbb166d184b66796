"""Circuit engine: tape -> kernel ops -> plan -> launches (forward and adjoint).

The reference dispatches every gate eagerly through ``gate_wrapper``
(torchquantum/functional/gate_wrapper.py:324-446) into a permute + bmm +
permute (``apply_unitary_bmm``, :96-140).  Here gate calls are recorded on a
per-device tape; when the state is needed the tape is compiled once per
circuit *structure* into a plan and executed by the HIP kernels in
``libtqsv.so``:

* ``pergate``  -- one strided kernel per gate (the plain baseline);
* ``fused``    -- greedy partition of the gate DAG into stages of <= m local
  bits, each stage one read + write of the state (tqsv_stage.hip), gates
  that fit no stage run as single per-gate launches in between.

Matrices and their parameter derivatives are evaluated on the host in
float64 (vectorised per gate type, derivatives by forward-mode AD) and
uploaded as one pool per call; the plan's program words are uploaded once.

Bit positions: wire w of an n-wire device is bit ``n - 1 - w`` of the in-row
amplitude index (torchquantum/device/devices.py:64-71).
"""
from __future__ import annotations

import math
import os
import threading
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import gates as G
from . import _native as N

R_BITS = 4            # register bits per thread in the stage kernel
MAX_FUSION_WIDTH = 6  # widest fused block: 5- and 6-wire blocks run on the matrix-core kernel between stages
MAX_STAGE_BITS = 13   # LDS: 2^13 * 8 B = 64 KiB per workgroup
MAX_GP_STAGE = 1024
MAX_P_PER_OP = 4

# stage header layout (must match tqsv_stage.hip)
H_NOPS, H_M, H_T, H_NNL, H_OFF_OPS, H_NGP, H_OFF_GSLOT, H_FLAGS, H_POS = range(9)
OP_MAT1, OP_MAT2, OP_DIAG, OP_MATK, OP_REMAP = 1, 2, 3, 4, 5
F_BATCHED, F_NOCTRL, F_REAL, F_XPERM, F_SIGN, F_MG = 1 << 8, 1 << 9, 1 << 10, 1 << 11, 1 << 12, 1 << 13
MG_ITEM = np.dtype([("mg_slot", "<i4"), ("P", "<i4"), ("slot", "<i4", (4,)), ("u_off", "<i8"), ("bstride", "<i8")])
MG_SUMS = 4     # Hermitian-generator sums S_I, S_X, S_Y, S_Z per F_MG op
# single-wire gates whose matrices the GPU builds from device parameters
# (tqsv_devgates.hip), and the fixed single-wire gates that may sit in a
# merged run built there
DEV_GATE_CODE = {"rx": 1, "ry": 2, "rz": 3, "phaseshift": 4, "u1": 4, "u2": 5, "u3": 6, "rot": 7}
DEV_FIXED = ("hadamard", "shadamard", "paulix", "pauliy", "sx", "sxdg")
DG_MAX_MEMBERS = 8
DGF_INVERSE, DGF_CONJ, DGF_DERIV, DGF_DIAG = 1, 2, 4, 8
DEV_OP = np.dtype([("dst", "<i8"), ("bstride", "<i8"), ("adj", "<i8"), ("abstride", "<i8"), ("rows", "<i4"),
                   ("m0", "<i4"), ("nm", "<i4"), ("flags", "<i4")])
DEV_MEMBER = np.dtype([("params", "<u8"), ("rstride", "<i8"), ("cstride", "<i8"), ("code", "<i4"), ("flags", "<i4"),
                       ("fixed", "<f4", (8,))])
FEAT_MAT2, FEAT_MATK = 1, 2
# H_FLAGS: bits 0-1 features; bits 2-14 / 15-27: register local bits (mask)
# of the layout the stage loads in / stores in (0 = the canonical layout)
FL_LOAD_SHIFT, FL_STORE_SHIFT = 2, 15


# ----------------------------------------------------------------------------
# options
# ----------------------------------------------------------------------------
@dataclass
class EngineOptions:
    """Execution options of one device (``QuantumDevice.engine_options``).

    mode:        "fused" (default) or "pergate".
    stage_bits:  m, local bits per stage (R_BITS + log2 threads per group); 0 = auto
                 (13 for states of more than 512 MiB, else min(n, 12)).
    low_bits:    lowest bit positions always kept local (coalescing).  With 4,
                 a wavefront's loads cover whole 128-byte lines; fewer leave
                 64-byte runs, more constrain the partition into more stages.
    merge_1q:    multiply runs of single-wire gates into one 2x2 (fused mode).
    fusion_width: k; >= 2 pre-multiplies runs of gates acting within k wires
                 into one dense 2^k x 2^k block (k <= 6; trainable blocks k <= 2).
                 Blocks of <= 4 wires run inside stages; 5- and 6-wire blocks
                 run as their own launches on the matrix cores.
    grid_cap:    workgroups per stage launch (0 = auto).
    adjoint_mg:  adjoint sweep accumulates, for trainable single-wire ops in
                 stages, the four sums Im<lam|s|psi> (s = I, X, Y, Z) that
                 carry every parameter's gradient through the Hermitian
                 generator i dU/dtheta U^dagger (else one inner product per
                 parameter).
    device_gates: single-wire ops (one gate or a merged run) whose
                 parameters -- shared [1, P] or per-row [bsz, P], RX, RY, RZ,
                 PhaseShift/U1, U2, U3, Rot, with fixed H, SH, X, Y, SX, SXDG
                 in runs -- live on the GPU as float32 have their matrices
                 and derivatives built by a kernel straight into the matrix
                 pool: no device-to-host copy of the parameters (trainable
                 weights after an optimizer step, encoder inputs) and no host
                 matrix work.
    jit:         "auto" | "on" | "off": run stages as kernels compiled at run
                 time from their programs (straight-line code, hipRTC) instead
                 of the program-interpreting kernels.  "auto" compiles for
                 states of >= 2^22 amplitudes (bsz * 2^n) on a background
                 thread (about 0.5-2 s per stage and direction; code objects
                 are cached on disk) while the interpreting kernels run, and
                 switches once they are ready (wait_for_compiled_stages()
                 blocks until then).  "on" compiles before the first pass.
    reorder_ops: within a stage, run commuting ops in an order that needs
                 fewer register-set changes (LDS remaps); ops sharing a wire
                 keep their order.
    grad_cap:    at most this many ops with gradient parameters per stage
                 (0: no limit); evens out the adjoint stages' arithmetic.
    io_layouts:  a stage may load in the register layout of its first ops
                 and store in that of its last (when the lowest 4 local bits
                 stay thread bits, so accesses still cover whole 128-byte
                 lines) instead of remapping from and to the canonical one.
    partition:   "greedy" (local bits taken in op order), "lookahead" (each
                 stage's local bits chosen one at a time, the bit that lets
                 the most ops into the stage first: fewer stages, about 0.3 s
                 more planning per thousand ops) or "auto" (lookahead for
                 states of more than 2^22 amplitudes, where a pass costs tens
                 of microseconds or more and a plan is reused; the autotuner's
                 many small candidate plans keep the greedy pass).
    """
    mode: str = "fused"
    stage_bits: int = 0
    low_bits: int = 4
    merge_1q: bool = True
    fusion_width: int = 1
    grid_cap: int = 0
    adjoint_mg: bool = True
    jit: str = "auto"
    device_gates: bool = True
    reorder_ops: bool = True
    io_layouts: bool = True
    grad_cap: int = 0
    partition: str = "auto"

    def key(self):
        return (self.mode, self.stage_bits, self.low_bits, self.merge_1q, self.fusion_width, self.grid_cap,
                self.adjoint_mg, self.jit, self.device_gates, self.reorder_ops, self.io_layouts, self.grad_cap,
                self.partition)


DEFAULT_OPTIONS = EngineOptions()
JIT_MIN_AMPS = 1 << 22
# TQSV_JIT_NT=1 compiles stages that stream the state with non-temporal
# hints (states of at least 2^25 amplitudes = 256 MiB, past the Infinity
# Cache).  Off by default: on the 24-qubit bsz-64 gradient it moved the
# forward stage 4.01 -> 3.95 ms and the adjoint stage 9.18 -> 9.27 ms.
JIT_NT_AMPS = 1 << 25

# Optional per-launch timing of stage kernels: when set to a list, every
# stage launch appends (direction "f"/"a", compiled?, start, end), two HIP
# events recorded on the launch stream around it.
_stage_timer: Optional[list] = None


def set_stage_timer(records: Optional[list]) -> None:
    """Record HIP events around every stage launch into ``records`` (None
    switches it off).  Used to time the stage kernels inside a timed loop."""
    global _stage_timer
    _stage_timer = records


def _timed(direction: str, compiled: bool):
    """Start event (or None) for a stage launch when the stage timer is on."""
    if _stage_timer is None:
        return None
    e0 = torch.cuda.Event(enable_timing=True)
    e0.record()
    return (direction, compiled, e0)


def _timed_end(tok):
    if tok is not None and _stage_timer is not None:
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        _stage_timer.append((tok[0], tok[1], tok[2], e1))


def wait_for_compiled_stages(timeout: Optional[float] = None) -> bool:
    """Block until the background compiles of stage kernels (jit="auto") have
    finished; later passes of those plans then run the compiled kernels.
    Returns False if the timeout passed first."""
    import sys
    sj = sys.modules.get(__package__ + ".stagejit")
    return True if sj is None else sj.wait_all(timeout)


# ----------------------------------------------------------------------------
# tape
# ----------------------------------------------------------------------------
@dataclass(slots=True)
class TapeOp:
    name: str                       # canonical gate name
    wires: Tuple[int, ...]
    params: Optional[torch.Tensor]  # [rows, P] real, or complex matrix [d,d] / [B,d,d]
    inverse: bool = False
    version: int = 0
    conj: bool = False              # apply the complex conjugate of the matrix (density column index)
    nonunitary: bool = False        # a channel's superoperator: no adjoint (gradient) sweep through it

    @property
    def is_matrix_param(self) -> bool:
        return self.params is not None and self.params.is_complex()

    def rows(self) -> int:
        if self.params is None:
            return 1
        if self.is_matrix_param:
            return self.params.shape[0] if self.params.dim() == 3 else 1
        return self.params.shape[0]

    def n_params(self) -> int:
        if self.params is None or self.is_matrix_param:
            return 0
        return self.params.shape[-1]

    def wants_grad(self) -> bool:
        return self.params is not None and self.params.requires_grad


# ----------------------------------------------------------------------------
# kernel-level ops
# ----------------------------------------------------------------------------
@dataclass
class KOp:
    kind: str                       # mat1 | mat2 | matk | diag | swap
    bits: Tuple[int, ...]           # mat1: (t,); mat2: (hi, lo); matk: MSB..LSB; diag: parity bits; swap: (a, b)
    cbits: Tuple[int, ...] = ()     # control bit positions
    cvals: Tuple[int, ...] = ()     # required values
    src: Tuple[int, ...] = ()       # tape indices (in application order)
    extract: str = "full"           # how to take this op's matrix from the tape op matrix
    diag_idx: Tuple[int, int] = (0, 0)
    grad_slots: Tuple[int, ...] = ()  # flat parameter slots (one per kernel parameter)
    batched: bool = False
    n_for_block: int = 0            # qubit count (blocks map tape wires to bit positions)
    real: bool = False              # matrix entries are real for every parameter value
    xperm: bool = False             # the 2x2 is Pauli X (a controlled exchange)
    sign: bool = False              # diagonal phase -1 where the controls match, no parity
    dev: bool = False               # matrix built on the GPU from device parameters

    @property
    def cmask(self) -> int:
        m = 0
        for b in self.cbits:
            m |= 1 << b
        return m

    @property
    def cval(self) -> int:
        v = 0
        for b, x in zip(self.cbits, self.cvals):
            if x:
                v |= 1 << b
        return v

    @property
    def allbits(self) -> frozenset:
        return frozenset(self.bits) | frozenset(self.cbits)

    @property
    def P(self) -> int:
        return len(self.grad_slots)

    def dim(self) -> int:
        if self.kind == "diag":
            return 0
        if self.kind == "mat1":
            return 2
        if self.kind in ("mat2", "swap"):
            return 4
        return 1 << len(self.bits)


# gate structure that lets the stage kernel take a cheaper path
X_GATES = {"paulix", "cnot", "toffoli", "c3x", "c4x", "multicnot", "multixcnot"}
SIGN_GATES = {"pauliz", "cz", "ccz"}
REAL_1Q_BLOCK = X_GATES | {"ry", "hadamard", "shadamard", "cry", "chadamard", "pauliz", "i"}


def dev_member_code(op: TapeOp, bsz: int) -> Optional[int]:
    """Gate code of a tape op the device builder can take as a member of a
    single-wire op (0: a fixed gate), None if it cannot."""
    p = op.params
    if len(op.wires) != 1:
        return None
    if p is None:
        return 0 if op.name in DEV_FIXED else None
    if p.dim() != 2 or p.shape[0] not in (1, bsz) or not p.is_cuda or p.dtype != torch.float32:
        return None
    return DEV_GATE_CODE.get(op.name)


def mark_device_built(kops: List[KOp], tape: Sequence[TapeOp], bsz: int) -> None:
    """Flag the single-wire kernel ops whose matrices the GPU builds: an
    uncontrolled dense 2x2, a diagonal, or a merged run, all of whose members
    the builder takes, with at least one parameterised member."""
    for k in kops:
        if k.kind not in ("mat1", "diag") or len(k.src) > DG_MAX_MEMBERS:
            continue
        if k.kind == "mat1" and (k.cbits or k.extract not in ("full", "chain")):
            continue
        if k.kind == "diag" and (len(k.src) != 1 or len(tape[k.src[0]].wires) != 1):
            continue
        codes = [dev_member_code(tape[t], bsz) for t in k.src]
        if all(c is not None for c in codes) and any(codes):
            k.dev = True


def lower(tape: Sequence[TapeOp], n: int, bsz: int, param_offsets: Sequence[int]) -> List[KOp]:
    """Tape ops -> kernel ops (no fusion).  Identity gates vanish."""
    out: List[KOp] = []
    for ti, op in enumerate(tape):
        s = G.spec(op.name)
        k = len(op.wires)
        pos = [n - 1 - w for w in op.wires]
        rows = op.rows()
        if rows not in (1, bsz):
            raise ValueError(f"gate '{op.name}': parameter batch {rows} does not match bsz {bsz}")
        batched = rows > 1
        slots = ()
        if op.wants_grad():
            slots = tuple(param_offsets[ti] + p for p in range(op.n_params()))
        if s.name == "i":
            continue
        if s.kind == "diag":
            cidx, pidx = G.diag_form(s, k)
            out.append(KOp("diag", tuple(pos[i] for i in pidx), tuple(pos[i] for i in cidx),
                           tuple(1 for _ in cidx), (ti,), "diag",
                           G.diag_phase_indices(cidx, pidx, k), slots, batched,
                           sign=s.name in SIGN_GATES, real=s.name in SIGN_GATES))
        elif s.kind == "ctrl1":
            ci, cv, t = G.ctrl_values(s, k)
            out.append(KOp("mat1", (pos[t],), tuple(pos[i] for i in ci), tuple(cv), (ti,),
                           "ctrl0" if s.cval0 else "ctrl1", (0, 0), slots, batched,
                           real=s.name in REAL_1Q_BLOCK, xperm=s.name in X_GATES))
        elif s.kind == "swap":
            out.append(KOp("swap", (pos[-2], pos[-1]), tuple(pos[:-2]), tuple(1 for _ in pos[:-2]), (ti,),
                           "swap", (0, 0), (), False))
        else:
            if op.is_matrix_param and op.params.requires_grad:
                raise NotImplementedError("gradients through a matrix-valued gate parameter "
                                          "(qubitunitary*) are not supported")
            kind = "mat1" if k == 1 else ("mat2" if k == 2 else "matk")
            out.append(KOp(kind, tuple(pos), (), (), (ti,), "full", (0, 0), slots, batched,
                           real=k == 1 and s.name in REAL_1Q_BLOCK, xperm=k == 1 and s.name in X_GATES))
    return out


def merge_single_wire(kops: List[KOp], tape: Sequence[TapeOp]) -> List[KOp]:
    """Fuse runs of uncontrolled single-wire ops on the same bit into one mat1.

    A later op on bit b joins the pending run on b if nothing touched b in
    between; the run's matrix is the product of its members.  Runs are capped
    at MAX_P_PER_OP trainable parameters."""
    out: List[Optional[KOp]] = []
    pending: Dict[int, int] = {}   # bit -> index in out of the open run

    def single(k: KOp) -> bool:
        if k.cbits:
            return False
        if k.kind == "mat1":
            return True
        return k.kind == "diag" and len(k.bits) == 1

    for k in kops:
        if single(k):
            b = k.bits[0]
            j = pending.get(b)
            if j is not None:
                run = out[j]
                if run.P + k.P <= MAX_P_PER_OP:
                    out[j] = KOp("mat1", (b,), (), (), run.src + k.src, "chain", (0, 0),
                                 run.grad_slots + k.grad_slots, run.batched or k.batched,
                                 real=run.real and k.real)
                    continue
            out.append(KOp("mat1", (b,), (), (), k.src, "chain", (0, 0), k.grad_slots, k.batched, real=k.real)
                       if k.kind == "diag" else k)
            pending[b] = len(out) - 1
        else:
            for b in k.allbits:
                pending.pop(b, None)
            out.append(k)
    return [k for k in out if k is not None]


def fuse_blocks(kops: List[KOp], tape: Sequence[TapeOp], n: int, width: int) -> List[KOp]:
    """Greedy gate fusion into dense blocks of <= ``width`` bits.

    Walking the ops in order, an op joins the blocks that last touched its
    bits.  Those blocks are merged with it into one block when each of them is
    still open (no other block has touched any of its bits since) and the
    union stays within ``width`` bits: the merged block then sits at the op's
    position, which is legal because nothing after the merged blocks acts on
    their bits.  Multi-member blocks become one dense op (mat1 / mat2 / matk)
    whose matrix is the ordered product of its members embedded in the
    block's bits.  Blocks with trainable members are kept to <= 2 bits and <=
    MAX_P_PER_OP parameters (the widths whose kernels return parameter
    gradients).  Blocks of 5 or 6 bits run on the matrix-core kernel."""
    width = max(1, min(width, MAX_FUSION_WIDTH))
    blocks: List[Optional[dict]] = []
    last: Dict[int, int] = {}

    def place(B: dict):
        # B follows every block last seen on its bits: those are closed (a
        # later block acts on one of their bits, so they can no longer move)
        for b in B["bits"]:
            j = last.get(b)
            if j is not None and blocks[j] is not None:
                blocks[j]["closed"] = True
        blocks.append(B)
        for b in B["bits"]:
            last[b] = len(blocks) - 1

    for i, k in enumerate(kops):
        bs = set(k.allbits)
        if len(k.src) != 1 or (k.kind == "matk" and len(k.bits) > R_BITS) or len(bs) > width:
            place(dict(bits=bs, members=[i], solo=True, P=k.P, closed=False))
            continue
        cands = sorted({last[b] for b in bs if b in last})
        cb = [blocks[j] for j in cands]
        if cands and not any(B["solo"] or B["closed"] for B in cb):
            nb = set(bs)
            P = k.P
            for B in cb:
                nb |= B["bits"]
                P += B["P"]
            lim = 2 if P > 0 else width
            if len(nb) <= min(width, lim) and P <= MAX_P_PER_OP:
                # the open candidates move to this op's position: nothing after
                # them acts on their bits
                members = sorted([m for B in cb for m in B["members"]] + [i])
                for j in cands:
                    blocks[j] = None
                place(dict(bits=nb, members=members, solo=False, P=P, closed=False))
                continue
        place(dict(bits=bs, members=[i], solo=False, P=k.P, closed=False))
    out = []
    for B in blocks:
        if B is None:
            continue
        if len(B["members"]) == 1:
            out.append(kops[B["members"][0]])
            continue
        bits = tuple(sorted(B["bits"], reverse=True))       # matrix MSB first
        kind = "mat1" if len(bits) == 1 else ("mat2" if len(bits) == 2 else "matk")
        src = tuple(kops[m].src[0] for m in B["members"])
        slots = tuple(s for m in B["members"] for s in kops[m].grad_slots)
        batched = any(kops[m].batched for m in B["members"])
        out.append(KOp(kind, bits, (), (), src, "block", (0, 0), slots, batched))
    return out


def embed(u: np.ndarray, op_bits: Sequence[int], block_bits: Sequence[int]) -> np.ndarray:
    """[rows, 2^k, 2^k] on op_bits (MSB first) -> [rows, 2^K, 2^K] on block_bits."""
    k, K = len(op_bits), len(block_bits)
    rest = [b for b in block_bits if b not in op_bits]
    rows = u.shape[0]
    full = np.kron(u, np.eye(1 << (K - k))[None]) if K > k else u          # order: op_bits + rest
    order = list(op_bits) + rest
    perm = [order.index(b) for b in block_bits]
    t = full.reshape([rows] + [2] * (2 * K))
    t = t.transpose([0] + [1 + p for p in perm] + [1 + K + p for p in perm])
    return t.reshape(rows, 1 << K, 1 << K)


# ----------------------------------------------------------------------------
# matrix values (host, float64)
# ----------------------------------------------------------------------------
class _GVLayout:
    """The structural part of GateValues for one plan (it depends only on the
    tape structure, which the plan cache key fixes): tape ops grouped by
    (gate, width, has-params), each op's group key, first row in its group,
    row count and whether its derivatives are wanted."""

    def __init__(self, tape: Sequence[TapeOp], skip: frozenset = frozenset()):
        T = len(tape)
        self.skip = skip        # tape ops whose matrices the GPU builds: no host values
        self.gkey: List[Optional[tuple]] = [None] * T
        self.grow = np.zeros(T, dtype=np.int64)
        self.rows = np.ones(T, dtype=np.int64)
        self.dwant = np.zeros(T, dtype=bool)
        self.mat_ops: List[int] = []
        groups: Dict[tuple, List[int]] = {}
        for i, op in enumerate(tape):
            if i in skip:
                continue
            if op.is_matrix_param:
                self.mat_ops.append(i)
                self.gkey[i] = ("__mat__", i)
                self.rows[i] = op.params.shape[0] if op.params.dim() == 3 else 1
                continue
            key = (op.name, len(op.wires), op.n_params() > 0)
            groups.setdefault(key, []).append(i)
        self.groups = []
        for key, idx in groups.items():
            has_p = key[2]
            inv_rows = None
            cj_rows = None
            if has_p:
                off = 0
                inv, cj = [], []
                for i in idx:
                    r = tape[i].params.shape[0]
                    self.grow[i], self.rows[i] = off, r
                    self.dwant[i] = tape[i].wants_grad()
                    inv.extend([tape[i].inverse] * r)
                    cj.extend([tape[i].conj] * r)
                    off += r
                inv_rows = np.array(inv, dtype=bool)
                cj_rows = np.array(cj, dtype=bool)
                P = tape[idx[0]].params.shape[-1]
            else:
                P = 0
                for i in idx:   # rows: U, U^dagger, conj(U), U^T
                    self.grow[i] = (1 if tape[i].inverse else 0) + (2 if tape[i].conj else 0)
            for i in idx:
                self.gkey[i] = key
            self.groups.append((key, idx, P, inv_rows, bool(self.dwant[idx].any()) if has_p else False, cj_rows))


class _GroupViews:
    """gv.U[i] / gv.D[i]: tape op i's rows of its group's matrices (a view)."""

    def __init__(self, gv: "GateValues", deriv: bool):
        self.gv, self.deriv = gv, deriv

    def __getitem__(self, i: int):
        gv = self.gv
        key = gv.gkey[i]
        if key is None:
            return None
        if self.deriv:
            if not gv.dwant[i]:
                return None
            M = gv.DG.get(key)
        else:
            M = gv.UG.get(key)
        if M is None:
            return None
        g, r = int(gv.grow[i]), int(gv.rows[i])
        return M[g:g + r]


class GateValues:
    """Gate matrices for the current parameter values, evaluated on the host.

    Tape ops are grouped by (gate, width, has-params) (``_GVLayout``); each
    group's matrices are built in one vectorised call into ``UG[key]``
    ([rows, d, d] complex128, inverse ops already conjugate-transposed) and,
    when any member is trainable, derivative matrices ``DG[key]``
    ([rows, P, d, d]).  ``gkey[i]``, ``grow[i]``, ``rows[i]`` locate tape op
    i; ``U[i]`` / ``D[i]`` are views.  Parameters living on the GPU are
    fetched with one copy (or reused from their host copies)."""

    def __init__(self, tape: Sequence[TapeOp], want_grad: bool, layout: Optional[_GVLayout] = None):
        lay = layout if layout is not None else _GVLayout(tape)
        self.gkey, self.grow, self.rows = lay.gkey, lay.grow, lay.rows
        self.dwant = lay.dwant if want_grad else np.zeros(len(tape), dtype=bool)
        self.UG: Dict[tuple, np.ndarray] = {}
        self.DG: Dict[tuple, Optional[np.ndarray]] = {}
        self.U = _GroupViews(self, False)
        self.D = _GroupViews(self, True)
        host = _host_params(tape, lay.skip)
        for i in lay.mat_ops:
            op = tape[i]
            m = host[i].astype(np.complex128)
            m = m.reshape((-1,) + m.shape[-2:]) if m.ndim == 3 else m[None]
            if op.inverse:
                m = np.conj(np.swapaxes(m, -1, -2))
            if op.conj:
                m = np.conj(m)
            self.UG[lay.gkey[i]] = m
            self.DG[lay.gkey[i]] = None
        for key, idx, P, inv, any_grad, cj in lay.groups:
            name, k, has_p = key
            builder = G.width_builder(name, k)
            if has_p:
                allp = torch.from_numpy(np.concatenate([host[i].reshape(-1, P) for i in idx], axis=0))
                U = builder(allp).numpy()
                D = None
                if want_grad and any_grad:
                    def one(q, _b=builder):
                        return torch.view_as_real(_b(q.unsqueeze(0))[0])
                    J = torch.func.vmap(torch.func.jacfwd(one))(allp)        # [N, d, d, 2, P]
                    D = torch.complex(J[..., 0, :], J[..., 1, :]).movedim(-1, 1).numpy()  # [N,P,d,d]
                if inv.any():
                    U = U.copy()
                    U[inv] = np.conj(np.swapaxes(U[inv], -1, -2))
                    if D is not None:
                        D = D.copy()
                        D[inv] = np.conj(np.swapaxes(D[inv], -1, -2))
                if cj.any():     # d conj(U) / d theta = conj(dU / d theta) for real theta
                    U = U.copy()
                    U[cj] = np.conj(U[cj])
                    if D is not None:
                        D = D.copy()
                        D[cj] = np.conj(D[cj])
                self.UG[key], self.DG[key] = U, D
            else:
                u = builder(torch.zeros(1, 1, dtype=torch.float64)).numpy()
                ud = np.conj(np.swapaxes(u, -1, -2))
                # rows: 0 U, 1 U^dagger, 2 conj(U), 3 conj(U^dagger) = U^T
                self.UG[key] = np.concatenate([u, ud, np.conj(u), np.conj(ud)], axis=0)
                self.DG[key] = None


class _HostMirror:
    """Host copies of GPU parameter tensors, reused while a tensor is unchanged.

    Gate matrices are evaluated on the host, so GPU-resident parameters are
    copied to the host once per flush.  That copy is ordered after all work
    already queued on the stream and would stall the host until the GPU
    drains.  A copy is reused while torch's version counter and the data
    pointer of the tensor (or of the base of a view) are unchanged: every
    in-place write (optimizer steps, copy_, .data assignment) invalidates it.
    Entries die with their tensors (weak keys)."""

    def __init__(self):
        self._d: Dict[int, tuple] = {}      # id(base) -> (weakref(base), stamp, {view: array})
        # reentrant: a weakref callback (_drop) can run from a collection
        # triggered while this thread holds the lock
        self._lock = threading.RLock()

    @staticmethod
    def _key(p: torch.Tensor):
        base = p._base
        if base is None:          # the common case: the parameter tensor itself
            return p, None, (p._version, p.data_ptr())
        view = (p.storage_offset(), tuple(p.shape), tuple(p.stride()), p.dtype)
        return base, view, (base._version, base.data_ptr())

    def _drop(self, key: int):
        with self._lock:
            self._d.pop(key, None)

    def get(self, p: torch.Tensor) -> Optional[np.ndarray]:
        base, view, stamp = self._key(p)
        with self._lock:
            e = self._d.get(id(base))
            if e is None or e[0]() is not base or e[1] != stamp:
                return None
            return e[2].get(view)

    def put(self, p: torch.Tensor, value: np.ndarray):
        import weakref
        base, view, stamp = self._key(p)
        key = id(base)
        with self._lock:
            e = self._d.get(key)
            if e is None or e[0]() is not base or e[1] != stamp:
                e = (weakref.ref(base, lambda _r, k=key: self._drop(k)), stamp, {})
                self._d[key] = e
            e[2][view] = value


_MIRROR = _HostMirror()


class _PinnedRing:
    """Matrix-pool uploads through a fixed ring of pinned host buffers.

    A flush copies its pool host -> device asynchronously, so the host can run
    ahead of the GPU.  A slot is refilled only after the copy that last used
    it has completed (its event), which bounds how far ahead the host runs;
    buffers grow on demand and are never freed, so steady-state flushes make
    no pinned allocations."""

    SLOTS = 8

    def __init__(self):
        self.slots: List[Optional[tuple]] = [None] * self.SLOTS
        self.i = 0
        self.lock = threading.Lock()

    def upload(self, host: np.ndarray, device: torch.device) -> torch.Tensor:
        flat = np.ascontiguousarray(host, dtype=np.complex64).reshape(-1)
        n = flat.size
        with self.lock:
            i = self.i
            self.i = (i + 1) % self.SLOTS
            buf, ev = self.slots[i] if self.slots[i] is not None else (None, None)
            if ev is not None:
                ev.synchronize()
            if buf is None or buf.numel() < n:
                cap = max(n, 2 * (buf.numel() if buf is not None else 0), 4096)
                buf = torch.empty(cap, dtype=torch.complex64, pin_memory=True)
                # the first use fills every empty slot too, so that later
                # flushes make no pinned allocation (slots that are still
                # empty have no copy in flight)
                for j in range(self.SLOTS):
                    if j != i and self.slots[j] is None:
                        self.slots[j] = (torch.empty(cap, dtype=torch.complex64, pin_memory=True), None)
            buf[:n].numpy()[:] = flat
            out = torch.empty(n, dtype=torch.complex64, device=device)
            out.copy_(buf[:n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(device))
            self.slots[i] = (buf, ev)
        return out.reshape(host.shape)


_RINGS: Dict[torch.device, _PinnedRing] = {}


_DEV_LAYOUT_OK = False


def _check_dev_layout():
    global _DEV_LAYOUT_OK
    if not _DEV_LAYOUT_OK:
        want = DEV_OP.itemsize | (DEV_MEMBER.itemsize << 16)
        if N.lib().tqsv_dev_gate_bytes() != want:
            raise RuntimeError("DevOp / DevMember layout mismatch with libtqsv.so")
        _DEV_LAYOUT_OK = True


def _upload(host: np.ndarray, device: torch.device) -> torch.Tensor:
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    ring = _RINGS.get(device)
    if ring is None:
        ring = _RINGS.setdefault(device, _PinnedRing())
    return ring.upload(host, device)


def _host_params(tape: Sequence[TapeOp], skip: frozenset = frozenset()) -> List[Optional[np.ndarray]]:
    """Parameter values of every tape op as numpy (float64 or complex128),
    except the ops in ``skip`` (their matrices are built on the GPU).  GPU
    tensors without a valid host copy are gathered with one concatenation
    and one device-to-host copy."""
    out: List[Optional[np.ndarray]] = [None] * len(tape)
    dev: Dict[tuple, List[int]] = {}
    cpu_real: List[int] = []
    entries = _MIRROR._d
    with _MIRROR._lock:    # one acquisition for the whole tape (reentrant)
        for i, op in enumerate(tape):
            p = op.params
            if p is None or i in skip:
                continue
            if not p.is_cuda:
                if p.is_complex():
                    out[i] = p.detach().numpy().astype(np.complex128)
                else:
                    cpu_real.append(i)
                continue
            if p._base is None:       # the common case, _HostMirror.get inlined
                e = entries.get(id(p))
                hit = e[2].get(None) if (e is not None and e[0]() is p
                                         and e[1] == (p._version, p.data_ptr())) else None
            else:
                hit = _MIRROR.get(p)
            if hit is not None:
                out[i] = hit
            else:
                dev.setdefault((p.device, p.dtype), []).append(i)
    if cpu_real:
        # real host parameters ([rows, P]): one concatenation and one
        # conversion per parameter width P
        byP: Dict[int, List[int]] = {}
        for i in cpu_real:
            byP.setdefault(tape[i].params.shape[-1], []).append(i)
        with torch.no_grad():
            for P, idx in byP.items():
                flat = torch.cat([tape[i].params for i in idx], 0).numpy().astype(np.float64)
                off = 0
                for i in idx:
                    r = tape[i].params.shape[0]
                    out[i] = flat[off:off + r]
                    off += r
    for (d, dt), idx in dev.items():
        flat = torch.cat([tape[i].params.detach().reshape(-1) for i in idx]).cpu().numpy()
        flat = flat.astype(np.complex128 if np.iscomplexobj(flat) else np.float64)
        off = 0
        for i in idx:
            sh = tuple(tape[i].params.shape)
            sz = int(np.prod(sh))
            out[i] = flat[off:off + sz].reshape(sh)
            out[i].flags.writeable = False
            _MIRROR.put(tape[i].params, out[i])
            off += sz
    return out


def _block(u: np.ndarray, extract: str) -> np.ndarray:
    """Take this op's matrix from a tape-op matrix [..., d, d]."""
    if extract == "ctrl1":
        return u[..., -2:, -2:]
    if extract == "ctrl0":
        return u[..., :2, :2]
    return u


def kop_data(k: KOp, gv: GateValues, tape: Sequence[TapeOp], with_grad: bool) -> np.ndarray:
    """Data of one kernel op: [rows, 1 + P, d*d] (or [rows, 1 + P, 2] for diag)."""
    if k.kind == "swap":
        m = G.FIXED_NP["swap"][None, None]
        return m.reshape(1, 1, 16)
    if k.kind == "diag":
        ti = k.src[0]
        u = gv.U[ti]
        a, b = k.diag_idx
        ph = np.stack([u[:, a, a], u[:, b, b]], axis=-1)[:, None]          # [rows, 1, 2]
        if with_grad and k.P:
            d = gv.D[ti]
            dph = np.stack([d[:, :, a, a], d[:, :, b, b]], axis=-1)         # [rows, P, 2]
            ph = np.concatenate([np.broadcast_to(ph, (dph.shape[0],) + ph.shape[1:]), dph], axis=1)
        return ph
    if k.extract == "chain":
        return _chain_data(k, gv, tape, with_grad)
    if k.extract == "block":
        return _block_data(k, gv, tape, with_grad)
    ti = k.src[0]
    u = _block(gv.U[ti], k.extract)
    d2 = u.shape[-1] ** 2
    parts = [u.reshape(u.shape[0], 1, d2)]
    if with_grad and k.P:
        d = _block(gv.D[ti], k.extract)
        rows = max(u.shape[0], d.shape[0])
        parts = [np.broadcast_to(parts[0], (rows, 1, d2)), d.reshape(d.shape[0], d.shape[1], d2)]
    return np.concatenate(parts, axis=1)


def _as2x2(k_src: int, gv: GateValues, tape, with_grad):
    """(U [rows,2,2], D list of [rows,2,2] or []) of a single-wire tape op."""
    op = tape[k_src]
    u = gv.U[k_src]
    d = gv.D[k_src] if with_grad and op.wants_grad() else None
    ds = [] if d is None else [d[:, p] for p in range(d.shape[1])]
    return u, ds


def _block_data(k: KOp, gv: GateValues, tape, with_grad) -> np.ndarray:
    n_bits = len(k.bits)
    mats, ders = [], []
    for mi, ti in enumerate(k.src):
        op = tape[ti]
        nq = k.n_for_block
        ob = [nq - 1 - w for w in op.wires]
        mats.append(embed(gv.U[ti], ob, k.bits))
        if with_grad and op.wants_grad():
            d = gv.D[ti]
            for p in range(d.shape[1]):
                ders.append((mi, embed(d[:, p], ob, k.bits)))
    return _product_data(mats, ders, 1 << n_bits, with_grad)


def _product_data(mats, ders, dim, with_grad) -> np.ndarray:
    rows = max(m.shape[0] for m in mats)
    ident = np.broadcast_to(np.eye(dim, dtype=np.complex128), (rows, dim, dim))
    prefix = [ident]
    for u in mats:
        prefix.append(np.matmul(u, prefix[-1]))
    parts = [prefix[-1].reshape(rows, 1, dim * dim)]
    if with_grad and ders:
        suffix = [ident] * (len(mats) + 1)
        for i in range(len(mats) - 1, -1, -1):
            suffix[i] = np.matmul(suffix[i + 1], mats[i])
        for mi, d in ders:
            dd = np.matmul(np.matmul(suffix[mi + 1], d), prefix[mi])
            parts.append(np.broadcast_to(dd, (rows, dim, dim)).reshape(rows, 1, dim * dim))
    return np.concatenate([np.broadcast_to(p, (rows,) + p.shape[1:]) for p in parts], axis=1)


def _chain_data(k: KOp, gv: GateValues, tape, with_grad) -> np.ndarray:
    mats = []
    ders = []   # list of (member index, derivative matrix)
    for mi, ti in enumerate(k.src):
        u, ds = _as2x2(ti, gv, tape, with_grad)
        mats.append(u)
        for d in ds:
            ders.append((mi, d))
    rows = max(m.shape[0] for m in mats)
    ident = np.broadcast_to(np.eye(2, dtype=np.complex128), (rows, 2, 2))
    # prefix[i] = U_{i-1} ... U_0 ; total = U_{m-1} ... U_0
    prefix = [ident]
    for u in mats:
        prefix.append(np.matmul(u, prefix[-1]))
    total = prefix[-1]
    parts = [total.reshape(rows, 1, 4)]
    if with_grad and ders:
        suffix = [ident] * (len(mats) + 1)
        for i in range(len(mats) - 1, -1, -1):
            suffix[i] = np.matmul(suffix[i + 1], mats[i]) if i < len(mats) else ident
        # suffix[i] = U_{m-1} ... U_i ; derivative of member i: suffix[i+1] @ dU_i @ prefix[i]
        for mi, d in ders:
            dd = np.matmul(np.matmul(suffix[mi + 1], d), prefix[mi])
            parts.append(np.broadcast_to(dd, (rows, 2, 2)).reshape(rows, 1, 4))
    return np.concatenate([np.broadcast_to(p, (rows,) + p.shape[1:]) for p in parts], axis=1)


# ----------------------------------------------------------------------------
# plans
# ----------------------------------------------------------------------------
@dataclass
class Stage:
    words: np.ndarray           # int32 program words of this stage
    ops: List[int]              # kernel-op indices, program order
    m: int
    enc: List[tuple]            # per encoded op: (kop index or -1 for remap, word offset, perm)


@dataclass
class Plan:
    n: int
    bsz: int
    options: EngineOptions
    kops: List[KOp]
    steps: List[tuple]          # ("stage", Stage) | ("solo", kop index)
    n_slots: int                # flat trainable-parameter count
    prog_dev: Optional[torch.Tensor] = None
    prog_host: Optional[np.ndarray] = None
    stage_offsets: List[int] = field(default_factory=list)


def _perm_index_bits(d_bits: int, src_of_dst: Sequence[int]) -> np.ndarray:
    """Index permutation: kernel index r -> gate index r' where kernel bit q
    takes gate bit src_of_dst[q] (bit numbers, LSB = 0)."""
    out = np.zeros(1 << d_bits, dtype=np.int64)
    for r in range(1 << d_bits):
        rr = 0
        for q in range(d_bits):
            if (r >> q) & 1:
                rr |= 1 << src_of_dst[q]
        out[r] = rr
    return out


WIDE_STAGE_BYTES = 1 << 29     # states above this use 13-bit stages by default


def choose_stage_bits(n: int, opts: EngineOptions, bsz: int = 1) -> int:
    """Local bits per stage: opts.stage_bits, or 13 for states of more than
    512 MiB and 12 below.  Measured with the compiled stage kernels: on
    HBM-resident states the wider stages need fewer passes (24 qubits bsz 64
    gradient: 30 -> 27 forward stages, 0.378 -> 0.358 s; 30 qubits bsz 4:
    forward 1.18 -> 0.97 s, gradient 3.62 -> 3.28 s); on a 128 MiB state held
    by the Infinity Cache they lose (20 qubits bsz 16 circuit: 1.54 -> 1.92 ms)."""
    if opts.stage_bits:
        return max(R_BITS, min(opts.stage_bits, n, MAX_STAGE_BITS))
    if n >= 13 and bsz * (8 << n) > WIDE_STAGE_BYTES:
        return 13
    return max(R_BITS, min(n, 12))


def stageable(k: KOp, m: int) -> bool:
    if len(k.allbits) > m:
        return False
    if k.kind == "matk" and len(k.bits) > R_BITS:
        return False
    return True


def use_lookahead(n: int, opts: EngineOptions, bsz: int = 1) -> bool:
    """Whether partition() chooses local bits by lookahead (opts.partition)."""
    if opts.partition in ("greedy", "lookahead"):
        return opts.partition == "lookahead"
    if opts.partition != "auto":
        raise ValueError(f"unknown partition {opts.partition!r}")
    return bsz * (1 << n) > JIT_MIN_AMPS


def _lookahead_stage(kops: List[KOp], remaining: List[int], n: int, m: int, forced: set,
                     grad_cap: int) -> Tuple[List[int], set]:
    """One stage of the lookahead partition: start from the first op's bits
    (with the forced low bits when they fit), then add, one at a time, the
    bit whose addition lets the most of the remaining ops into the stage
    (ties: the lowest bit).  An op enters when all its bits are local, none
    of them is blocked by an earlier op left out, and the gradient cap
    allows; a left-out op blocks its bits, so the stage's ops stay a prefix
    of the dependency order."""
    first = remaining[0]
    mask = {i: sum(1 << b for b in kops[i].allbits) for i in remaining}
    ok = {i: stageable(kops[i], m) for i in remaining}

    def scan(L: int) -> List[int]:
        blocked, inc, ng = 0, [], 0
        for i in remaining:
            b = mask[i]
            capped = grad_cap > 0 and kops[i].P > 0 and ng >= grad_cap and i != first
            if (b & blocked) or (b & ~L) or not ok[i] or capped:
                blocked |= b
                continue
            inc.append(i)
            ng += kops[i].P > 0
        return inc

    L = mask[first] | sum(1 << b for b in forced)
    if bin(L).count("1") > m:
        L = mask[first]
    while bin(L).count("1") < m:
        best, best_n = -1, -1
        for b in range(n):
            if not (L >> b) & 1:
                k = len(scan(L | (1 << b)))
                if k > best_n:
                    best, best_n = b, k
        if best < 0:
            break
        L |= 1 << best
    return scan(L), {b for b in range(n) if (L >> b) & 1}


def partition(kops: List[KOp], n: int, m: int, low_bits: int, grad_cap: int = 0,
              lookahead: bool = False) -> List[tuple]:
    """DAG partition into stages of <= m local bits (+ solo ops): greedy (a
    stage takes the bits of the ops in order while they fit) or lookahead
    (_lookahead_stage).  grad_cap > 0: at most that many ops with gradient
    parameters per stage (their adjoint arithmetic, not memory, bounds a
    stage)."""
    steps = []
    remaining = list(range(len(kops)))
    forced = set(range(min(low_bits, n)))
    while remaining:
        first = remaining[0]
        if not stageable(kops[first], m):
            steps.append(("solo", first))
            remaining = remaining[1:]
            continue
        if lookahead:
            inc, L = _lookahead_stage(kops, remaining, n, m, forced, grad_cap)
            p = 0
            while len(L) < m:
                if p not in L:
                    L.add(p)
                p += 1
            steps.append(("stage", (inc, sorted(L))))
            taken = set(inc)
            remaining = [i for i in remaining if i not in taken]
            continue
        # keep the low (coalescing) bits local unless the first op cannot fit with them
        L = set(forced) if len(set(forced) | kops[first].allbits) <= m else set()
        blocked = set()
        inc, skip = [], []
        n_grad = 0
        for idx in remaining:
            b = kops[idx].allbits
            capped = grad_cap > 0 and kops[idx].P > 0 and n_grad >= grad_cap and idx != first
            if b & blocked or not stageable(kops[idx], m) or capped:
                skip.append(idx)
                blocked |= b
                continue
            nl = L | b
            if len(nl) <= m:
                L = nl
                inc.append(idx)
                n_grad += kops[idx].P > 0
            else:
                skip.append(idx)
                blocked |= b
        if not inc:   # cannot happen (the first op always fits); guard against a stall
            steps.append(("solo", first))
            remaining = remaining[1:]
            continue
        # fill to m local bits with the lowest unused positions
        p = 0
        while len(L) < m:
            if p not in L:
                L.add(p)
            p += 1
        steps.append(("stage", (inc, sorted(L))))
        remaining = skip
    return steps


def reg_targets(k: KOp) -> Tuple[int, ...]:
    if k.kind == "mat1":
        return k.bits
    if k.kind in ("mat2", "swap", "matk"):
        return k.bits
    return ()


def order_for_registers(inc: List[int], kops: List[KOp], lbit: Dict[int, int], canon: List[int]) -> List[int]:
    """Reorder a stage's ops so that runs of them share a register set (fewer
    LDS remaps).  Ops that share a bit keep their relative order, so the
    product is unchanged; others commute.  Greedy list scheduling: run every
    ready op whose targets are register bits now; when none is, take the new
    register set from the earliest ready op's targets plus those of the other
    ready ops that fit beside them."""
    k = len(inc)
    bits = [kops[i].allbits for i in inc]
    need = [frozenset(lbit[b] for b in reg_targets(kops[i])) for i in inc]
    npred = [0] * k
    succ: List[List[int]] = [[] for _ in range(k)]
    for a in range(k):
        for b in range(a):
            if bits[a] & bits[b]:
                npred[a] += 1
                succ[b].append(a)
    ready = sorted(a for a in range(k) if npred[a] == 0)
    S = set(canon)
    order: List[int] = []
    while ready:
        fit = [a for a in ready if need[a] <= S]
        if not fit:
            first = ready[0]
            new = set(need[first])
            for a in ready[1:]:
                if len(new | need[a]) <= R_BITS:
                    new |= need[a]
            S = new
            continue
        a = fit[0]
        ready.remove(a)
        order.append(a)
        for b in succ[a]:
            npred[b] -= 1
            if npred[b] == 0:
                ready.append(b)
        ready.sort()
    assert len(order) == k
    return [inc[a] for a in order]


def encode_stage(inc: List[int], L: List[int], kops: List[KOp], n: int, data_off: Dict[int, tuple],
                 slot_of_kop: Dict[int, List[int]], stage_base: int,
                 mg_slot_of: Optional[Dict[int, List[int]]] = None, reorder: bool = False,
                 io_layouts: bool = False) -> Stage:
    """Program words of one stage (see tqsv_stage.hip for the format);
    reorder: schedule commuting ops for fewer remaps (order_for_registers);
    io_layouts: load / store in the layouts of the first / last ops where
    they keep accesses on whole lines (H_FLAGS masks)."""
    m = len(L)
    T = m - R_BITS
    lbit = {g: i for i, g in enumerate(L)}           # global position -> local bit
    canon = list(range(T, m))                        # register local bits, canonical
    nonlocal_pos = [p for p in range(n) if p not in lbit]

    header = np.zeros(H_POS + n, dtype=np.int32)
    header[H_M], header[H_T], header[H_NNL] = m, T, n - m
    header[H_POS:H_POS + T] = [L[i] for i in range(T)]
    header[H_POS + T:H_POS + m] = [L[i] for i in canon]
    header[H_POS + m:H_POS + n] = nonlocal_pos

    # register-set schedule with look-ahead
    if reorder:
        inc = order_for_registers(inc, kops, lbit, canon)
    needs = [tuple(lbit[b] for b in reg_targets(kops[i])) for i in inc]
    S = list(canon)
    seq = []   # ("remap", old, new) | ("op", i, S)
    for j, i in enumerate(inc):
        need = needs[j]
        if need and not set(need) <= set(S):
            new = list(dict.fromkeys(need))
            for later in needs[j + 1:]:
                for b in later:
                    if len(new) < R_BITS and b not in new:
                        new.append(b)
                if len(new) >= R_BITS:
                    break
            # fill from the current set, then from the high local bits: the
            # lowest local bits stay thread bits where possible, so the
            # stage can load and store in these layouts (below)
            for b in S + list(range(m - 1, -1, -1)):
                if len(new) >= R_BITS:
                    break
                if b not in new:
                    new.append(b)
            new = sorted(new)
            seq.append(("remap", list(S), new))
            S = new
        seq.append(("op", i, list(S)))
    # A layout whose register bits leave the lowest 4 local bits as thread
    # bits 0..3 reads and writes whole 128-byte lines like the canonical one,
    # so the stage loads in the layout of its first ops and stores in the
    # layout of its last instead of remapping to and from the canonical one.
    low = set(range(4))
    load_regs = store_regs = canon
    if io_layouts and T >= 4:
        if seq and seq[0][0] == "remap" and not set(seq[0][2]) & low:
            load_regs = seq[0][2]
            seq = seq[1:]
        if S != canon and not set(S) & low:
            store_regs = S
    if S != store_regs:
        seq.append(("remap", list(S), list(canon)))

    def layout_vals(regs):
        thr = [b for b in range(m) if b not in regs]
        return [1 << b for b in thr] + [1 << b for b in regs]

    gslots: List[int] = []
    ops_words = []
    remap_words = []
    enc = []
    feat = 0
    n_ops = len(seq)
    off_gslot = H_POS + n
    # gslot table size known after the loop; ops after it; remap records last
    for item in seq:
        w = [0] * 8
        if item[0] == "remap":
            rec = layout_vals(item[1]) + layout_vals(item[2])
            w[0] = OP_REMAP
            w[7] = len(remap_words)          # patched to an absolute offset below
            remap_words.extend(rec)
            ops_words.append(w)
            enc.append((-1, None))
            continue
        _, i, regs = item
        k = kops[i]
        thr = [b for b in range(m) if b not in regs]
        tpos = {b: t for t, b in enumerate(thr)}
        rpos = {b: r for r, b in enumerate(regs)}
        tm = tv = rm = rv = 0
        for b, v in zip(k.cbits, k.cvals):
            lb = lbit[b]
            if lb in rpos:
                rm |= 1 << rpos[lb]
                rv |= (v & 1) << rpos[lb]
            else:
                tm |= 1 << tpos[lb]
                tv |= (v & 1) << tpos[lb]
        off, bstride = data_off[i]
        P = k.P
        perm = None
        if k.kind == "mat1":
            w[0] = OP_MAT1
            w[3] = rpos[lbit[k.bits[0]]]
        elif k.kind in ("mat2", "swap"):
            w[0] = OP_MAT2
            ph, pl = rpos[lbit[k.bits[0]]], rpos[lbit[k.bits[1]]]
            if ph > pl:
                w[3] = ph | (pl << 4)
            else:
                w[3] = pl | (ph << 4)
                perm = _perm_index_bits(2, [1, 0])
        elif k.kind == "matk":
            w[0] = OP_MATK
            kb = len(k.bits)
            posl = [rpos[lbit[b]] for b in k.bits]         # bits[0] = gate MSB
            order = sorted(range(kb), key=lambda q: posl[q])
            # kernel index bit q <-> q-th ascending register position <-> gate bit (kb-1-order[q])
            perm = _perm_index_bits(kb, [kb - 1 - order[q] for q in range(kb)])
            if kb == 4:
                w[3] = 4
            else:
                used = set(posl)
                w[3] = [r for r in range(R_BITS) if r not in used][0]
        else:  # diag
            w[0] = OP_DIAG
            tpm = rpm = 0
            for b in k.bits:
                lb = lbit[b]
                if lb in rpos:
                    rpm |= 1 << rpos[lb]
                else:
                    tpm |= 1 << tpos[lb]
            w[3] = tpm
            rm |= 0
            w[6] = (rpm & 255) << 16
        w[0] |= (P << 4) | (F_BATCHED if k.batched else 0)
        mg = mg_slot_of is not None and i in mg_slot_of
        if k.kind == "mat1":
            if k.xperm and P == 0:
                w[0] |= F_XPERM
            elif tm == 0 and rm == 0:
                w[0] |= F_NOCTRL | (F_REAL if k.real else 0)
            if mg:
                w[0] |= F_MG
        elif k.kind == "diag" and k.sign and not k.bits and P == 0:
            w[0] |= F_SIGN
        if k.kind in ("mat2", "swap"):
            feat |= FEAT_MAT2
        elif k.kind == "matk":
            feat |= FEAT_MAT2 | FEAT_MATK
        w[1] = off
        w[2] = bstride
        w[4], w[5] = tm, tv
        w[6] |= (rm & 255) | ((rv & 255) << 8)
        if P:
            w[7] = len(gslots)
            gslots.extend(mg_slot_of[i] if mg else slot_of_kop[i])
        else:
            w[7] = -1
        ops_words.append(w)
        enc.append((i, perm))
    if len(gslots) > MAX_GP_STAGE:
        raise ValueError("stage exceeds the gradient-parameter limit")
    off_ops = off_gslot + len(gslots)
    off_remap = off_ops + 8 * len(ops_words)
    for w in ops_words:
        if w[0] == OP_REMAP:
            w[7] += off_remap
    def regmask(regs):
        return 0 if list(regs) == canon else sum(1 << b for b in regs)

    header[H_NOPS] = n_ops
    header[H_FLAGS] = feat | (regmask(load_regs) << FL_LOAD_SHIFT) | (regmask(store_regs) << FL_STORE_SHIFT)
    header[H_OFF_OPS] = off_ops
    header[H_NGP] = len(gslots)
    header[H_OFF_GSLOT] = off_gslot
    words = np.concatenate([header, np.array(gslots, dtype=np.int32),
                            np.array(ops_words, dtype=np.int32).reshape(-1),
                            np.array(remap_words, dtype=np.int32)]).astype(np.int32)
    return Stage(words, list(inc), m, enc)


# ----------------------------------------------------------------------------
# compiling
# ----------------------------------------------------------------------------
class Compiled:
    """A plan plus the per-op data layout inside the matrix pool."""

    def __init__(self, tape: Sequence[TapeOp], n: int, bsz: int, opts: EngineOptions, device):
        self.n, self.bsz, self.opts, self.device = n, bsz, opts, device
        self.param_offsets = []
        off = 0
        for op in tape:
            self.param_offsets.append(off)
            off += op.n_params()
        self.n_slots = off
        # tape ops with real parameters, and whether any of them is trainable
        # (both fixed by the structure key, which holds shapes and requires_grad)
        self.param_idx = [i for i, op in enumerate(tape) if op.params is not None and not op.is_matrix_param]
        self.any_param_grad = any(tape[i].params.requires_grad for i in self.param_idx)
        kops = lower(tape, n, bsz, self.param_offsets)
        if opts.mode == "fused" and opts.fusion_width >= 2:
            kops = fuse_blocks(kops, tape, n, opts.fusion_width)
            for k in kops:
                if k.extract == "block":
                    k.n_for_block = n
        elif opts.mode == "fused" and opts.merge_1q:
            kops = merge_single_wire(kops, tape)
        if opts.device_gates:
            mark_device_built(kops, tape, bsz)
        self.kops = kops
        # pool layout: kop data (fwd, with derivative matrices when trainable)
        self.data_off: Dict[int, tuple] = {}
        self.data_len: Dict[int, int] = {}
        pos = 0
        for i, k in enumerate(kops):
            per = (2 if k.kind == "diag" else k.dim() ** 2) * (1 + k.P)
            rows = bsz if k.batched else 1
            self.data_off[i] = (pos, per if k.batched else 0)
            self.data_len[i] = per * rows
            pos += per * rows
        # adjoint data (U^dagger) for non-trainable ops run by per-gate launches
        self.adj_off: Dict[int, tuple] = {}
        for i, k in enumerate(kops):
            if k.P == 0 and k.kind != "swap":
                per = 2 if k.kind == "diag" else k.dim() ** 2
                rows = bsz if k.batched else 1
                self.adj_off[i] = (pos, per if k.batched else 0)
                pos += per * rows
        self.pool_len = max(pos, 1)
        # ops whose matrices the GPU builds (device parameters): one DEV_OP
        # record each plus one DEV_MEMBER per member gate, appended to the
        # uploaded pool; parameter pointers and strides are filled per flush
        self.dev_kops = [i for i, k in enumerate(kops) if k.dev]
        self.dev_tape = frozenset(t for i in self.dev_kops for t in kops[i].src)
        self.dev_ops = np.zeros(len(self.dev_kops), dtype=DEV_OP)
        members = []
        self.dev_param_mem: List[int] = []      # member record index of each parameterised member
        self.dev_param_tape: List[int] = []     # its tape index
        deriv_ops = []
        for j, i in enumerate(self.dev_kops):
            k = kops[i]
            dop = self.dev_ops[j:j + 1]
            dop["dst"], dop["bstride"] = self.data_off[i]
            dop["adj"], dop["abstride"] = self.adj_off.get(i, (-1, 0))
            dop["rows"] = bsz if k.batched else 1
            dop["m0"], dop["nm"] = len(members), len(k.src)
            if k.kind == "diag":
                dop["flags"] = DGF_DIAG | ((k.diag_idx[0] & 1) << 8) | ((k.diag_idx[1] & 1) << 9)
            if k.P:
                deriv_ops.append(j)
            for t in k.src:
                op = tape[t]
                code = dev_member_code(op, bsz)
                rec = np.zeros(1, dtype=DEV_MEMBER)
                rec["code"] = code
                fl = (DGF_INVERSE if op.inverse else 0) | (DGF_CONJ if op.conj else 0)
                if code:
                    fl |= DGF_DERIV if op.wants_grad() else 0
                    self.dev_param_mem.append(len(members))
                    self.dev_param_tape.append(t)
                else:
                    u = G.FIXED_NP[op.name]
                    if op.inverse:
                        u = u.conj().T
                    if op.conj:
                        u = u.conj()
                    rec["fixed"][0] = np.stack([u.reshape(-1).real, u.reshape(-1).imag], -1).reshape(-1)
                    fl = 0
                rec["flags"] = fl
                members.append(rec)
        self.dev_members = np.concatenate(members) if members else np.zeros(0, dtype=DEV_MEMBER)
        self.dev_deriv_ops = np.array(deriv_ops, dtype=np.int64)
        self.dev_tail = (self.dev_ops.nbytes + self.dev_members.nbytes) // 8     # complex64 elements
        # every parameterised op built on the GPU: the host part of the pool
        # depends on the plan only and is built once
        self.host_static = all(k.dev or all(tape[t].params is None for t in k.src) for k in kops)
        self._static_host: Optional[np.ndarray] = None
        self.steps: List[tuple] = []
        self.perms: Dict[int, np.ndarray] = {}
        self._fill = None
        self._gv_layout: Optional[_GVLayout] = None
        self._jit: Dict[str, Optional[Dict[int, int]]] = {}
        self._jit_jobs: Dict[str, object] = {}
        self._prune: Optional[tuple] = None
        self._work_floats: Optional[int] = None
        self.mg_items = np.zeros(0, dtype=MG_ITEM)
        self.grad_width = max(self.n_slots, 1)
        if opts.mode == "fused" and n >= R_BITS:
            m = choose_stage_bits(n, opts, bsz)
            slot_of = {i: list(k.grad_slots) for i, k in enumerate(kops)}
            # trainable single-wire ops accumulate the four sums S = Im<lam|s|psi>
            # (s = I, X, Y, Z; past the parameter columns) instead of one
            # inner product per parameter; one small kernel folds them with
            # the coefficients of H_p = i dU/dtheta_p U^dagger after the sweep
            mg_slot_of = {}
            items = []
            for i, k in enumerate(kops):
                if opts.adjoint_mg and k.kind == "mat1" and k.P >= 1 and not k.xperm:
                    base = self.n_slots + MG_SUMS * len(items)
                    mg_slot_of[i] = list(range(base, base + MG_SUMS))
                    sl = list(k.grad_slots) + [0] * (4 - k.P)
                    items.append((base, k.P, sl, self.data_off[i][0], self.data_off[i][1]))
            self.mg_items = np.array(items, dtype=MG_ITEM) if items else np.zeros(0, dtype=MG_ITEM)
            self.grad_width = max(self.n_slots + MG_SUMS * len(items), 1)
            words = []
            base = 0
            for kind, payload in partition(kops, n, m, opts.low_bits, opts.grad_cap,
                                           use_lookahead(n, opts, bsz)):
                if kind == "solo":
                    self.steps.append(("solo", payload))
                    continue
                inc, L = payload
                st = encode_stage(inc, L, kops, n, self.data_off, slot_of, base, mg_slot_of,
                                  reorder=opts.reorder_ops, io_layouts=opts.io_layouts)
                for (ki, perm) in st.enc:
                    if ki >= 0 and perm is not None:
                        self.perms[ki] = perm
                self.steps.append(("stage", (base, st)))
                words.append(st.words)
                base += len(st.words)
            self.prog_host = np.concatenate(words).astype(np.int32) if words else np.zeros(1, np.int32)
            self.prog_dev = torch.from_numpy(self.prog_host.copy()).to(device)
        else:
            self.steps = [("solo", i) for i in range(len(kops))]
            self.prog_host = None
            self.prog_dev = None
        # Per-gate adjoint launches write P consecutive gradient columns.  A
        # trainable op whose slots are not consecutive (a fused block of
        # several gates) gets its own columns past the others; they are
        # added into its slots after the sweep.
        self.scatter: List[tuple] = []
        for kind, payload in self.steps:
            if kind == "solo":
                sl = list(kops[payload].grad_slots)
                if sl and sl != list(range(sl[0], sl[0] + len(sl))):
                    self.scatter.append((payload, self.grad_width, sl))
                    self.grad_width += len(sl)
        self.scatter_col = {i: col for i, col, _ in self.scatter}
        self.fwd_steps = self._forward_schedule()

    def _forward_schedule(self) -> List[tuple]:
        """The forward steps with each run of per-gate launches packed into one
        array of gate records for tqsv_run_gates: ("stage", payload) |
        ("gates", int64 [count, GATE_WORDS])."""
        out: List[tuple] = []
        run: List[np.ndarray] = []
        for kind, payload in self.steps:
            if kind == "solo":
                off, bs = self.data_off[payload]
                run.append(gate_record(self.kops[payload], off, bs))
                continue
            if run:
                out.append(("gates", np.stack(run)))
                run = []
            out.append((kind, payload))
        if run:
            out.append(("gates", np.stack(run)))
        return out

    def jit_kernels(self, kind: str) -> Optional[Dict[int, int]]:
        """Compiled stage kernels of this plan ({stage word offset: kernel
        handle}; kind "f" forward, "a" adjoint), or None when stages run on
        the interpreting kernels: option off, small state, nothing eligible,
        or -- with jit="auto" -- while the background compile is running.
        jit="on" compiles on the calling thread before returning."""
        if kind in self._jit:
            return self._jit[kind]
        mode = self.opts.jit
        if mode not in ("auto", "on", "off"):
            raise ValueError(f"EngineOptions.jit must be 'auto', 'on' or 'off', not {mode!r}")
        want = mode == "on" or (mode == "auto" and (self.bsz << self.n) >= JIT_MIN_AMPS)
        if not (want and torch.device(self.device).type == "cuda" and self.prog_host is not None):
            self._jit[kind] = None
            return None
        from . import stagejit
        if kind == "p":
            stop, skip = self.prune_info()
            stages = [(base, (st.words, skip.get(si, ()), si == stop)) for si, (kind_, (base, st)) in
                      [(i, (k, p)) for i, (k, p) in enumerate(self.steps) if k == "stage"]
                      if stagejit.supported(st.words) and (si in skip or si == stop)]
        else:
            stages = [(base, st.words) for kind_, (base, st) in
                      [(k, p) for k, p in self.steps if k == "stage"] if stagejit.supported(st.words)]
        if not stages:
            self._jit[kind] = None
            return None
        dev = torch.device(self.device)
        index = dev.index if dev.index is not None else torch.cuda.current_device()
        nt = os.environ.get("TQSV_JIT_NT") == "1" and (self.bsz << self.n) >= JIT_NT_AMPS
        if mode == "on":
            funcs = stagejit.get_kernels([w for _, w in stages], kind, index, nt)
        else:
            job = self._jit_jobs.get(kind)
            if job is None:
                self._jit_jobs[kind] = stagejit.submit([w for _, w in stages], kind, index, nt)
                return None
            if not job.done.is_set():
                return None
            funcs = job.result
            if job.error is not None:
                import warnings
                warnings.warn(f"run-time stage compile failed, staying on the interpreting kernels: {job.error}")
                funcs = None
        table = None if funcs is None else {base: f for (base, _), f in zip(stages, funcs)}
        self._jit[kind] = table
        return table

    def adjoint_work_floats(self) -> int:
        """Floats of the partial buffer the adjoint launches of this plan need
        (the largest over its stages and per-gate steps)."""
        if self._work_floats is None:
            lib = N.lib()
            need = 0
            if any(kind == "solo" and self.kops[p].P for kind, p in self.steps):
                need = int(lib.tqsv_adjoint_work_floats(self.bsz, self.n))
            for kind, payload in self.steps:
                if kind == "stage":
                    hp = self.prog_host[payload[0]:]
                    w = int(lib.tqsv_stage_adjoint_work_floats(hp.ctypes.data_as(N.c_int_p), self.n, self.bsz,
                                                               self.opts.grid_cap))
                    if w < 0:
                        raise RuntimeError("stage adjoint workspace query failed: "
                                           + lib.tqsv_last_error().decode(errors="replace"))
                    need = max(need, w)
            self._work_floats = need
        return self._work_floats

    def prune_info(self) -> tuple:
        """How the reverse sweep shrinks when no gradient with respect to the
        input state is wanted: (stop, skip).  stop: index in self.steps of
        the first step, in forward order, holding a trainable op -- the sweep
        ends with it (nothing earlier is run).  skip: {stage step index:
        program positions of ops whose un-apply is left out}.

        An op's un-apply is needed only when an earlier op (forward order)
        that is trainable or itself un-applied shares a wire with it.  The
        others touch wires that no such op uses: they commute with every
        generator and every un-applied op left in the sweep, so leaving them
        applied to both psi and lam changes none of the sums
        Im<lam|H_p|psi>.  Ops of per-gate steps and of stages without
        compiled kernels are always un-applied."""
        if self._prune is not None:
            return self._prune
        from . import stagejit
        used: set = set()          # wires of trainable or un-applied ops seen so far
        stop = None
        skip: Dict[int, tuple] = {}
        for si, (kind, payload) in enumerate(self.steps):
            if kind == "solo":
                k = self.kops[payload]
                if k.P and stop is None:
                    stop = si
                used |= k.allbits
                continue
            _, st = payload
            ok = stagejit.supported(st.words)
            left = []
            for pos, (ki, _) in enumerate(st.enc):
                if ki < 0:
                    continue
                k = self.kops[ki]
                if k.P and stop is None:
                    stop = si
                if ok and not (k.allbits & used):
                    left.append(pos)
                    if k.P:
                        used |= k.allbits
                else:
                    used |= k.allbits
            if left:
                skip[si] = tuple(left)
        self._prune = (len(self.steps) if stop is None else stop, skip)
        return self._prune

    # ---------------------------------------------------------------- values
    def pool(self, tape: Sequence[TapeOp], with_grad: bool, with_adj: bool) -> torch.Tensor:
        """The matrix pool on the device: staged through pinned host memory and
        copied asynchronously on the current stream (no wait for queued work)."""
        host = self.pool_host(tape, with_grad, with_adj)
        if torch.device(self.device).type != "cuda":
            return torch.from_numpy(host).to(self.device)
        pool = _upload(host, torch.device(self.device))
        if len(self.dev_ops):
            _check_dev_layout()
            base = pool.data_ptr() + 8 * self.pool_len
            N.call("tqsv_build_dev_gates", N.stream_of(pool), pool.data_ptr(), base, base + self.dev_ops.nbytes,
                   len(self.dev_ops), self.bsz)
        return pool

    def pool_host(self, tape: Sequence[TapeOp], with_grad: bool, with_adj: bool) -> np.ndarray:
        """The matrix pool (complex64) for the current parameter values."""
        tail = None
        if len(self.dev_ops):
            ops = self.dev_ops
            if with_grad and len(self.dev_deriv_ops):
                ops = ops.copy()
                ops["flags"][self.dev_deriv_ops] |= DGF_DERIV
            mem = self.dev_members.copy()
            ps = [tape[t].params for t in self.dev_param_tape]
            mem["params"][self.dev_param_mem] = [p.data_ptr() for p in ps]
            mem["rstride"][self.dev_param_mem] = [p.stride(0) if p.shape[0] > 1 else 0 for p in ps]
            mem["cstride"][self.dev_param_mem] = [p.stride(1) for p in ps]
            tail = (ops, mem)
        if self._static_host is not None:
            host = np.empty(self.pool_len + self.dev_tail, dtype=np.complex64)
            host[:self.pool_len] = self._static_host
            self._put_tail(host, tail)
            return host
        if self.host_static:
            with_adj = True          # built once for both uses
        if self._gv_layout is None:
            self._gv_layout = _GVLayout(tape, self.dev_tape)
        gv = GateValues(tape, with_grad, self._gv_layout)
        host = np.zeros(self.pool_len + self.dev_tail, dtype=np.complex64)
        self._put_tail(host, tail)
        if self._fill is None:
            self._fill = self._build_fill(tape)
        singles, chains, others = self._fill
        for key, (tidx, poff, aoff) in singles:
            gkey, extract, didx, perm, P = key
            rows = gv.grow[tidx]
            U = gv.UG[gkey][rows]
            if extract == "diag":
                a, b = didx
                data = np.stack([U[:, a, a], U[:, b, b]], axis=-1)                     # [G, 2]
                adj = np.conj(data)
                if with_grad and P:
                    D = gv.DG[gkey][rows]
                    data = np.concatenate([data, np.stack([D[:, :, a, a], D[:, :, b, b]], -1).reshape(len(tidx), -1)], 1)
            else:
                M = _block(U, extract)
                if perm is not None:
                    M = M[:, perm][:, :, perm]
                dd = M.shape[-1]
                data = M.reshape(len(tidx), dd * dd)
                adj = None
                if aoff is not None:
                    Mu = _block(U, extract)
                    adj = np.conj(np.swapaxes(Mu, -1, -2)).reshape(len(tidx), dd * dd)
                if with_grad and P:
                    D = _block(gv.DG[gkey][rows], extract)
                    if perm is not None:
                        D = D[:, :, perm][:, :, :, perm]
                    data = np.concatenate([data, D.reshape(len(tidx), -1)], 1)
            host[poff[:, None] + np.arange(data.shape[1])] = data
            if with_adj and aoff is not None:
                host[aoff[:, None] + np.arange(adj.shape[1])] = adj
        for key, (tidx, poff, aoff) in chains:
            gkeys, Ps = key
            G_, L = tidx.shape
            mats = [gv.UG[gkeys[m]][gv.grow[tidx[:, m]]] for m in range(L)]
            eye = np.broadcast_to(np.eye(2, dtype=np.complex128), (G_, 2, 2))
            pre = [eye]
            for u in mats:
                pre.append(np.matmul(u, pre[-1]))
            parts = [pre[-1].reshape(G_, 4)]
            if with_grad and sum(Ps):
                suf = [eye] * (L + 1)
                for m in range(L - 1, -1, -1):
                    suf[m] = np.matmul(suf[m + 1], mats[m])
                for m in range(L):
                    if Ps[m]:
                        D = gv.DG[gkeys[m]][gv.grow[tidx[:, m]]]
                        for q in range(Ps[m]):
                            parts.append(np.matmul(np.matmul(suf[m + 1], D[:, q]), pre[m]).reshape(G_, 4))
            data = np.concatenate(parts, axis=1)
            host[poff[:, None] + np.arange(data.shape[1])] = data
            if with_adj and aoff is not None:
                adj = np.conj(np.swapaxes(pre[-1], -1, -2)).reshape(G_, 4)
                host[aoff[:, None] + np.arange(4)] = adj
        for i in others:
            self._fill_one(i, tape, gv, host, with_grad, with_adj)
        if self.host_static:
            self._static_host = host[:self.pool_len].copy()
        return host

    def _put_tail(self, host: np.ndarray, tail):
        if tail is not None:
            ops, mem = tail
            raw = host[self.pool_len:].view(np.uint8)
            raw[:ops.nbytes] = ops.view(np.uint8)
            raw[ops.nbytes:ops.nbytes + mem.nbytes] = mem.view(np.uint8)

    def _build_fill(self, tape):
        """Group kernel ops by how their pool data is formed, once per plan."""
        singles: Dict[tuple, list] = {}
        chains: Dict[tuple, list] = {}
        others: List[int] = []
        gk = [None] * len(tape)
        for i, op in enumerate(tape):
            gk[i] = ("__mat__", i) if op.is_matrix_param else (op.name, len(op.wires), op.n_params() > 0)
        for i, k in enumerate(self.kops):
            if k.dev:
                continue                      # written on the GPU (tqsv_build_dev_gates)
            aoff = self.adj_off.get(i, (None, 0))[0]
            if k.batched or k.kind == "swap" or k.extract == "block" or any(tape[t].is_matrix_param for t in k.src):
                others.append(i)
                continue
            if k.extract == "chain":
                key = (tuple(gk[t] for t in k.src), tuple(len(_member_slots(tape[t])) for t in k.src))
                chains.setdefault(key, []).append((list(k.src), self.data_off[i][0], aoff))
                continue
            perm = self.perms.get(i)
            key = (gk[k.src[0]], k.extract, tuple(k.diag_idx), None if perm is None else tuple(perm), k.P)
            singles.setdefault(key, []).append((k.src[0], self.data_off[i][0], aoff))

        def pack(d):
            out = []
            for key, lst in d.items():
                tidx = np.array([x[0] for x in lst], dtype=np.int64)
                poff = np.array([x[1] for x in lst], dtype=np.int64)
                ao = [x[2] for x in lst]
                aoff = None if any(a is None for a in ao) else np.array(ao, dtype=np.int64)
                out.append((key, (tidx, poff, aoff)))
            return out

        singles_l = [((g, e, di, None if pm is None else np.array(pm), P), v)
                     for (g, e, di, pm, P), v in pack(singles)]
        return singles_l, pack(chains), others

    def _fill_one(self, i, tape, gv, host, with_grad, with_adj):
        k = self.kops[i]
        d = kop_data(k, gv, tape, with_grad and k.P > 0)       # [rows, 1+P, dd]
        perm = self.perms.get(i)
        if perm is not None:
            dd = int(round(math.sqrt(d.shape[-1])))
            d = d.reshape(d.shape[0], d.shape[1], dd, dd)[:, :, perm][:, :, :, perm]
            d = d.reshape(d.shape[0], d.shape[1], dd * dd)
        rows = self.bsz if k.batched else 1
        if not (with_grad and k.P > 0):
            d = d[:, :1]
        d = np.broadcast_to(d, (rows,) + d.shape[1:])
        off, _ = self.data_off[i]
        flat = d.reshape(rows, -1)
        per = self.data_len[i] // rows
        blk = host[off:off + per * rows].reshape(rows, per)
        blk[:, :flat.shape[1]] = flat
        if with_adj and i in self.adj_off:
            aoff, _ = self.adj_off[i]
            if k.kind == "diag":
                a = np.conj(d[:, 0])
            else:
                dd = k.dim()
                u = d[:, 0].reshape(rows, dd, dd)
                if perm is not None:
                    inv = np.argsort(perm)
                    u = u[:, inv][:, :, inv]
                a = np.conj(np.swapaxes(u, -1, -2)).reshape(rows, -1)
            host[aoff:aoff + a.size] = a.reshape(-1)


def _member_slots(op: TapeOp):
    return range(op.n_params()) if op.wants_grad() else range(0)


_cache_lock = threading.Lock()
_plan_cache: Dict[tuple, Compiled] = {}


def structure_key(tape: Sequence[TapeOp], n: int, bsz: int, opts: EngineOptions, device) -> tuple:
    items = []
    add = items.append
    for op in tape:
        p = op.params
        if p is None:
            ps = None
        elif p.is_complex():
            ps = ("m", p.shape)                 # torch.Size: a tuple (hash and equality)
        else:
            ps = (p.shape, p.requires_grad, p.is_cuda, p.dtype)
        add((op.name, op.wires, op.inverse, op.conj, ps))
    return (n, bsz, opts.key(), str(device), tuple(items))


def compile_tape(tape: Sequence[TapeOp], n: int, bsz: int, opts: EngineOptions, device) -> Compiled:
    key = structure_key(tape, n, bsz, opts, device)
    with _cache_lock:
        c = _plan_cache.get(key)
    if c is None:
        c = Compiled(tape, n, bsz, opts, device)
        with _cache_lock:
            if len(_plan_cache) > 256:
                _plan_cache.clear()
            _plan_cache[key] = c
    return c


# ----------------------------------------------------------------------------
# execution
# ----------------------------------------------------------------------------
GATE_WORDS = 16
GATE_KIND = {"mat1": 1, "mat2": 2, "swap": 3, "diag": 4, "matk": 5}


def gate_record(k: KOp, off: int, bstride: int) -> np.ndarray:
    """One tqsv_run_gates record: kind, k, bits[10], cmask, cval, matrix
    offset (complex elements into the pool), matrix row stride."""
    if len(k.bits) > 10:
        raise ValueError(f"dense gate on {len(k.bits)} wires: the per-gate kernels take at most 10")
    w = np.zeros(GATE_WORDS, dtype=np.int64)
    w[0] = GATE_KIND[k.kind]
    w[1] = len(k.bits)
    w[2:2 + len(k.bits)] = k.bits
    w[12] = k.cmask
    w[13] = k.cval
    w[14] = off
    w[15] = bstride
    return w


def _pool_ptr(pool: torch.Tensor, off: int) -> int:
    return pool.data_ptr() + 8 * off


def run_solo_forward(st: torch.Tensor, k: KOp, c: Compiled, pool: torch.Tensor, off: int, bstride: int,
                     stream: int):
    n, bsz = c.n, c.bsz
    p = _pool_ptr(pool, off)
    sp = st.data_ptr()
    if k.kind == "mat1":
        N.call("tqsv_apply_1q", stream, sp, bsz, n, k.bits[0], k.cmask, k.cval, p, bstride)
    elif k.kind == "mat2":
        N.call("tqsv_apply_2q", stream, sp, bsz, n, k.bits[0], k.bits[1], k.cmask, k.cval, p, bstride)
    elif k.kind == "swap":
        N.call("tqsv_apply_swap", stream, sp, bsz, n, k.bits[0], k.bits[1], k.cmask, k.cval)
    elif k.kind == "diag":
        pm = 0
        for b in k.bits:
            pm |= 1 << b
        N.call("tqsv_apply_diag", stream, sp, bsz, n, k.cmask, k.cval, pm, p, bstride)
    else:
        bits = N.int_array(k.bits)
        N.call("tqsv_apply_kq", stream, sp, bsz, n, len(k.bits), bits, k.cmask, k.cval, p, bstride)


def forward(c: Compiled, state_in: torch.Tensor, pool: torch.Tensor, init_zero: bool = False) -> torch.Tensor:
    """Run the plan out of place; returns the new state tensor ([bsz, 2^n])."""
    n, bsz = c.n, c.bsz
    out = torch.empty((bsz, 1 << n), dtype=torch.complex64, device=c.device)
    stream = N.stream_of(out)
    cur_in = state_in
    fresh = False           # out holds the current state
    if init_zero and (not c.steps or c.steps[0][0] != "stage"):
        N.call("tqsv_init_basis", stream, out.data_ptr(), bsz, n, 0)
        fresh = True
    jit = c.jit_kernels("f") if c.prog_host is not None else None
    for kind, payload in c.fwd_steps:
        if kind == "stage":
            base, stg = payload
            src = out if fresh else cur_in
            hp = c.prog_host[base:]
            f = jit.get(base) if jit else None
            tok = _timed("f", f is not None)
            if f is not None:
                N.call("tqsv_jit_stage_forward", stream, f, out.data_ptr(),
                       (src.data_ptr() if src is not None else None), bsz, n, hp.ctypes.data_as(N.c_int_p),
                       pool.data_ptr(), 1 if (init_zero and not fresh) else 0, c.opts.grid_cap)
            else:
                N.call("tqsv_stage_forward", stream, out.data_ptr(), (src.data_ptr() if src is not None else None),
                       bsz, n, hp.ctypes.data_as(N.c_int_p), c.prog_dev.data_ptr() + 4 * base, pool.data_ptr(),
                       1 if (init_zero and not fresh) else 0, c.opts.grid_cap)
            _timed_end(tok)
            fresh = True
        else:
            if not fresh:
                if init_zero:
                    N.call("tqsv_init_basis", stream, out.data_ptr(), bsz, n, 0)
                else:
                    out.copy_(cur_in.reshape(bsz, -1))
                fresh = True
            N.call("tqsv_run_gates", stream, out.data_ptr(), bsz, n, payload.ctypes.data_as(N.c_i64_p),
                   len(payload), pool.data_ptr())
    if not fresh:
        if init_zero:
            N.call("tqsv_init_basis", stream, out.data_ptr(), bsz, n, 0)
        else:
            out.copy_(cur_in.reshape(bsz, -1))
    return out


def adjoint(c: Compiled, psi_src: torch.Tensor, lam_src: torch.Tensor, pool: torch.Tensor,
            need_input_grad: bool = True):
    """Reverse sweep from psi_src (forward output) and lam_src (dL/dpsi),
    which are left unchanged.  Returns (grads [bsz, n_slots] float32, lam
    after the sweep = dL/d(input state), or None without need_input_grad).
    When the sweep starts with a compiled stage, that stage reads psi_src /
    lam_src and writes fresh buffers; otherwise they are copied first.
    Without need_input_grad the sweep stops at the first trainable step and,
    once their kernels are compiled, runs the pruned stages of
    Compiled.prune_info."""
    n, bsz = c.n, c.bsz
    grad = torch.zeros((bsz, c.grad_width), dtype=torch.float32, device=c.device)
    stream = N.stream_of(psi_src)
    gstride = grad.shape[1]
    # per-workgroup gradient partials of one launch (added into grad in a
    # fixed order by the library right after it; every launch reuses them)
    work = torch.empty(max(c.adjoint_work_floats(), 1), dtype=torch.float32, device=c.device)
    wptr, wlen = work.data_ptr(), work.numel()
    jit = c.jit_kernels("a") if c.prog_host is not None else None
    steps = list(reversed(c.steps))
    pj = None
    stop = 0
    if not need_input_grad and c.prog_host is not None:
        stop, _ = c.prune_info()
        steps = steps[:len(c.steps) - stop]
        pj = c.jit_kernels("p")          # (queued behind the full kernels' compile when it runs in the background)
        if not jit:
            pj = None
    last_base = c.steps[stop][1][0] if stop < len(c.steps) and c.steps[stop][0] == "stage" else None
    first_src = None
    if steps and steps[0][0] == "stage" and jit and jit.get(steps[0][1][0]) is not None:
        psi, lam = torch.empty_like(psi_src), torch.empty_like(lam_src)
        first_src = (psi_src.data_ptr(), lam_src.data_ptr())
    else:
        psi, lam = psi_src.clone(), lam_src.clone()
    for kind, payload in steps:
        if kind == "stage":
            base, stg = payload
            hp = c.prog_host[base:]
            f = jit.get(base) if jit else None
            tag = "a"
            if pj and base in pj:
                f = pj[base]
                # pruned kernels: fewer un-applies; the sweep's last stage stores nothing
                tag = "a_read" if base == last_base else "a_pruned"
            tok = _timed(tag, f is not None)
            if f is not None:
                pin, lin = first_src if first_src is not None else (None, None)
                first_src = None
                N.call("tqsv_jit_stage_adjoint", stream, f, psi.data_ptr(), lam.data_ptr(), pin, lin, bsz, n,
                       hp.ctypes.data_as(N.c_int_p), c.prog_dev.data_ptr() + 4 * base, pool.data_ptr(),
                       N.fptr(grad), gstride, wptr, wlen, c.opts.grid_cap)
            else:
                N.call("tqsv_stage_adjoint", stream, psi.data_ptr(), lam.data_ptr(), bsz, n,
                       hp.ctypes.data_as(N.c_int_p), c.prog_dev.data_ptr() + 4 * base, pool.data_ptr(),
                       N.fptr(grad), gstride, wptr, wlen, c.opts.grid_cap)
            _timed_end(tok)
            continue
        i = payload
        k = c.kops[i]
        if k.P > 0:
            off, bs = c.data_off[i]
            p = _pool_ptr(pool, off)
            g0 = c.scatter_col.get(i, k.grad_slots[0])
            if k.kind == "mat1":
                N.call("tqsv_adjoint_1q", stream, psi.data_ptr(), lam.data_ptr(), bsz, n, k.bits[0], k.cmask,
                       k.cval, p, bs, k.P, N.fptr(grad), gstride, g0, wptr, wlen)
            elif k.kind == "mat2":
                N.call("tqsv_adjoint_2q", stream, psi.data_ptr(), lam.data_ptr(), bsz, n, k.bits[0], k.bits[1],
                       k.cmask, k.cval, p, bs, k.P, N.fptr(grad), gstride, g0, wptr, wlen)
            elif k.kind == "diag":
                pm = 0
                for b in k.bits:
                    pm |= 1 << b
                N.call("tqsv_adjoint_diag", stream, psi.data_ptr(), lam.data_ptr(), bsz, n, k.cmask, k.cval, pm,
                       p, bs, k.P, N.fptr(grad), gstride, g0, wptr, wlen)
            else:
                raise NotImplementedError(f"adjoint of trainable '{k.kind}' op")
        else:
            if k.kind == "swap":
                off, bs = 0, 0
            else:
                off, bs = c.adj_off[i]
            run_solo_forward(psi, k, c, pool, off, bs, stream)
            run_solo_forward(lam, k, c, pool, off, bs, stream)
    if len(c.mg_items):
        if getattr(c, "_mg_dev", None) is None or c._mg_dev.device != psi.device:
            c._mg_dev = torch.from_numpy(c.mg_items.view(np.uint8).copy()).to(psi.device)
        N.call("tqsv_mg_to_grad", stream, N.fptr(grad), gstride, bsz, c._mg_dev.data_ptr(), len(c.mg_items),
               pool.data_ptr())
    for i, col, sl in c.scatter:
        idx = torch.as_tensor(sl, dtype=torch.long, device=grad.device)
        grad.index_add_(1, idx, grad[:, col:col + len(sl)].clone())
    return grad[:, :max(c.n_slots, 1)], (lam if need_input_grad else None)
