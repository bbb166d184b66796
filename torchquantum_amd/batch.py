"""Many independent circuits of uneven width and depth in one launch.

Each circuit (bsz = 1) is compiled by the fused planner into stages, and
every stage into (circuit, stage, group range) work items.  By default
(``mode="queue"``) all items of all circuits run in ONE persistent launch
(``tqsv_batch_queue``): listed level by level, taken in order from a device
counter, each waiting only for its own circuit's previous stage, so a
circuit's next stage starts as soon as its last one is done rather than when
the slowest circuit's is.  ``mode="levels"`` launches one
``tqsv_batch_forward`` per level instead (max stages over circuits launches,
each ending when its slowest item does).  Both end with one reduction.  All
device buffers (states, programs, matrix pool, work lists, queue counters)
are built once, so the launch sequence can be captured in a HIP graph
(``capture`` / ``replay``).

The reference runs such a batch as one Python loop per circuit, each gate a
separate permute + bmm (torchquantum/functional/gate_wrapper.py:96-140).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from . import engine as E
from . import gates as G

BATCH_ITEM = np.dtype([("state_off", "<i8"), ("prog_off", "<i4"), ("n", "<i4"), ("init_zero", "<i4"),
                       ("pad", "<i4"), ("g0", "<i8"), ("g1", "<i8")])
QUEUE_ITEM = np.dtype([("b", BATCH_ITEM), ("circ", "<i4"), ("need", "<i4")])
ZCIRC = np.dtype([("state_off", "<i8"), ("n", "<i4"), ("zmask", "<u4"), ("part_off", "<i8")])
ZCHUNK_LG = 14
MAX_M = 12
SPIN_LIMIT = 1 << 20      # polls of a waiting workgroup (each a short sleep): about a second, far above any stage


def ops_to_tape(ops: Sequence[dict]) -> List[E.TapeOp]:
    tape = []
    for op in ops:
        p = op.get("params")
        s = G.spec(op["name"])
        if p is not None:
            p = torch.as_tensor(np.asarray(p, dtype=np.float64)).reshape(1, -1) if s.num_params >= 0 else \
                torch.as_tensor(np.asarray(p)).to(torch.complex64).unsqueeze(0)
        tape.append(E.TapeOp(G.canonical(op["name"]), tuple(int(w) for w in op["wires"]), p,
                             bool(op.get("inverse", False))))
    return tape


class CircuitBatch:
    def __init__(self, circuits: Sequence[tuple], device="cuda",
                 options: Optional[Sequence[E.EngineOptions]] = None, groups_per_item: Optional[int] = None,
                 zwire: int = 0, mode: str = "queue"):
        """circuits: [(n_wires, ops)], ops as {name, wires, params, inverse} dicts.
        options: per-circuit EngineOptions (mode is forced to 'fused').
        mode: "queue" (one persistent launch) or "levels" (a launch per level).
        groups_per_item: groups of 2^m amplitudes per work item (default 8
        with the queue, 4 with a launch per level: the measured best of
        2..32 on the 256-circuit workload, 21.6 and 24.8 ms)."""
        lib = N.lib()
        if (lib.tqsv_batch_item_bytes() != BATCH_ITEM.itemsize or lib.tqsv_batch_zcirc_bytes() != ZCIRC.itemsize
                or lib.tqsv_batch_queue_item_bytes() != QUEUE_ITEM.itemsize):
            raise RuntimeError("BatchItem / QueueItem / ZCirc layout mismatch with libtqsv.so")
        if mode not in ("queue", "levels"):
            raise ValueError(f"unknown batch mode {mode!r}")
        self.mode = mode
        if groups_per_item is None:
            groups_per_item = 8 if mode == "queue" else 4
        if groups_per_item < 1:
            raise ValueError("groups_per_item must be positive")
        self.device = torch.device(device)
        self.n_circ = len(circuits)
        self.ns = [int(n) for n, _ in circuits]
        self.compiled = []
        pools, progs = [], []
        pool_base, prog_base = 0, 0
        self.stage_offsets: List[List[int]] = []
        self.stage_m: List[List[int]] = []
        for i, (n, ops) in enumerate(circuits):
            if n < E.R_BITS:
                raise ValueError("batched circuits need >= 4 qubits")
            o = options[i] if options is not None else E.EngineOptions()
            o = E.EngineOptions(**{**o.__dict__, "mode": "fused",
                                   "stage_bits": min(o.stage_bits or MAX_M, MAX_M, n)})
            tape = ops_to_tape(ops)
            c = E.Compiled(tape, n, 1, o, "cpu")
            if any(kind != "stage" for kind, _ in c.steps):
                raise ValueError(f"circuit {i}: an op does not fit a stage (batched circuits must be "
                                 f"fully stageable)")
            pool = c.pool_host(tape, False, False)
            offs, ms = [], []
            for _, (base, st) in c.steps:
                words = st.words.copy()
                # relocate matrix offsets into the concatenated pool
                h = words
                off_ops = int(h[E.H_OFF_OPS])
                for q in range(int(h[E.H_NOPS])):
                    w = words[off_ops + 8 * q: off_ops + 8 * q + 8]
                    if (w[0] & 15) != E.OP_REMAP:
                        w[1] += pool_base
                offs.append(prog_base)
                ms.append(st.m)
                progs.append(words)
                prog_base += len(words)
            pools.append(pool)
            pool_base += len(pool)
            self.compiled.append(c)
            self.stage_offsets.append(offs)
            self.stage_m.append(ms)
        self.max_m = max(max(m) for m in self.stage_m)
        self.state_off = np.zeros(self.n_circ, dtype=np.int64)
        tot = 0
        for i, n in enumerate(self.ns):
            self.state_off[i] = tot
            tot += 1 << n
        self.total_amps = tot
        self.states = torch.zeros(tot, dtype=torch.complex64, device=self.device)
        self.progs = torch.from_numpy(np.concatenate(progs).astype(np.int32)).to(self.device)
        self.mats = torch.from_numpy(np.concatenate(pools).astype(np.complex64)).to(self.device)
        # work lists per level, and the same items as one level-major queue
        n_levels = max(len(o) for o in self.stage_offsets)
        self.levels = []
        queue = []
        seen = [0] * self.n_circ          # items of each circuit in earlier levels
        for L in range(n_levels):
            items = []
            per_circ = [0] * self.n_circ
            for i in range(self.n_circ):
                if L >= len(self.stage_offsets[i]):
                    continue
                m = self.stage_m[i][L]
                ng = 1 << (self.ns[i] - m)
                for g0 in range(0, ng, groups_per_item):
                    b = (self.state_off[i], self.stage_offsets[i][L], self.ns[i], 1 if L == 0 else 0, 0,
                         g0, min(g0 + groups_per_item, ng))
                    items.append(b)
                    queue.append((b, i, seen[i]))
                    per_circ[i] += 1
            for i in range(self.n_circ):
                seen[i] += per_circ[i]
            arr = np.array(items, dtype=BATCH_ITEM)
            self.levels.append((torch.from_numpy(arr.view(np.uint8).copy()).to(self.device), len(items)))
        qa = np.array(queue, dtype=QUEUE_ITEM)
        self.queue = torch.from_numpy(qa.view(np.uint8).copy()).to(self.device)
        self.n_queue = len(queue)
        self.ctrl = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.done = torch.zeros(self.n_circ, dtype=torch.int32, device=self.device)
        self.spin_limit = SPIN_LIMIT
        # <Z_zwire> per circuit
        zc = np.zeros(self.n_circ, dtype=ZCIRC)
        chunk_circ = []
        off = 0
        for i, n in enumerate(self.ns):
            nch = max(1, ((1 << n) + (1 << ZCHUNK_LG) - 1) >> ZCHUNK_LG)
            zc[i] = (self.state_off[i], n, 1 << (n - 1 - zwire), off)
            chunk_circ.extend([i] * nch)
            off += nch
        self.zcirc = torch.from_numpy(zc.view(np.uint8).copy()).to(self.device)
        self.chunk_circ = torch.tensor(chunk_circ, dtype=torch.int64, device=self.device)
        self.n_chunks = len(chunk_circ)
        self.part = torch.empty(self.n_chunks, dtype=torch.float32, device=self.device)
        self.out = torch.empty(self.n_circ, dtype=torch.float32, device=self.device)
        self.graph = None

    @property
    def n_launches(self) -> int:
        return (2 if self.mode == "queue" else len(self.levels)) + 2

    def launch(self):
        """Enqueue every stage of every circuit and the <Z> read-out on the
        current stream."""
        stream = N.stream_of(self.states)
        if self.mode == "queue":
            N.call("tqsv_batch_queue", stream, self.states.data_ptr(), self.queue.data_ptr(), self.n_queue,
                   self.progs.data_ptr(), self.mats.data_ptr(), self.max_m, self.ctrl.data_ptr(),
                   self.done.data_ptr(), self.n_circ, self.spin_limit, 0)
        else:
            for items, cnt in self.levels:
                N.call("tqsv_batch_forward", stream, self.states.data_ptr(), items.data_ptr(), cnt,
                       self.progs.data_ptr(), self.mats.data_ptr(), self.max_m, 0)
        N.call("tqsv_batch_zexpval", stream, self.states.data_ptr(), self.zcirc.data_ptr(), self.n_circ,
               self.chunk_circ.data_ptr(), self.n_chunks, N.fptr(self.part), N.fptr(self.out))
        return self.out

    def run(self) -> torch.Tensor:
        if self.graph is not None:
            self.graph.replay()
            return self.out
        return self.launch()

    def capture(self):
        """Record the launch sequence in a HIP graph; later ``run`` replays it."""
        self.launch()                      # warm-up outside capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        return g

    def check(self):
        """Raise if a wait of the last queue launch ran out (synchronizes)."""
        if self.mode == "queue" and int(self.ctrl[1].item()) != 0:
            raise RuntimeError("batch queue: a workgroup's wait for its circuit's previous stage ran out; "
                               "the batch results are not valid")

    def state(self, i: int) -> torch.Tensor:
        self.check()
        o = int(self.state_off[i])
        return self.states[o:o + (1 << self.ns[i])]


def random_circuits(n_circ=256, n_range=(8, 22), depth_range=(4, 40), seed=0,
                    gate_set=("hadamard", "rx", "ry", "rz", "u3", "cnot", "cz", "swap")):
    """The uneven-batch workload: n ~ U{8..22}, depth ~ U{4..40} layers, each layer
    n gates drawn uniformly from gate_set on distinct random wires, angles U[0, 2pi)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_circ):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        d = int(rng.integers(depth_range[0], depth_range[1] + 1))
        ops = []
        for _ in range(d):
            for _ in range(n):
                nm = gate_set[rng.integers(len(gate_set))]
                s = G.spec(nm)
                w = [int(x) for x in rng.choice(n, s.num_wires, replace=False)]
                p = None if s.num_params == 0 else rng.uniform(0, 2 * math.pi, s.num_params)
                ops.append(dict(name=nm, wires=w, params=p, inverse=False))
        out.append((n, ops))
    return out
