"""Plan selection: exhaustive autotuning with a persistent JSON database, and
the learned planner that replaces it once the cost model is trained.

Candidates are EngineOptions (fusion width k = dense-block width, stage bits
m = local bits per stage).  ``Autotuner.tune`` compiles and times every
candidate on the GPU (HIP events, median of repeats), records every
(plan features, time) pair -- the cost model's training data -- and keeps the
winner per circuit structure.  ``LearnedPlanner.choose`` scores candidates
with the MLP from host-side plan features only and runs nothing.
"""
from __future__ import annotations

import hashlib
import json
import os
import threading
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import engine as E
from .costmodel import MLPCostModel, plan_features

DEFAULT_CANDIDATES = ([E.EngineOptions(fusion_width=k, stage_bits=m) for k in (1, 2, 3, 4) for m in (10, 12, 13)]
                      + [E.EngineOptions(fusion_width=k, stage_bits=12) for k in (5, 6)])
# Plans of circuits run by CircuitBatch: every op inside stages (fusion width
# <= 4, stage width <= 12) and the program-interpreting stage code, which is
# what the batch kernel executes; timed with the interpreting stage kernels.
BATCH_CANDIDATES = [E.EngineOptions(fusion_width=k, stage_bits=12, jit="off") for k in (1, 2, 3, 4)]


def circuit_key(tape: Sequence[E.TapeOp], n: int, bsz: int) -> str:
    h = hashlib.sha1()
    h.update(f"{n}:{bsz}".encode())
    for op in tape:
        shape = None if op.params is None else tuple(op.params.shape)
        h.update(f"{op.name}{op.wires}{op.inverse}{op.conj}{shape};".encode())
    return h.hexdigest()


def _opts_dict(o: E.EngineOptions) -> dict:
    return dict(o.__dict__)


def time_plan(tape, n, bsz, opts: E.EngineOptions, device="cuda", reps: int = 5) -> float:
    """Median ms of one forward of the compiled plan (state from |0..0>),
    with the kernels the plan runs in steady state: when its stages compile
    at run time, the timing waits for them."""
    c = E.Compiled(tape, n, bsz, opts, torch.device(device))
    pool = c.pool(tape, with_grad=False, with_adj=False)
    E.forward(c, None, pool, init_zero=True)          # warm-up (starts any stage compile)
    if c._jit_jobs:
        E.wait_for_compiled_stages()
        E.forward(c, None, pool, init_zero=True)
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        E.forward(c, None, pool, init_zero=True)
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


class Autotuner:
    def __init__(self, db_path: Optional[str] = None):
        self.db_path = db_path
        self.db: Dict[str, dict] = {}
        self._lock = threading.Lock()
        if db_path and os.path.exists(db_path):
            with open(db_path) as f:
                self.db = json.load(f)

    def save(self):
        if self.db_path:
            with self._lock, open(self.db_path, "w") as f:
                json.dump(self.db, f)

    def tune(self, tape, n, bsz, device="cuda", candidates: Sequence[E.EngineOptions] = DEFAULT_CANDIDATES,
             reps: int = 5) -> E.EngineOptions:
        key = circuit_key(tape, n, bsz)
        if key in self.db:
            return E.EngineOptions(**self.db[key]["best"])
        recs = []
        for o in candidates:
            o = E.EngineOptions(**{**o.__dict__, "stage_bits": min(o.stage_bits or 12, n)})
            if any(r["opts"] == _opts_dict(o) for r in recs):
                continue
            c = E.Compiled(tape, n, bsz, o, "cpu")
            ms = time_plan(tape, n, bsz, o, device, reps)
            recs.append({"opts": _opts_dict(o), "ms": ms, "features": plan_features(c).tolist()})
        best = min(recs, key=lambda r: r["ms"])
        with self._lock:
            self.db[key] = {"best": best["opts"], "records": recs}
        return E.EngineOptions(**best["opts"])

    def dataset(self, keys: Optional[Sequence[str]] = None):
        X, y = [], []
        for k, v in self.db.items():
            if keys is not None and k not in keys:
                continue
            for r in v["records"]:
                X.append(r["features"])
                y.append(r["ms"])
        return np.array(X), np.array(y)


class LearnedPlanner:
    """Pick the candidate with the smallest predicted time (no launches)."""

    def __init__(self, model: MLPCostModel, candidates: Sequence[E.EngineOptions] = DEFAULT_CANDIDATES):
        self.model = model
        self.candidates = list(candidates)

    def score(self, tape, n, bsz):
        opts, feats = [], []
        for o in self.candidates:
            o = E.EngineOptions(**{**o.__dict__, "stage_bits": min(o.stage_bits or 12, n)})
            if any(_opts_dict(o) == _opts_dict(p) for p in opts):
                continue
            c = E.Compiled(tape, n, bsz, o, "cpu")
            opts.append(o)
            feats.append(plan_features(c))
        pred = self.model.predict_ms(np.array(feats))
        return opts, pred

    def choose(self, tape, n, bsz) -> E.EngineOptions:
        opts, pred = self.score(tape, n, bsz)
        return opts[int(np.argmin(pred))]

    def choose_from_records(self, records: Sequence[dict]) -> E.EngineOptions:
        """The candidate with the smallest predicted time among an autotuner
        entry's records, from their stored plan features (no plan rebuilt)."""
        pred = self.model.predict_ms(np.array([r["features"] for r in records]))
        return E.EngineOptions(**records[int(np.argmin(pred))]["opts"])


def tune_circuits(tuner: "Autotuner", tapes: Sequence[tuple], candidates: Sequence[E.EngineOptions],
                  reps: int = 3) -> List[E.EngineOptions]:
    """Exhaustive autotuning of many (tape, n) circuits at bsz 1; returns the
    winner per circuit (the tuner keeps every timing as training data)."""
    return [tuner.tune(tape, n, 1, candidates=candidates, reps=reps) for tape, n in tapes]
