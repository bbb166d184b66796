"""GPU: the learned plan cost model.

* its HIP forward/backward/Adam kernels against the float64 numpy oracle
  (oracle/mlp.py), tolerance 1e-4 relative;
* trained on autotuner timings, the learned planner's choices on held-out
  circuits are compared with (a) the default plan without the model and
  (b) exhaustive autotuning, whose measured times are the reference.
"""
import json
import os
import time

import numpy as np
import pytest
import torch

from oracle import mlp as OM

gpu = pytest.mark.gpu


@gpu
def test_mlp_kernels_vs_oracle():
    from torchquantum_amd.costmodel import MLPCostModel
    rng = np.random.default_rng(0)
    F, M = 11, 37
    model = MLPCostModel(n_features=F, hidden=(19, 13), device="cuda", seed=3)
    X = rng.normal(size=(M, F))
    y = rng.normal(size=M)
    Xt = torch.tensor(X, dtype=torch.float32, device="cuda")
    yt = torch.tensor(y, dtype=torch.float32, device="cuda")
    Ws = [w.double().cpu().numpy() for w in model.W]
    bs = [b.double().cpu().numpy() + 0.1 * rng.normal(size=b.shape) for b in model.b]
    for b, bn in zip(model.b, bs):
        b.copy_(torch.tensor(bn, dtype=torch.float32))
    pred, _ = model.forward(Xt)
    pref, _ = OM.forward(Ws, bs, X)
    assert np.allclose(pred.cpu().numpy(), pref, rtol=1e-4, atol=1e-4)
    loss, gW, gb = model.grads(Xt, yt)
    lo, gWo, gbo = OM.loss_and_grads(Ws, bs, X, y)
    assert abs(float(loss) - lo) < 1e-4 * (1 + abs(lo))
    for a, b in zip(gW + gb, gWo + gbo):
        assert np.allclose(a.cpu().numpy(), b, rtol=1e-3, atol=1e-5)
    # one Adam step
    ps = [p.double().cpu().numpy().copy() for p in model.W + model.b]
    ms = [np.zeros_like(p) for p in ps]
    vs = [np.zeros_like(p) for p in ps]
    model.adam_step(gW, gb, lr=1e-2)
    for p, g, m, v in zip(ps, gWo + gbo, ms, vs):
        OM.adam(p, g, m, v, 1, 1e-2)
    for a, b in zip(model.W + model.b, ps):
        assert np.allclose(a.cpu().numpy(), b, rtol=1e-4, atol=1e-5)


@gpu
def test_learned_planner_gain():
    """The learned planner against the method without it.  Plans are fusion
    widths 1..4 on the program-interpreting stage code (what the uneven batch
    runs); every candidate of every circuit is timed (exhaustive autotuning),
    the MLP is trained by this package's kernels on the timings of 24
    circuits of 12..20 qubits and then chooses, from plan features alone,
    for 8 held-out circuits of 20..23 qubits (larger than any it saw).  The
    learned choices must be faster in sum than the default plan (width 1)
    and within 15% of the exhaustive optimum."""
    from torchquantum_amd import batch as B
    from torchquantum_amd.autotune import Autotuner, LearnedPlanner, BATCH_CANDIDATES, circuit_key, tune_circuits
    from torchquantum_amd.costmodel import MLPCostModel

    train_c = B.random_circuits(n_circ=24, n_range=(12, 20), depth_range=(4, 24), seed=42)
    test_c = B.random_circuits(n_circ=8, n_range=(20, 23), depth_range=(4, 24), seed=43)
    assert max(n for n, _ in test_c) >= 22
    tapes = [(B.ops_to_tape(ops), n) for n, ops in train_c + test_c]
    tuner = Autotuner()
    t0 = time.perf_counter()
    tune_circuits(tuner, tapes, BATCH_CANDIDATES, reps=3)
    t_exhaustive = (time.perf_counter() - t0) / len(tapes)
    keys = [circuit_key(tape, n, 1) for tape, n in tapes]
    X, y = tuner.dataset(keys[:24])
    model = MLPCostModel(device="cuda", seed=0)
    hist = model.fit(X, y, epochs=1500, lr=3e-3)
    assert hist[-1] < hist[0]
    planner = LearnedPlanner(model, BATCH_CANDIDATES)
    tot_default = tot_learned = tot_best = 0.0
    t_learned = 0.0
    rel_err = []
    for (tape, n), key in zip(tapes[24:], keys[24:]):
        recs = tuner.db[key]["records"]
        by = {json.dumps(r["opts"], sort_keys=True): r["ms"] for r in recs}
        t1 = time.perf_counter()
        pick = planner.choose(tape, n, 1)
        t_learned += time.perf_counter() - t1
        default = dict(BATCH_CANDIDATES[0].__dict__, stage_bits=min(12, n))
        tot_default += by[json.dumps(default, sort_keys=True)]
        tot_learned += by[json.dumps(pick.__dict__, sort_keys=True)]
        tot_best += min(r["ms"] for r in recs)
        opts, pred = planner.score(tape, n, 1)
        for o, p in zip(opts, pred):
            rel_err.append(abs(p - by[json.dumps(o.__dict__, sort_keys=True)]) / by[json.dumps(o.__dict__, sort_keys=True)])
    t_learned /= len(test_c)
    report = {"held_out_circuits": len(test_c), "held_out_qubits": [n for n, _ in test_c],
              "sum_ms_default_plan": tot_default, "sum_ms_learned_plan": tot_learned,
              "sum_ms_exhaustive_best": tot_best, "gain_vs_default": tot_default / tot_learned,
              "regret_vs_exhaustive": tot_learned / tot_best - 1,
              "median_rel_prediction_error": float(np.median(rel_err)),
              "selection_s_learned": t_learned, "selection_s_exhaustive": t_exhaustive,
              "train_loss_first_last": [hist[0], hist[-1]]}
    out_dir = os.environ.get("TQSV_REPORT_DIR")
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "costmodel_report.json"), "w") as f:
            json.dump(report, f, indent=1)
    print(report)
    assert tot_learned < tot_default, report
    assert tot_learned <= 1.15 * tot_best, report
    assert t_learned < t_exhaustive
