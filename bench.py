"""Benchmark: amplitude updates/s & adjoint grads/s vs the plain per-gate kernel,
and % of HBM peak, on one MI355X.

Headline (``value``): adjoint gradient evaluations/s of the fused engine on
24 qubits, bsz 64, complex64: 10 x [RX, RY, RZ on every wire (angles
N(0, 0.1)), CNOT ring], a per-row RY input layer, loss = mean_b [sum_i
<Z_i Z_i+1> + 0.5 sum_i <X_i>], d loss / d theta for all 720 parameters.
One step = one forward pass + the loss + the adjoint backward.

Fields of the same JSON line:
  * roofline: the dominant kernel of that step, the compiled adjoint stage
    (HIP events around every launch inside the timed loop); the compiled
    forward stage and the dense 1-qubit gate on the same HBM-resident state
    are reported beside it;
  * pergate_gradient: the same gradient through the plain per-gate kernels
    (same adjoint algorithm, no fusion) -- the ">= 5x" denominator;
  * config2 / small_circuit / dense_block / batch / config5: the other
    workloads, each with its own rate and, where it applies, roofline;
  * cpu_baseline: the reference's per-gate torch bmm path (restated) with
    torch autograd on the host cores, on a bounded sample of the headline
    workload, scaled to gradient evaluations/s;
  * oracle_error: errors against the complex128 oracle for every workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_config2(q, ops, tqf):
    for op in ops:
        p = op["params"]
        tqf.func_name_dict[op["name"]](q, op["wires"], params=None if p is None else
                                       torch.tensor(p, dtype=torch.float32, device=q.device))


def time_forward(n, bsz, ops, mode, steps, warmup, tq, jit="auto"):
    import torchquantum_amd.functional as tqf
    params = [None if op["params"] is None else torch.tensor(op["params"], dtype=torch.float32, device="cuda")
              for op in ops]

    def step():
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
        q.engine_options = tq.EngineOptions(mode=mode, jit=jit, **_plan_opts(mode))
        for op, p in zip(ops, params):
            tqf.func_name_dict[op["name"]](q, op["wires"], params=p)
        return tq.MeasureAll(tq.PauliZ)(q)

    for _ in range(warmup):
        step()
    if warmup and tq.wait_for_compiled_stages():   # compiled stages (jit="auto") are ready: run them once
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    t_issue = time.perf_counter() - t0
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    time_forward.host_issue_s = t_issue / steps
    return dt, out


STAGE_BITS = 0      # local bits per stage of the fused plans (0 = the engine's choice)
LOW_BITS = None     # lowest bits kept local in every stage (None = the engine's choice)
IO_LAYOUTS = True   # stages load / store in the layouts of their first / last ops where lines stay whole
GRAD_CAP = 0        # ops with gradient parameters per stage (0 = no limit)


def _plan_opts(mode="fused"):
    """EngineOptions fields set from the command line (fused plans only)."""
    if mode != "fused":
        return {}
    d = {} if LOW_BITS is None else {"low_bits": LOW_BITS}
    if STAGE_BITS:
        d["stage_bits"] = STAGE_BITS
    if not IO_LAYOUTS:
        d["io_layouts"] = False
    if GRAD_CAP:
        d["grad_cap"] = GRAD_CAP
    return d


def time_replay(n, bsz, ops, steps, warmup, tq):
    """The same forward recorded once with tq.CircuitReplay and replayed:
    seconds per batch, host issue seconds per batch, the <Z_i> output and the
    stage-kernel milliseconds per batch (HIP events)."""
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E
    params = [None if op["params"] is None else torch.tensor(op["params"], dtype=torch.float32, device="cuda")
              for op in ops]

    def record(q):
        for op, p in zip(ops, params):
            tqf.func_name_dict[op["name"]](q, op["wires"], params=p)

    rep = tq.CircuitReplay(record, n, bsz, device="cuda")
    meas = tq.MeasureAll(tq.PauliZ)
    for _ in range(warmup):
        meas(rep())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = meas(rep())
    t_issue = time.perf_counter() - t0
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    rec = []
    E.set_stage_timer(rec)
    try:
        for _ in range(steps):
            meas(rep())
    finally:
        E.set_stage_timer(None)
    torch.cuda.synchronize()
    stage_ms = sum(e0.elapsed_time(e1) for _, _, e0, e1 in rec) / steps
    return dt, t_issue / steps, out, stage_ms


def _ansatz_loss_step(n, bsz, layers, mode, tq, jit, params, phi, ops, obs, coef):
    import torchquantum_amd.functional as tqf
    for p in params:
        p.grad = None
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
    q.engine_options = tq.EngineOptions(mode=mode, jit=jit, **_plan_opts(mode))
    for w in range(n):
        tqf.ry(q, w, params=phi)
    k = 0
    for op in ops:
        if op["params"] is None:
            tqf.func_name_dict[op["name"]](q, op["wires"])
        else:
            tqf.func_name_dict[op["name"]](q, op["wires"], params=params[k])
            k += 1
    loss = tq.expval_pauli_sum(q, obs, coef).mean()
    if params[0].requires_grad:
        loss.backward()
    return loss


def ansatz_inputs(n, bsz, layers, theta=None):
    """Circuit, observables, per-row input angles and parameter tensors of the
    gradient workload (angles N(0, 0.1) from seed 0, input RY angles U[0, pi)
    from seed 1)."""
    from circuits import config3_ops, config3_observables
    ops = config3_ops(seed=0, n=n, layers=layers)
    if theta is None:
        theta = [float(np.asarray(op["params"]).reshape(-1)[0]) for op in ops if op["params"] is not None]
    params = [torch.tensor([[t]], dtype=torch.float32, device="cuda", requires_grad=True) for t in theta]
    obs, coef = config3_observables(n)
    g = torch.Generator().manual_seed(1)
    phi = (torch.rand(bsz, 1, generator=g) * np.pi).to("cuda")
    return ops, obs, coef, phi, params


def time_gradient(n, bsz, layers, mode, steps, warmup, tq, jit="auto", timer=None):
    """Seconds per gradient evaluation (mean of ``steps`` after ``warmup``);
    with ``timer`` a list, HIP events around every stage launch of the timed
    steps are collected into it."""
    from torchquantum_amd import engine as E
    ops, obs, coef, phi, params = ansatz_inputs(n, bsz, layers)

    def step():
        return _ansatz_loss_step(n, bsz, layers, mode, tq, jit, params, phi, ops, obs, coef)

    for i in range(warmup):
        step()
        log(f"[gradient {n}q {mode}] warm-up step {i + 1}/{warmup}")
    if warmup and tq.wait_for_compiled_stages():
        step()
        log(f"[gradient {n}q {mode}] compiled stage kernels ready")
    torch.cuda.synchronize()
    if timer is not None:
        E.set_stage_timer(timer)
    try:
        t0 = time.perf_counter()
        for _ in range(steps):
            loss = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
    finally:
        E.set_stage_timer(None)
    grads = torch.cat([p.grad.reshape(-1) for p in params])
    return dt, float(loss.detach()), grads


def stage_launch_stats(records, steps, direction, amps, bytes_per_amp):
    """Average duration (HIP events) of the compiled stage launches of one
    direction, and the algorithmic rate: bytes_per_amp x amps per launch."""
    ms = [e0.elapsed_time(e1) for d, comp, e0, e1 in records if d == direction and comp]
    if not ms:
        return None
    avg = float(np.mean(ms))
    nbytes = bytes_per_amp * amps
    return {"avg_launch_ms": avg, "launches_per_step": len(ms) / steps, "ms_per_step": sum(ms) / steps,
            "bytes_per_launch": nbytes, "achieved_GBs": nbytes / (avg * 1e-3) / 1e9}


def parameter_shift_check(n, bsz, layers, tq, n_params=8, seed=7):
    """Full-size check of the fused adjoint gradient: the parameter-shift rule
    (f(t + pi/2) - f(t - pi/2)) / 2, exact for RX/RY/RZ
    (util/utils.py:1017-1075), from GPU forward passes only, on seeded
    parameters of the workload."""
    ops, obs, coef, phi, params = ansatz_inputs(n, bsz, layers)
    _ansatz_loss_step(n, bsz, layers, "fused", tq, "auto", params, phi, ops, obs, coef)
    g = torch.cat([p.grad.reshape(-1) for p in params]).cpu().numpy().astype(np.float64)
    theta = [float(p.detach()) for p in params]
    idx = np.random.default_rng(seed).choice(len(theta), n_params, replace=False)
    errs = []
    with torch.no_grad():
        for i in idx:
            vals = []
            for sh in (np.pi / 2, -np.pi / 2):
                th = list(theta)
                th[i] += sh
                _, _, _, _, ps = ansatz_inputs(n, bsz, layers, theta=th)
                for p in ps:
                    p.requires_grad_(False)
                vals.append(float(_ansatz_loss_step(n, bsz, layers, "fused", tq, "auto", ps, phi, ops, obs, coef)))
            errs.append(abs(g[i] - (vals[0] - vals[1]) / 2))
    return {"params_checked": [int(i) for i in idx], "grad_max_abs_vs_parameter_shift": float(max(errs))}


def kernel_roofline(n, bsz, iters=50):
    """Average duration of the dense 1-qubit gate (RY, target cycling over every
    bit 0..n-1, as RY on all wires does: k_1q_t0 for bit 0, k_1q_tile for bits
    1..8, k_1q_pair above) with HIP events on the launch stream; algorithmic
    bytes = 16 B x bsz x 2^n per launch."""
    from torchquantum_amd import _native as N
    st = torch.zeros((bsz, 1 << n), dtype=torch.complex64, device="cuda")
    st[:, 0] = 1
    th = 0.7
    m = torch.tensor([[np.cos(th / 2), -np.sin(th / 2)], [np.sin(th / 2), np.cos(th / 2)]],
                     dtype=torch.complex64, device="cuda")
    stream = N.stream_of(st)
    for i in range(5):
        N.call("tqsv_apply_1q", stream, st.data_ptr(), bsz, n, i % n, 0, 0, m.data_ptr(), 0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for i in range(iters):
        N.call("tqsv_apply_1q", stream, st.data_ptr(), bsz, n, i % n, 0, 0, m.data_ptr(), 0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    nbytes = 16.0 * bsz * (1 << n)
    return ms, nbytes


def dense_block_bench(n=26, bsz=16, iters=6):
    """Dense k-qubit blocks (k = 5, 6) on an 8 GiB state: the matrix-core
    kernel (one matrix for all rows) against the VALU kernel (the same matrix
    given per row, which takes the LDS/VALU path).  Real FLOPs per amplitude:
    8 * 2^k (a 2^(k+1) x 2^(k+1) real GEMM per group of 2^k amplitudes)."""
    from torchquantum_amd import _native as N
    st = torch.zeros((bsz, 1 << n), dtype=torch.complex64, device="cuda")
    st[:, 0] = 1
    stream = N.stream_of(st)
    out = {}
    rng = np.random.default_rng(0)
    for k in (5, 6):
        z = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
        u, _ = np.linalg.qr(z)
        m1 = torch.tensor(u, dtype=torch.complex64, device="cuda")
        mb = m1.unsqueeze(0).expand(bsz, -1, -1).contiguous()
        bits = N.int_array([n - 1 - w for w in range(3, 3 + k)])      # wires 3..3+k-1
        res = {}
        for name, mat, bstride in (("mfma", m1, 0), ("valu", mb, (1 << k) ** 2)):
            for _ in range(2):
                N.call("tqsv_apply_kq", stream, st.data_ptr(), bsz, n, k, bits, 0, 0, mat.data_ptr(), bstride)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(iters):
                N.call("tqsv_apply_kq", stream, st.data_ptr(), bsz, n, k, bits, 0, 0, mat.data_ptr(), bstride)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / iters
            amps = float(bsz) * (1 << n)
            res[name] = {"ms": ms, "tflops": 8 * (2 ** k) * amps / (ms * 1e-3) / 1e12,
                         "gbs": 16 * amps / (ms * 1e-3) / 1e9}
        res["mfma_frac_of_fp32_peak"] = res["mfma"]["tflops"] / MFMA_FP32_PEAK_TFLOPS
        res["speedup_mfma_vs_valu"] = res["valu"]["ms"] / res["mfma"]["ms"]
        out[f"k{k}"] = res
    return {"workload": f"{n} qubits, bsz {bsz}, dense 2^k x 2^k block on wires 3..3+k-1", **out}


def time_batch(n_circ, steps, tq):
    """256 circuits, n ~ U{8..22}, depth ~ U{4..40}: the batch as one
    persistent work-queue launch (and, for comparison, one launch per level),
    captured in a HIP graph, vs one engine call per circuit."""
    from torchquantum_amd import batch as B
    import torchquantum_amd.functional as tqf
    circs = B.random_circuits(n_circ=n_circ, n_range=(8, 22), depth_range=(4, 40), seed=0)
    upd = float(sum(len(ops) * (1 << n) for n, ops in circs))
    t0 = time.perf_counter()
    cb = B.CircuitBatch(circs, device="cuda")
    t_build = time.perf_counter() - t0
    cb.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cb.launch()
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / steps
    cb.capture()
    cb.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = cb.run()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    cb.check()
    cl = B.CircuitBatch(circs, device="cuda", mode="levels")
    t_levels = _replay_ms(cl, steps)
    same_as_levels = bool(torch.equal(cl.states, cb.states))
    del cl
    # one engine call per circuit (fused plans, the same interpreting op code as
    # the batch kernel, Python loop over circuits)
    def per_circuit():
        res = []
        for n, ops in circs:
            q = tq.QuantumDevice(n_wires=n, bsz=1, device="cuda")
            q.engine_options = tq.EngineOptions(jit="off")
            for op in ops:
                p = op["params"]
                tqf.func_name_dict[op["name"]](q, op["wires"], params=None if p is None else
                                               torch.tensor(p, dtype=torch.float32).reshape(1, -1))
            res.append(tq.MeasureAll(tq.PauliZ)(q)[0, 0])
        return torch.stack(res)
    ref = per_circuit()
    torch.cuda.synchronize()
    per_circuit()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = per_circuit()
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    alg = batch_pass_bytes(cb)
    tuned = tune_batch(circs, tq)
    return {"workload": f"{n_circ} circuits, n ~ U{{8..22}}, depth ~ U{{4..40}} layers of n random gates "
                        "from {H,RX,RY,RZ,U3,CNOT,CZ,SWAP}, <Z_0> each, bsz 1",
            "roofline": {"bound": "hbm", "algorithmic_bytes_per_batch": alg,
                         "achieved": alg / t_graph / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / t_graph / 1e9 / HBM_PEAK_GBS,
                         "stages_total": int(sum(len(o) for o in cb.stage_offsets)),
                         "pergate_equivalent_bytes": pergate_bytes(circs),
                         "note": "16 B x 2^n per executed stage of each circuit + the <Z_0> read, per graph "
                                 "replay; states of <= 64 MiB are cache-resident, so this is not an HBM figure"},
            "mode": "one persistent work-queue launch (items level-major, each waiting for its circuit's "
                    "previous stage)",
            "levels": len(cb.levels), "launches_per_batch": cb.n_launches,
            "ms_graph_replay": t_graph * 1e3, "ms_eager_batch": t_eager * 1e3, "ms_per_circuit_loop": t_loop * 1e3,
            "ms_graph_replay_launch_per_level": t_levels, "states_equal_launch_per_level": same_as_levels,
            "amp_updates_per_s_graph": upd / t_graph, "speedup_vs_per_circuit_loop": t_loop / t_graph,
            "build_s": t_build, "z0_max_abs_diff_vs_per_circuit": float((out - ref).abs().max()),
            "plans": tuned}, \
        (out.reshape(-1).cpu().numpy(), circs)


def encoder_training(tq, n, bsz, blocks, steps, warmup):
    """One QML training step = encoder (per-row RY on every wire from a
    [bsz, n] input, matrices built on the GPU) + RXYZCXLayer0 ansatz +
    <Z_i> read-out + a linear head + cross entropy, backward, and an Adam
    update of the ansatz and head inside the timed loop (the training loop of
    examples/mnist/mnist.py:82-135 at a larger size).  New synthetic inputs
    every step."""
    enc = tq.GeneralEncoder([{"input_idx": [i], "func": "ry", "wires": [i]} for i in range(n)])
    ansatz = tq.RXYZCXLayer0({"n_wires": n, "n_blocks": blocks}).to("cuda")
    head = torch.nn.Linear(n, 4).to("cuda")
    opt = torch.optim.Adam(list(ansatz.parameters()) + list(head.parameters()), lr=1e-2)
    g = torch.Generator(device="cuda").manual_seed(0)
    xs = [torch.rand(bsz, n, device="cuda", generator=g) * np.pi for _ in range(4)]
    ys = [torch.randint(0, 4, (bsz,), device="cuda", generator=g) for _ in range(4)]

    def step(i):
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
        enc(q, xs[i % 4])
        ansatz(q)
        logits = head(tq.MeasureAll(tq.PauliZ)(q))
        loss = torch.nn.functional.cross_entropy(logits, ys[i % 4])
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss

    def timed(step):
        for i in range(warmup):
            step(i)
        tq.wait_for_compiled_stages()
        step(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            loss = step(i)
        t_host = time.perf_counter() - t0
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        return {"ms_per_step": dt * 1e3, "steps_per_s": 1.0 / dt, "host_issue_ms_per_step": t_host / steps * 1e3,
                "loss_last": float(loss.detach())}

    res = timed(step)
    # the same training loop on a circuit recorded once (tq.CircuitReplay):
    # the encoder's input columns are gathered on the GPU per call, the
    # ansatz parameters are read after every Adam step
    rep = tq.CircuitReplay(lambda q, x: (enc(q, x), ansatz(q)), n_wires=n, bsz=bsz, example_inputs=(xs[0],))

    def step_replay(i):
        logits = head(tq.MeasureAll(tq.PauliZ)(rep(xs[i % 4])))
        loss = torch.nn.functional.cross_entropy(logits, ys[i % 4])
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss

    res["replay"] = timed(step_replay)
    n_par = sum(p.numel() for p in ansatz.parameters())
    return {"workload": f"{n} qubits, bsz {bsz}: RY encoder from a [bsz, {n}] input + RXYZCXLayer0 x {blocks} "
                        f"({n_par} params) + <Z_i> + linear head, cross entropy, Adam step; new inputs each step",
            **res}


def _replay_ms(cb, steps=5):
    cb.capture()
    cb.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cb.run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    cb.check()
    return ms


def tune_batch(circs, tq):
    """Per-circuit plans for the uneven batch: every circuit autotuned over
    fusion widths 1..4 (exhaustive: each candidate timed with the
    program-interpreting stage kernels, the op code the batch kernel runs),
    and the learned planner trained on a quarter of those timings choosing
    for all circuits from plan features alone.  Reports the batch time with
    the default plans, the exhaustive winners and the learned choices, and
    the learned planner's regret on the held-out circuits."""
    from torchquantum_amd import batch as B
    from torchquantum_amd.autotune import Autotuner, LearnedPlanner, BATCH_CANDIDATES, circuit_key, tune_circuits
    from torchquantum_amd.costmodel import MLPCostModel
    tapes = [(B.ops_to_tape(ops), n) for n, ops in circs]
    tuner = Autotuner()
    t0 = time.perf_counter()
    best = tune_circuits(tuner, tapes, BATCH_CANDIDATES, reps=3)
    t_tune = time.perf_counter() - t0
    log(f"[bench] batch: exhaustive autotune {t_tune:.1f} s")
    keys = [circuit_key(tape, n, 1) for tape, n in tapes]
    train = set(keys[::4])
    X, y = tuner.dataset(sorted(train))
    model = MLPCostModel(device="cuda", seed=0)
    hist = model.fit(X, y, epochs=1500, lr=3e-3)
    planner = LearnedPlanner(model, BATCH_CANDIDATES)
    t0 = time.perf_counter()
    picks = [planner.choose(tape, n, 1) for tape, n in tapes]
    t_pick = time.perf_counter() - t0
    held = [i for i, k in enumerate(keys) if k not in train]

    def ms_of(i, o):
        by = {json.dumps(r["opts"], sort_keys=True): r["ms"] for r in tuner.db[keys[i]]["records"]}
        o = dict(o.__dict__, stage_bits=min(o.stage_bits or 12, tapes[i][1]))
        return by[json.dumps(o, sort_keys=True)]

    sum_def = sum(ms_of(i, BATCH_CANDIDATES[0]) for i in held)
    sum_best = sum(ms_of(i, best[i]) for i in held)
    sum_pick = sum(ms_of(i, picks[i]) for i in held)
    out = {"candidates": "fusion width 1..4, stage width min(12, n), interpreting stage kernels",
           "autotune_s": t_tune, "learned_selection_s": t_pick, "train_circuits": len(train),
           "held_out_circuits": len(held), "train_loss_first_last": [hist[0], hist[-1]],
           "held_out_sum_ms_default": sum_def, "held_out_sum_ms_exhaustive": sum_best,
           "held_out_sum_ms_learned": sum_pick, "learned_gain_vs_default": sum_def / sum_pick,
           "learned_regret_vs_exhaustive": sum_pick / sum_best - 1,
           "widths_picked": {str(k): sum(1 for o in picks if o.fusion_width == k) for k in (1, 2, 3, 4)}}
    for label, opts in (("exhaustive", best), ("learned", picks)):
        cb = B.CircuitBatch(circs, device="cuda", options=opts)
        out[f"ms_graph_replay_{label}_plans"] = _replay_ms(cb)
        del cb
        torch.cuda.empty_cache()
    return out


def copy_calibration(nbytes=1 << 31):
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    a = torch.empty(nbytes // 8, dtype=torch.complex64, device="cuda").fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    return nbytes, e0.elapsed_time(e1)


def load_traffic(name):
    """HBM bytes per launch of a kernel from a committed counter summary
    (profiles/<name>.json: two rocprofv3 --pmc runs, FETCH_SIZE and
    WRITE_SIZE, scaled by what they read for a device copy of known size)."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def config5(args, tq):
    """30 qubits, bsz 4 (32 GiB): 20 x [RY all, CZ even/odd] + a Toffoli or CSWAP
    on seeded wires after every 5th layer; loss = sum_i <Z_i Z_i+1> (open chain)."""
    import torchquantum_amd.functional as tqf
    from circuits import config5_ops, chain_zz
    n, bsz = args.c5_qubits, args.c5_bsz
    ops = config5_ops(seed=0, n=n)
    obs = chain_zz(n)
    params = [torch.tensor(op["params"], dtype=torch.float32, device="cuda", requires_grad=True)
              if op["params"] is not None else None for op in ops]
    upd = len(ops) * bsz * float(1 << n)

    def build(mode):
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
        q.engine_options = (tq.EngineOptions(mode=mode, fusion_width=args.c5_fusion_width, **_plan_opts(mode))
                            if mode == "fused" else tq.EngineOptions(mode=mode))
        for op, p in zip(ops, params):
            tqf.func_name_dict[op["name"]](q, op["wires"], params=p)
        return q

    res = {"workload": f"{n} qubits, bsz {bsz}, 20 x [RY all, CZ even+odd] + Toffoli/CSWAP every 5 layers",
           "gates": len(ops), "fusion_width": args.c5_fusion_width}
    with torch.no_grad():
        # one untimed fused pass plans the circuit and starts the background
        # compile of its stage kernels; the timed passes run the compiled ones
        log("[config5] fused forward (warm-up, compile)")
        t0 = time.perf_counter()
        q = build("fused")
        tq.expval_joint_analytical_batch(q, obs[:4])
        del q
        torch.cuda.empty_cache()
        tq.wait_for_compiled_stages()
        res["warmup_and_compile_s"] = time.perf_counter() - t0
        for mode in ("fused", "pergate"):
            log(f"[config5] {mode} forward")
            q = build(mode)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ev = tq.expval_joint_analytical_batch(q, obs[:4])
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            res[f"{mode}_forward_s"] = dt
            res[f"{mode}_amp_updates_per_s"] = upd / dt
            res[f"{mode}_zz_first4"] = ev[0].tolist()
            if mode == "fused":
                sv = q.get_states_1d()
                res["norm_max_dev"] = float((torch.linalg.vector_norm(sv, dim=1) - 1).abs().max())
                sample = sv[:, :: (1 << 12)].clone()
            else:
                res["fused_vs_pergate_amp_max_abs"] = float((q.get_states_1d()[:, :: (1 << 12)] - sample).abs().max())
            del q
            torch.cuda.empty_cache()
    log("[config5] fused gradient (warm-up, compile)")
    q = build("fused")
    loss = tq.expval_pauli_sum(q, obs, [1.0] * len(obs)).mean()
    loss.backward()
    del q, loss
    for p in params:
        if p is not None:
            p.grad = None
    tq.wait_for_compiled_stages()
    log("[config5] fused gradient")
    q = build("fused")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = tq.expval_pauli_sum(q, obs, [1.0] * len(obs)).mean()
    loss.backward()
    torch.cuda.synchronize()
    res["fused_gradient_s"] = time.perf_counter() - t0
    res["n_params"] = sum(p.numel() for p in params if p is not None)
    res["loss"] = float(loss.detach())
    return res


def config5_plans(args, tq):
    """Plan choice for the 30-qubit workload.  Candidates: fusion width
    1..6 (dense blocks; widths 5 and 6 run on the matrix-core kernel), stage
    width 12, timed with the interpreting stage kernels (no run-time compile
    per candidate).  The learned planner -- the MLP trained by this package's
    kernels on exhaustive timings of six smaller circuits of the same family
    (16, 20 and 24 qubits, bsz 1) -- chooses from plan features alone;
    exhaustive timing of every candidate at full size scores the choice."""
    from torchquantum_amd import batch as B
    from torchquantum_amd.autotune import Autotuner, LearnedPlanner, circuit_key, time_plan
    from torchquantum_amd.costmodel import MLPCostModel
    from circuits import config5_ops
    cands = [tq.EngineOptions(fusion_width=k, stage_bits=12, jit="off") for k in range(1, 7)]
    tuner = Autotuner()
    keys = []
    t0 = time.perf_counter()
    for n in (16, 20, 24):
        for seed in (1, 2):
            tape = B.ops_to_tape(config5_ops(seed=seed, n=n))
            tuner.tune(tape, n, 1, candidates=cands, reps=3)
            keys.append(circuit_key(tape, n, 1))
            log(f"[config5-plans] training circuit {len(keys)}/6 ({n} qubits) timed")
    t_data = time.perf_counter() - t0
    X, y = tuner.dataset(keys)
    model = MLPCostModel(device="cuda", seed=0)
    hist = model.fit(X, y, epochs=1500, lr=3e-3)
    planner = LearnedPlanner(model, cands)
    n, bsz = args.c5_qubits, args.c5_bsz
    tape = B.ops_to_tape(config5_ops(seed=0, n=n))
    t0 = time.perf_counter()
    opts, pred = planner.score(tape, n, bsz)
    t_pick = time.perf_counter() - t0
    pick = opts[int(np.argmin(pred))]
    measured = []
    t0 = time.perf_counter()
    for o in opts:
        measured.append(time_plan(tape, n, bsz, o, reps=1))
        torch.cuda.empty_cache()
        log(f"[config5-plans] fusion width {o.fusion_width}: {measured[-1]:.0f} ms")
    t_exhaustive = time.perf_counter() - t0
    best = int(np.argmin(measured))
    ip = opts.index(pick)
    return {"workload": f"{n} qubits, bsz {bsz}, the config-5 circuit, forward",
            "candidates": "fusion width 1..6 (5, 6: matrix-core blocks), stage width 12, interpreting stage kernels",
            "train_circuits": "config-5 family, seeds 1-2, 16/20/24 qubits, bsz 1 (36 timed plans)",
            "train_timing_s": t_data, "train_loss_first_last": [hist[0], hist[-1]],
            "ms_by_fusion_width": {str(o.fusion_width): m for o, m in zip(opts, measured)},
            "predicted_ms_by_fusion_width": {str(o.fusion_width): float(p) for o, p in zip(opts, pred)},
            "learned_pick": pick.fusion_width, "exhaustive_pick": opts[best].fusion_width,
            "learned_regret_vs_exhaustive": measured[ip] / measured[best] - 1,
            "default_plan_ms": measured[0], "learned_selection_s": t_pick, "exhaustive_timing_s": t_exhaustive}


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(ops_sample, n, bsz, n_gates_total_per_step):
    from oracle import refpath
    threads = cpu_threads()
    torch.set_num_threads(threads)
    t0 = time.perf_counter()
    refpath.run(ops_sample, n, bsz, device="cpu")
    dt = time.perf_counter() - t0
    upd = len(ops_sample) * bsz * (1 << n)
    return upd / dt, threads, dt


def cpu_baseline_gradient(n, bsz_full, layers, n_gates):
    """The reference's per-gate torch path under torch autograd (what the
    reference computes for a gradient: gate_wrapper.py:96-140, complex64,
    every bmm input saved for backward) on the host cores: the first
    ``n_gates`` gates of the gradient workload at bsz 1, forward + backward,
    loss restricted to the ZZ ring (one diagonal pass).  Scaled to gradient
    evaluations/s of the whole workload: 1 / (t_sample x gates / n_gates x bsz)."""
    from oracle import refpath as R
    from circuits import config3_ops
    threads = cpu_threads()
    torch.set_num_threads(threads)
    ops = config3_ops(seed=0, n=n, layers=layers)
    total = len(ops) + n                      # + the per-row input layer
    sample = ops[:n_gates]
    params = [None if op["params"] is None else
              torch.tensor(np.asarray(op["params"], dtype=np.float32).reshape(1, -1), requires_grad=True)
              for op in sample]
    j = torch.arange(1 << n)
    sign = torch.zeros(1 << n)
    for i in range(n):
        b0 = (j >> (n - 1 - i)) & 1
        b1 = (j >> (n - 1 - (i + 1) % n)) & 1
        sign += (1 - 2 * (b0 ^ b1)).float()
    del j
    t0 = time.perf_counter()
    st = R.run_autograd(sample, params, n, 1, torch.complex64)
    loss = ((st.reshape(1, -1).abs() ** 2) * sign).sum()
    loss.backward()
    dt = time.perf_counter() - t0
    return 1.0 / (dt * (total / n_gates) * bsz_full), threads, dt, total


F_PASS = {"cnot": 0.5, "swap": 0.5, "cz": 0.25, "toffoli": 0.25, "cswap": 0.25}


def pergate_bytes(circuits):
    """Sum over circuits and gates of 16 B x 2^n x f (f = 1 dense, 1/2 CNOT/SWAP,
    1/4 CZ/Toffoli/CSWAP), plus the <Z_0> read-out (8 B x 2^n) per circuit:
    the traffic of a per-gate execution."""
    tot = 0.0
    for n, ops in circuits:
        a = float(1 << n)
        tot += sum(16.0 * a * F_PASS.get(op["name"], 1.0) for op in ops) + 8.0 * a
    return tot


def batch_pass_bytes(cb):
    """Algorithmic bytes of the batch as executed: every stage of circuit c is
    one read + write of its 2^n_c amplitudes (16 B each), plus the <Z_0>
    read-out (8 B per amplitude)."""
    return float(sum((16.0 * len(offs) + 8.0) * (1 << n) for n, offs in zip(cb.ns, cb.stage_offsets)))


def oracle_errors(tq, batch_out=None):
    """Max-abs error of the GPU results against the complex128 oracle, per
    workload, at sizes the oracle finishes in seconds (part of the CPU leg:
    the oracle is the checker here, never the thing timed).  Configs 1 and 2
    run at full size (their rows share parameters and start in |0..0>, so one
    oracle row stands for all); the gradient ansatz and the 30-qubit circuit
    run through the same builders at fewer qubits; the uneven batch is checked
    on its circuits of at most 12 qubits."""
    from oracle import statevector as O
    import torchquantum_amd.functional as tqf
    from circuits import config1_ops, config2_ops, config3_ops, config3_observables, config5_ops
    res = {}

    def run_gpu(n, bsz, ops, opts):
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
        q.engine_options = opts
        build_config2(q, ops, tqf)
        z = tq.MeasureAll(tq.PauliZ)(q)
        return q.get_states_1d().cpu().numpy(), z.cpu().numpy()

    forward_cases = (("config1: 10 q, bsz 32 (full size)", 10, 32, config1_ops(seed=0)),
                     ("config2: 20 q, bsz 16 (full size)", 20, 16, config2_ops(seed=0, n=20, layers=20)),
                     ("config5 circuit at 16 q, bsz 4", 16, 4, config5_ops(seed=0, n=16, layers=20)))
    for name, n, bsz, ops in forward_cases:
        log(f"[bench] oracle check {name}")
        ref = O.run_ops(ops, n, 1)
        zref = O.expval_z_all(ref, n)
        e = {}
        for label, opts in (("pergate", tq.EngineOptions(mode="pergate")),
                            ("fused", tq.EngineOptions(mode="fused", jit="on")),
                            ("fused_interpreting", tq.EngineOptions(mode="fused", jit="off"))):
            st, z = run_gpu(n, bsz, ops, opts)
            e[label] = {"amp_max_abs": float(np.abs(st - ref).max()), "expval_max_abs": float(np.abs(z - zref).max())}
        res[name] = e

    # gradient ansatz (10 x [RX,RY,RZ all wires, CNOT ring], per-row RY input
    # layer, loss = mean_b sum_t c_t <P_t>) at 10 qubits bsz 4: loss and a
    # seeded 32-parameter subset of the gradient against central differences
    # of the oracle (h = 1e-3)
    n, bsz = 10, 4
    log(f"[bench] oracle check gradient ansatz at {n} q")
    ops = config3_ops(seed=0, n=n, layers=10)
    obs, coef = config3_observables(n)
    phi = np.random.default_rng(1).random(bsz) * np.pi
    theta0 = np.array([float(np.asarray(op["params"]).reshape(-1)[0]) for op in ops if op["params"] is not None])

    def oracle_loss(theta):
        seq = [dict(name="ry", wires=[w], params=phi.reshape(bsz, 1), inverse=False) for w in range(n)]
        k = 0
        for op in ops:
            if op["params"] is None:
                seq.append(op)
            else:
                seq.append(dict(op, params=np.array([[theta[k]]])))
                k += 1
        st = O.run_ops(seq, n, bsz)
        return float(np.mean(sum(c * O.expval_pauli(st, o) for o, c in zip(obs, coef))))

    idx = np.random.default_rng(2).choice(len(theta0), 32, replace=False)
    g_ref = O.central_fd(oracle_loss, theta0, 1e-3, idx)[idx]
    loss_ref = oracle_loss(theta0)
    e = {}
    for label, opts in (("pergate", tq.EngineOptions(mode="pergate")),
                        ("fused", tq.EngineOptions(mode="fused", jit="on")),
                        ("fused_interpreting", tq.EngineOptions(mode="fused", jit="off"))):
        params = [torch.tensor([[t]], dtype=torch.float32, device="cuda", requires_grad=True) for t in theta0]
        q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cuda")
        q.engine_options = opts
        ph = torch.tensor(phi, dtype=torch.float32, device="cuda").reshape(bsz, 1)
        for w in range(n):
            tqf.ry(q, w, params=ph)
        k = 0
        for op in ops:
            if op["params"] is None:
                tqf.func_name_dict[op["name"]](q, op["wires"])
            else:
                tqf.func_name_dict[op["name"]](q, op["wires"], params=params[k])
                k += 1
        loss = tq.expval_pauli_sum(q, obs, coef).mean()
        loss.backward()
        g = torch.cat([p.grad.reshape(-1) for p in params]).cpu().numpy().astype(np.float64)
        e[label] = {"loss_abs": abs(float(loss.detach()) - loss_ref), "grad_max_abs_vs_fd": float(np.abs(g[idx] - g_ref).max())}
    res[f"config3 ansatz at {n} q, bsz {bsz} (32-param FD subset)"] = e

    if batch_out is not None:
        z0, circs = batch_out
        log("[bench] oracle check uneven batch (circuits of <= 12 qubits)")
        err, cnt = 0.0, 0
        for c, (nq, cops) in enumerate(circs):
            if nq > 12:
                continue
            ref = O.expval_z_all(O.run_ops(cops, nq, 1), nq)[0, 0]
            err = max(err, abs(float(z0[c]) - float(ref)))
            cnt += 1
        res["config4 batch"] = {"circuits_checked": cnt, "z0_max_abs": err}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed gradient evaluations of the headline workload")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--adj-qubits", type=int, default=24)
    ap.add_argument("--adj-bsz", type=int, default=64)
    ap.add_argument("--pergate-steps", type=int, default=1, help="timed per-gate gradient evaluations")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-config5", action="store_true")
    ap.add_argument("--skip-extras", action="store_true",
                    help="headline gradient (fused, per-gate) and its roofline only")
    ap.add_argument("--only-roofline", action="store_true",
                    help="run only the dense 1-qubit kernel timing loops (for rocprofv3 summaries)")
    ap.add_argument("--config5", action="store_true", help="run the 30-qubit workload only")
    ap.add_argument("--only-dense", action="store_true", help="run only the dense k = 5, 6 block timings")
    ap.add_argument("--only-fused", action="store_true",
                    help="run only the fused forward of the 20-qubit RY/CZ circuit (--steps, --warmup)")
    ap.add_argument("--only-pergate", action="store_true",
                    help="run only the per-gate forward of the 20-qubit RY/CZ circuit (--steps, --warmup)")
    ap.add_argument("--low-bits", type=int, default=None,
                    help="lowest state bits kept local in every stage of the gradient workload (default: the engine's)")
    ap.add_argument("--grad-cap", type=int, default=0,
                    help="at most this many ops with gradient parameters per stage (0 = no limit)")
    ap.add_argument("--canonical-io", action="store_true",
                    help="fused plans load and store every stage in the canonical layout (remaps at both ends)")
    ap.add_argument("--stage-bits", type=int, default=0,
                    help="local bits per stage of the fused plans (default: the engine's choice)")
    ap.add_argument("--c5-fusion-width", type=int, default=1,
                    help="fusion width of the 30-qubit workload's fused plan (1 = merged single-wire runs)")
    ap.add_argument("--config5-plans", action="store_true",
                    help="only the 30-qubit plan choice: learned planner vs exhaustive timing")
    ap.add_argument("--only-small", action="store_true",
                    help="only the 10-qubit bsz-32 circuit: functional API and CircuitReplay")
    ap.add_argument("--only-gradient", action="store_true",
                    help="run only the fused adjoint gradient (--adj-qubits, --adj-bsz, --steps)")
    ap.add_argument("--only-batch", action="store_true", help="run only the uneven batch")
    ap.add_argument("--only-encoder", action="store_true", help="run only the encoder training steps")
    ap.add_argument("--c5-qubits", type=int, default=30)
    ap.add_argument("--c5-bsz", type=int, default=4)
    ap.add_argument("--cpu-layers", type=int, default=2)
    ap.add_argument("--cpu-grad-gates", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (sharding over several "
                                   "GPUs is designed but not built)"}))
        sys.exit(2)
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no ROCm GPU visible"}))
        sys.exit(2)
    import torchquantum_amd as tq
    from circuits import config2_ops
    global STAGE_BITS, LOW_BITS, IO_LAYOUTS, GRAD_CAP
    GRAD_CAP = args.grad_cap
    STAGE_BITS = args.stage_bits
    LOW_BITS = args.low_bits
    IO_LAYOUTS = not args.canonical_io

    if args.only_roofline:
        cb, cms = copy_calibration()
        ms, nbytes = kernel_roofline(20, 16, iters=60)
        ms_big, nbytes_big = kernel_roofline(24, 64, iters=24)
        print(json.dumps({"copy_bytes": cb, "copy_ms": cms, "k_1q_20q_bsz16_ms": ms,
                          "k_1q_20q_bsz16_GBs": nbytes / ms / 1e6, "k_1q_24q_bsz64_ms": ms_big,
                          "k_1q_24q_bsz64_GBs": nbytes_big / ms_big / 1e6}), flush=True)
        return
    if args.only_dense:
        print(json.dumps({"dense_block": dense_block_bench()}), flush=True)
        return
    if args.only_encoder:
        print(json.dumps({"encoder_training": {"small": encoder_training(tq, 10, 256, 4, 20, 3),
                                               "large": encoder_training(tq, 22, 64, 4, 3, 2)}}), flush=True)
        return
    if args.only_batch:
        res, _ = time_batch(256, 5, tq)
        print(json.dumps({"batch": res}), flush=True)
        return
    if args.only_small:
        from circuits import config1_ops
        ops1 = config1_ops(seed=0)
        t1_f, ev_f1 = time_forward(10, 32, ops1, "fused", 50, 3, tq)
        host1 = time_forward.host_issue_s
        t1_r, host1_r, ev_r1, stage_r1 = time_replay(10, 32, ops1, 50, 3, tq)
        print(json.dumps({"small_circuit": {"fused_ms_per_circuit_batch": t1_f * 1e3, "fused_host_issue_ms": host1 * 1e3,
                                            "replay_ms_per_circuit_batch": t1_r * 1e3,
                                            "replay_host_issue_ms": host1_r * 1e3, "replay_stage_kernel_ms": stage_r1,
                                            "max_abs_diff": float((ev_r1 - ev_f1).abs().max())}}), flush=True)
        return
    if args.only_gradient:
        an, ab = args.adj_qubits, args.adj_bsz
        rec = []
        dt, loss, _ = time_gradient(an, ab, 10, "fused", args.steps, 1, tq, timer=rec)
        amps = float(ab) * (1 << an)
        print(json.dumps({"fused_gradient_s": dt, "workload": f"{an} qubits, bsz {ab}", "loss": loss,
                          "adjoint_stage": stage_launch_stats(rec, args.steps, "a", amps, 32.0),
                          "pruned_adjoint_stage": stage_launch_stats(rec, args.steps, "a_pruned", amps, 32.0),
                          "final_adjoint_stage": stage_launch_stats(rec, args.steps, "a_read", amps, 16.0),
                          "forward_stage": stage_launch_stats(rec, args.steps, "f", amps, 16.0)}), flush=True)
        return
    if args.only_fused or args.only_pergate:
        n, bsz = 20, 16
        mode = "fused" if args.only_fused else "pergate"
        ops = config2_ops(seed=0, n=n, layers=20)
        dt, _ = time_forward(n, bsz, ops, mode, args.steps, args.warmup, tq)
        print(json.dumps({f"{mode}_forward_ms": dt * 1e3, "workload": f"{n} qubits, bsz {bsz}, RY/CZ circuit",
                          "amp_updates_per_s": len(ops) * bsz * (1 << n) / dt}), flush=True)
        return
    if args.config5:
        print(json.dumps({"config5": config5(args, tq)}), flush=True)
        return
    if args.config5_plans:
        print(json.dumps({"config5_plans": config5_plans(args, tq)}), flush=True)
        return

    # ---------------------------------------------------------------- headline: adjoint gradient
    an, ab = args.adj_qubits, args.adj_bsz
    amps = float(ab) * (1 << an)
    log(f"[bench] fused adjoint gradient, {an} qubits bsz {ab} (warm-up compiles the stage kernels)")
    rec = []
    dt_af, loss_f, g_f = time_gradient(an, ab, 10, "fused", args.steps, args.warmup, tq, timer=rec)
    log(f"[bench] fused {dt_af:.4f} s/gradient")
    adj = stage_launch_stats(rec, args.steps, "a", amps, 32.0)
    adj_last = stage_launch_stats(rec, args.steps, "a_read", amps, 16.0)
    adj_pruned = stage_launch_stats(rec, args.steps, "a_pruned", amps, 32.0)
    fwd = stage_launch_stats(rec, args.steps, "f", amps, 16.0)
    del rec
    torch.cuda.empty_cache()
    log("[bench] per-gate adjoint gradient")
    dt_ap, loss_p, g_p = time_gradient(an, ab, 10, "pergate", args.pergate_steps, 1, tq)
    log(f"[bench] per-gate {dt_ap:.3f} s/gradient")
    log("[bench] fused gradient with the program-interpreting stage kernels")
    dt_ai, _, _ = time_gradient(an, ab, 10, "fused", 1, 1, tq, jit="off")
    torch.cuda.empty_cache()
    log("[bench] dense 1-qubit gate roofline (same state size, and a cache-resident one)")
    ms_d, nbytes_d = kernel_roofline(an, ab, iters=2 * an)
    ms_c, nbytes_c = kernel_roofline(20, 16, iters=60)
    torch.cuda.empty_cache()
    n_params = len(g_f)
    traffic = load_traffic("adjoint_stage_24q_traffic") or {}
    extra = {
        "pergate_gradient": {"grads_per_s": 1.0 / dt_ap, "s_per_gradient": dt_ap,
                             "fused_speedup_vs_pergate": dt_ap / dt_af,
                             "grad_max_abs_diff_fused_vs_pergate": float((g_f - g_p).abs().max())},
        "fused_interpreting_stage_kernels": {"grads_per_s": 1.0 / dt_ai, "s_per_gradient": dt_ai},
        "forward_stage": None if fwd is None else {**fwd, "frac": fwd["achieved_GBs"] / HBM_PEAK_GBS,
                                                   "kernel": "compiled forward stage tqsv_jf_* (16 B/amplitude)"},
        "pruned_adjoint_stages": None if adj_pruned is None else {
            **adj_pruned, "kernel": "pruned compiled adjoint stages tqsv_jp_* (ops on wires no earlier trainable op "
                                    "touches are not un-applied; psi and lam read and written: 32 B/amplitude)"},
        "final_adjoint_stage": None if adj_last is None else {
            **adj_last, "frac": adj_last["achieved_GBs"] / HBM_PEAK_GBS,
            "kernel": "last stage of the pruned reverse sweep (no input-state gradient): reads psi and lam, "
                      "stores nothing (16 B/amplitude)"},
        "dense_gate": {"kernel": "dense 1-qubit gate k_1q_t0/k_1q_tile/k_1q_pair, RY cycling over every target "
                                 f"bit, {an} qubits bsz {ab} ({nbytes_d / 2**30:.0f} GiB, HBM-resident)",
                       "avg_launch_ms": ms_d, "achieved": nbytes_d / (ms_d * 1e-3) / 1e9,
                       "frac": nbytes_d / (ms_d * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "cache_resident_20q_bsz16": {
                           "avg_launch_ms": ms_c, "achieved": nbytes_c / (ms_c * 1e-3) / 1e9,
                           "note": "256 MiB state held by the 256 MB Infinity Cache: not an HBM figure"}},
    }

    if not args.skip_extras:
        n, bsz = 20, 16
        ops = config2_ops(seed=0, n=n, layers=20)
        upd = len(ops) * bsz * (1 << n)
        log("[bench] 20-qubit RY/CZ circuit: per-gate and fused forward")
        dt_pg, ev_pg = time_forward(n, bsz, ops, "pergate", 10, 3, tq)
        dt_f, ev_f = time_forward(n, bsz, ops, "fused", 10, 3, tq)
        dt_fi, _ = time_forward(n, bsz, ops, "fused", 10, 3, tq, jit="off")
        extra["config2"] = {"workload": "20 qubits, bsz 16, 20 x [RY all wires, CZ even+odd pairs], <Z_i> (780 gates)",
                            "pergate_amp_updates_per_s": upd / dt_pg, "pergate_ms_per_step": dt_pg * 1e3,
                            "fused_amp_updates_per_s": upd / dt_f, "fused_ms_per_step": dt_f * 1e3,
                            "fused_speedup_vs_pergate": dt_pg / dt_f,
                            "fused_interpreting_ms_per_step": dt_fi * 1e3,
                            "expval_max_abs_diff_fused_vs_pergate": float((ev_pg - ev_f).abs().max())}
        from circuits import config1_ops
        ops1 = config1_ops(seed=0)
        upd1 = len(ops1) * 32 * (1 << 10)
        log("[bench] 10-qubit bsz-32 circuit (160 gates)")
        t1_pg, _ = time_forward(10, 32, ops1, "pergate", 20, 3, tq)
        t1_f, _ = time_forward(10, 32, ops1, "fused", 50, 3, tq)
        host1 = time_forward.host_issue_s
        rec1 = []
        from torchquantum_amd import engine as E
        E.set_stage_timer(rec1)
        time_forward(10, 32, ops1, "fused", 20, 0, tq)
        E.set_stage_timer(None)
        torch.cuda.synchronize()
        stage_ms = sum(e0.elapsed_time(e1) for _, _, e0, e1 in rec1) / 20
        _, ev_f1 = time_forward(10, 32, ops1, "fused", 1, 0, tq)
        t1_r, host1_r, ev_r1, stage_r1 = time_replay(10, 32, ops1, 50, 3, tq)
        extra["small_circuit"] = {"workload": "10 qubits, bsz 32, 8 x [U3 all wires, CNOT ring], <Z_i> (160 gates)",
                                  "pergate_ms_per_circuit_batch": t1_pg * 1e3,
                                  "fused_ms_per_circuit_batch": t1_f * 1e3, "fused_amp_updates_per_s": upd1 / t1_f,
                                  "fused_host_issue_ms": host1 * 1e3, "fused_stage_kernel_ms": stage_ms,
                                  "replay": {"api": "tq.CircuitReplay (recorded once, compiled stage kernel)",
                                             "ms_per_circuit_batch": t1_r * 1e3, "host_issue_ms": host1_r * 1e3,
                                             "stage_kernel_ms": stage_r1, "amp_updates_per_s": upd1 / t1_r,
                                             "expval_max_abs_diff_vs_functional":
                                                 float((ev_r1 - ev_f1).abs().max())}}
        log("[bench] dense k = 5, 6 blocks")
        extra["dense_block"] = dense_block_bench()
        log("[bench] encoder + ansatz training steps")
        extra["encoder_training"] = {"small": encoder_training(tq, 10, 256, 4, 20, 3),
                                     "large": encoder_training(tq, 22, 64, 4, 3, 2)}
        torch.cuda.empty_cache()

    batch_out = None
    if not args.skip_batch and not args.skip_extras:
        log("[bench] uneven batch (256 circuits)")
        extra["batch"], batch_out = time_batch(256, 5, tq)
        log(f"[bench] batch graph {extra['batch']['ms_graph_replay']:.2f} ms")
        torch.cuda.empty_cache()

    if not args.skip_config5 and not args.skip_extras:
        extra["config5"] = config5(args, tq)
        torch.cuda.empty_cache()

    cpu = None
    if not args.skip_cpu:
        log(f"[bench] cpu baseline: reference path + autograd, {args.cpu_grad_gates} gates at {an} q bsz 1")
        rate, threads, cdt, total = cpu_baseline_gradient(an, ab, 10, args.cpu_grad_gates)
        cpu = {"value": rate, "unit": "gradient evaluations/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"first {args.cpu_grad_gates} of the {total} gates at {an} qubits bsz 1, forward + "
                         f"torch autograd backward through the restated reference bmm path "
                         f"(gate_wrapper.py:96-140, complex64), loss = ZZ ring; {cdt:.1f} s, scaled by "
                         f"{total}/{args.cpu_grad_gates} gates x bsz {ab}"}
        if not args.skip_extras:
            sample = config2_ops(seed=0, n=20, layers=20)[: 39 * args.cpu_layers]
            r2, _, c2 = cpu_baseline(sample, 20, 16, 780)
            cpu["config2_forward"] = {"amp_updates_per_s": r2, "sample": f"first {len(sample)} gates of the "
                                      f"780-gate circuit, 20 qubits bsz 16, forward ({c2:.1f} s)"}
            from circuits import config1_ops
            r1, _, c1 = cpu_baseline(config1_ops(seed=0), 10, 32, 160)
            cpu["small_circuit"] = {"ms_per_circuit_batch": c1 * 1e3, "amp_updates_per_s": r1}
        log("[bench] oracle errors")
        extra["oracle_error"] = oracle_errors(tq, batch_out)
        log("[bench] full-size parameter-shift check")
        extra["oracle_error"][f"gradient at {an} q bsz {ab}: fused adjoint vs parameter shift"] = \
            parameter_shift_check(an, ab, 10, tq)

    line = {
        "metric": "amplitude updates/s & adjoint grads/s vs plain per-gate kernel, % HBM peak",
        "value": 1.0 / dt_af,
        "unit": "gradient evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt_af * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64 (fp32)", "data": "synthetic (seeded angles)",
        "config": {"workload": f"{an} qubits, bsz {ab}, 10 x [RX, RY, RZ all wires, CNOT ring] + per-row RY input "
                               f"layer, loss = mean_b [sum <Z_i Z_i+1> + 0.5 sum <X_i>], adjoint gradient of "
                               f"{n_params} parameters, fused stages", "n_qubits": an, "bsz": ab,
                   "params": n_params, "loss": loss_f},
        "roofline": None if adj is None else {
            "bound": "hbm", "achieved": adj["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": adj["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": traffic.get("bytes_per_launch"),
            "kernel": "compiled adjoint stage tqsv_ja_* (reads psi and lam, writes both: 32 B per amplitude, "
                      f"{adj['bytes_per_launch'] / 1e9:.2f} GB per launch)",
            "avg_launch_ms": adj["avg_launch_ms"], "launches_per_step": adj["launches_per_step"],
            "share_of_step": adj["ms_per_step"] / (dt_af * 1e3)},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
