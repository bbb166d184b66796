"""Code generation of the run-time compiled stage kernels, checked on the CPU
(hipRTC cross-compiles for gfx950 without a GPU): the generated HIP compiles,
carries no packed-FP32 instructions, and reflects the stage programs it was
generated from.  Their numerics are checked on the GPU (test_gpu_jit.py)."""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from circuits import apply_ops_tq, random_ops

NS = 16          # amplitudes per lane in a stage kernel


def _plan(n, bsz, ops, **kw):
    import torchquantum_amd as tq
    from torchquantum_amd import engine as E
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cpu")
    apply_ops_tq(q, ops, "cpu", requires_grad=True)
    return E.Compiled(q._tape, n, bsz, tq.EngineOptions(**kw), "cpu")


def _stages(c):
    return [st for kind, (base, st) in [(k, p) for k, p in c.steps if k == "stage"]]


def _objdump():
    for p in (shutil.which("llvm-objdump"), "/opt/rocm/llvm/bin/llvm-objdump"):
        if p and Path(p).exists():
            return p
    return None


def test_generated_stages_compile(tmp_path):
    """Forward and adjoint kernels of a random 12-qubit circuit with dense
    3-bit blocks, controls and per-row parameters compile for gfx950; the
    code object has no packed-FP32 instructions."""
    from torchquantum_amd import stagejit as J
    rng = np.random.default_rng(3)
    ops = random_ops(12, 40, 2, rng)
    c = _plan(12, 2, ops, fusion_width=3)
    stages = [s for s in _stages(c) if J.supported(s.words)]
    assert stages
    # the stage with the most distinct op kinds
    def kinds(s):
        return len({(w[0] & 15, w[0] & ~0xF0) for w in J.StageProgram(s.words).ops})
    pick = [max(stages, key=kinds)]
    src = J.PRELUDE + "".join(J.gen_forward(s.words, f"kf{i}") + J.gen_adjoint(s.words, f"ka{i}")
                              for i, s in enumerate(pick))
    code = J.compile_source(src)
    assert len(code) > 1000
    od = _objdump()
    if od is None:
        pytest.skip("llvm-objdump not found")
    f = tmp_path / "stage.co"
    f.write_bytes(code)
    dis = subprocess.run([od, "-d", "--mcpu=gfx950", str(f)], capture_output=True, text=True, check=True).stdout
    assert "v_fma_f32" in dis or "v_fmac_f32" in dis
    assert "v_pk_fma_f32" not in dis and "v_pk_mul_f32" not in dis and "v_pk_add_f32" not in dis


def test_compile_error_is_reported():
    from torchquantum_amd import stagejit as J
    with pytest.raises(RuntimeError, match="compile failed"):
        J.compile_source('extern "C" __global__ void k() { undefined_thing(); }')


def test_supported_needs_full_waves():
    """Stages with fewer than 64 threads per group (m < 10) stay on the
    interpreting kernels."""
    from torchquantum_amd import stagejit as J
    rng = np.random.default_rng(4)
    small = _plan(8, 1, random_ops(8, 10, 1, rng))
    assert not any(J.supported(s.words) for s in _stages(small))
    big = _plan(11, 1, random_ops(11, 10, 1, rng))
    assert all(J.supported(s.words) for s in _stages(big))


def test_generated_code_follows_program():
    """Structure of the generated forward code: one LDS round trip per remap
    op, no selects for uncontrolled ops, renaming (no code) for an X on a
    register bit without thread-bit controls."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E, stagejit as J
    n = 12
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    for w in range(n):
        tqf.ry(q, w, params=torch.tensor([[0.1 * (w + 1)]]))
    c = E.Compiled(q._tape, n, 1, tq.EngineOptions(), "cpu")
    for s in _stages(c):
        sp = J.StageProgram(s.words)
        src = J.gen_forward(s.words, "k")
        n_remap = sum(1 for w in sp.ops if (w[0] & 15) == E.OP_REMAP)
        # the body appears twice: the |0..0>-start loop and the loop that reads
        assert src.count("__syncthreads();") == 4 * n_remap
        assert "csel" not in src
    # a lone X on wire 0 (bit 11, a register bit of the canonical layout):
    # no instructions, the stores name the loaded values exchanged in pairs
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    tqf.paulix(q, 0)
    c = E.Compiled(q._tape, n, 1, tq.EngineOptions(merge_1q=False), "cpu")
    (s,) = _stages(c)
    sp = J.StageProgram(s.words)
    assert [w[0] & 15 for w in sp.ops] == [E.OP_MAT1]
    K = sp.ops[0][3]
    src = J.gen_forward(s.words, "k")
    assert "const float2 v" not in src
    mode = J._mem_mode(sp, False, J.CODEGEN["buf"])
    for j in range(16):
        assert J._store(sp, "out", j, f"i{j ^ (1 << K)}", mode) in src


def test_code_object_disk_cache(tmp_path, monkeypatch):
    """A compiled source is stored under TQSV_JIT_CACHE and read back."""
    from torchquantum_amd import stagejit as J
    monkeypatch.setenv("TQSV_JIT_CACHE", str(tmp_path))
    src = 'extern "C" __global__ void k(float* p) { p[threadIdx.x] *= 2.f; }\n'
    a = J.compile_source(src)
    files = list(tmp_path.glob("*.co"))
    assert len(files) == 1 and files[0].read_bytes() == a
    files[0].write_bytes(b"cached")          # a second compile returns the stored bytes
    assert J.compile_source(src) == b"cached"
    monkeypatch.setenv("TQSV_JIT_CACHE", "off")
    assert J.compile_source(src) == a


def test_pruned_adjoint_code():
    """Pruned adjoint kernels (engine.Compiled.prune_info): the ops at the
    skip positions keep their gradient partials but are not un-applied, and
    the sweep's last stage stores nothing; both compile for gfx950."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E, stagejit as J
    n, bsz = 12, 2
    q = tq.QuantumDevice(n_wires=n, bsz=bsz, device="cpu")
    for w in range(n):
        tqf.ry(q, w, params=torch.rand(bsz, 1))                       # input layer, no gradient
    for w in range(n):
        tqf.rx(q, w, params=torch.rand(1, 1, requires_grad=True))
    for w in range(n - 1):
        tqf.cnot(q, [w, w + 1])
    for w in range(n):
        tqf.ry(q, w, params=torch.rand(1, 1, requires_grad=True))
    c = E.Compiled(q._tape, n, bsz, tq.EngineOptions(), "cpu")
    stop, skip = c.prune_info()
    assert skip and stop < len(c.steps)
    si = min(skip)
    st = c.steps[si][1][1]
    full = J.gen_adjoint(st.words, "kf")
    pruned = J.gen_adjoint(st.words, "kp", skip=skip[si], no_store=(si == stop))
    # the merged input + first-layer rotations keep their sums (same number
    # of gradient accumulations) but lose their U^dagger products
    assert pruned.count("atomicAdd(&gacc") == full.count("atomicAdd(&gacc")
    assert pruned.count("cmulc(") < full.count("cmulc(")
    if si == stop:
        assert "tq_st(" not in pruned and "tq_st(" in full
    assert J.kernel_name("p", (st.words, skip[si], si == stop)) != J.kernel_name("a", st.words)
    code = J.compile_source(J.PRELUDE + pruned, J.COMPILE_OPTS["p"])
    assert len(code) > 1000


def _lane_positions(src, var):
    """Bit positions of `const unsigned <var> = ((((unsigned)tid >> i) & 1u) << p) | ...;` in thread-bit order."""
    m = re.search(rf"const unsigned {var} = ([^;]*);", src)
    assert m, var
    return [int(p) for _, p in sorted((int(i), p) for i, p in re.findall(r"tid >> (\d+)\) & 1u\) << (\d+)\)", m.group(1)))]


def _lane_offsets(positions, nthr):
    tid = np.arange(nthr)
    out = np.zeros(nthr, dtype=np.int64)
    for i, p in enumerate(positions):
        out |= ((tid >> i) & 1) << p
    return out


def test_load_store_layouts_follow_program():
    """Stages that load in the layout of their first ops or store in that of
    their last (H_FLAGS masks): the generated forward code reads in the load
    layout and writes in the store layout, the adjoint code the other way
    round, each lane and slot at the offsets the program emulator uses
    (tests/emulator.py layout_offsets); the kernels compile for gfx950."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    import emulator as EM
    from circuits import config3_ops
    from torchquantum_amd import engine as E, stagejit as J
    n = 16
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    for op in config3_ops(seed=3, n=n, layers=3):
        p = op["params"]
        tqf.func_name_dict[op["name"]](q, op["wires"], params=None if p is None else
                                       torch.tensor(np.asarray(p, dtype=np.float32).reshape(1, -1),
                                                    requires_grad=True))
    c = E.Compiled(q._tape, n, 1, tq.EngineOptions(stage_bits=12), "cpu")
    picked = [s for s in _stages(c) if int(s.words[E.H_FLAGS]) >> E.FL_LOAD_SHIFT]
    seen = set()
    srcs = []
    for s in picked:
        sp = J.StageProgram(s.words)
        seen |= {k for k, mk in (("ld", sp.lmask), ("st", sp.smask)) if mk}
        nthr = 1 << sp.T
        want = {k: EM.layout_offsets(s.words, k) for k in ("ld", "st")}
        var = {"toff": "ld", "toffs": "st"}
        fwd = J.gen_forward(s.words, f"kf{len(srcs)}")
        adj = J.gen_adjoint(s.words, f"ka{len(srcs)}")
        srcs += [fwd, adj]
        lanes = {v: _lane_offsets(_lane_positions(src, v), nthr) for src in (fwd,)
                 for v in ("toff", "toffs") if f"const unsigned {v} =" in src}
        assert ("toffs" in lanes) == (sp.lmask != sp.smask)

        def check(src, pattern, lay):
            hits = re.findall(pattern, src)
            assert len(hits) >= NS and len(hits) % NS == 0, pattern
            for k in range(0, len(hits), NS):
                for j, (v, const) in enumerate(hits[k:k + NS]):
                    assert var[v] == lay or (v == "toff" and sp.lmask == sp.smask)
                    assert np.array_equal(lanes[v], want[lay][0])
                    assert int(const) == 8 * int(want[lay][1][j])
        check(fwd, r"const float2 q\d+ = tq_ld\(r_inn, (toffs?)8, (\d+)\)", "ld")
        check(fwd, r"tq_st\(r_out, (toffs?)8, (\d+), ", "st")
        check(adj, r"const float2 p\d+ = tq_ld\(r_psi_in, (toffs?)8, (\d+)\)", "st")
        check(adj, r"tq_st\(r_lam, (toffs?)8, (\d+), ", "ld")
    assert seen == {"ld", "st"}
    code = J.compile_source(J.PRELUDE + "".join(srcs[:4]), J.COMPILE_OPTS["a"])
    assert len(code) > 1000



def test_matrix_core_products_in_forward_code():
    """Forward stages run the 2x2 products of single-wire ops without
    thread-bit controls on the matrix cores (tq_mm, one lane row each):
    present for the merged RX.RY.RZ rotations, absent with mfma=False and
    for ops with thread-bit controls; the code compiles for gfx950."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E, stagejit as J
    n = 12
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    for w in range(n):
        for g in ("rx", "ry", "rz"):
            tqf.func_name_dict[g](q, w, params=torch.tensor([[0.1 * (w + 1)]]))
    c = E.Compiled(q._tape, n, 1, tq.EngineOptions(), "cpu")
    (s,) = _stages(c)
    on = J.gen_forward(s.words, "kon", mfma=True)
    off = J.gen_forward(s.words, "koff", mfma=False)
    n_mat1 = sum(1 for w in J.StageProgram(s.words).ops if (w[0] & 15) == E.OP_MAT1)
    # 8 pairs per op; the body appears twice (the |0..0>-start loop and the loop that reads)
    assert on.count("tq_mm(") == 16 * n_mat1 and "tq_mm(" not in off
    assert "tq_row_u(" in on and "mhi" in on
    # a controlled rotation whose control is a thread bit: VALU code
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    tqf.crx(q, [n - 1, 0], params=torch.tensor([[0.3]]))
    c = E.Compiled(q._tape, n, 1, tq.EngineOptions(merge_1q=False), "cpu")
    (s2,) = _stages(c)
    assert "tq_mm(" not in J.gen_forward(s2.words, "kc", mfma=True)
    assert len(J.compile_source(J.PRELUDE + on)) > 1000


def _xl_model(b, x0, x1):
    """What tq_xl<b> leaves in each of the 64 lanes (the semantics the
    permlane / DPP helpers implement): a lane with bit b set takes x1 of
    lane l - 2^b into x0, a lane with bit b clear takes x0 of lane l + 2^b
    into x1."""
    n0, n1 = x0.copy(), x1.copy()
    for l in range(64):
        if (l >> b) & 1:
            n0[l] = x1[l - (1 << b)]
        else:
            n1[l] = x0[l + (1 << b)]
    return n0, n1


def _local_index(layout, T, tid, j):
    return (sum(layout[i] for i in range(T) if (tid >> i) & 1)
            + sum(layout[T + k] for k in range(4) if (j >> k) & 1))


def _lane_plan_holds(T, F, G):
    """Run lane_plan's exchanges on every wave of a 2^T-thread group of local
    indices; True if each (thread, slot) ends holding the local index that
    layout G assigns it (what the LDS remap produces)."""
    from torchquantum_amd import stagejit as J
    steps, perm = J.lane_plan(T, F, G)
    for w in range(1 << (T - 6)):
        tids = [64 * w + l for l in range(64)]
        regs = [np.array([_local_index(F, T, t, j) for t in tids]) for j in range(NS)]
        for b, sb in steps:
            for j0 in range(NS):
                if not j0 & (1 << sb):
                    regs[j0], regs[j0 | (1 << sb)] = _xl_model(b, regs[j0], regs[j0 | (1 << sb)])
        want = [np.array([_local_index(G, T, t, j) for t in tids]) for j in range(NS)]
        if not all((regs[perm[j]] == want[j]).all() for j in range(NS)):
            return False
    return True


def test_lane_remap_plan_matches_lds_remap():
    """Remaps that keep the wave bits and only exchange lane bits with
    register slots run as cross-lane moves in compiled adjoint stages; the
    exchange sequence reproduces the LDS remap's permutation on random
    layouts, and every other remap is left to the LDS."""
    import random
    from torchquantum_amd import stagejit as J
    rnd = random.Random(1)
    seen, lane_bits = 0, set()
    for _ in range(600):
        T = rnd.choice([6, 7, 9])
        m = T + 4
        F = [1 << i for i in range(m)]
        rnd.shuffle(F)
        G = list(F)
        lanes = rnd.sample(range(6), rnd.randint(1, 4))
        for b, r in zip(lanes, rnd.sample(range(T, m), len(lanes))):
            G[b], G[r] = G[r], G[b]
        regs = G[T:]
        rnd.shuffle(regs)
        G[T:] = regs
        wave_moved = T > 6 and rnd.random() < 0.2
        if wave_moved:
            i = rnd.randrange(6, T)
            r = rnd.randrange(T, m)
            G[i], G[r] = G[r], G[i]
        plan = J.lane_plan(T, F, G)
        if wave_moved or all(F[i] == G[i] for i in range(6)):
            assert plan is None
            continue
        assert plan is not None
        assert _lane_plan_holds(T, F, G)
        seen += 1
        lane_bits |= {b for b, _ in plan[0]}
    assert seen > 300 and lane_bits == set(range(6))


def test_lane_remaps_in_adjoint_code():
    """Adjoint stages of the 24-qubit gradient ansatz (13-bit stages): the
    wave-local remaps become permlane / DPP exchanges (fewer barriers than
    remaps through the LDS), while the forward kernels keep the LDS."""
    import torchquantum_amd as tq
    import torchquantum_amd.functional as tqf
    from torchquantum_amd import engine as E, stagejit as J
    from circuits import config3_ops
    n = 24
    q = tq.QuantumDevice(n_wires=n, bsz=1, device="cpu")
    for w in range(n):
        tqf.ry(q, w, params=torch.full((1, 1), 0.3))
    for op in config3_ops(seed=0, n=n, layers=2):
        p = op["params"]
        tqf.func_name_dict[op["name"]](q, op["wires"], params=None if p is None else
                                       torch.tensor(np.asarray(p, dtype=np.float32), requires_grad=True))
    c = E.Compiled(q._tape, n, 64, tq.EngineOptions(stage_bits=13), "cpu")
    found = 0
    for st in _stages(c):
        sp = J.StageProgram(st.words)
        lane = [w for w in sp.ops if (w[0] & 15) == E.OP_REMAP
                and J.lane_plan(sp.T, *reversed(sp.remap_layouts(w))) is not None]
        if not lane:
            continue
        found += 1
        adj = J.gen_adjoint(st.words, "k")
        adj_lds = J.gen_adjoint(st.words, "k", lane_remap=False)
        assert "tq_xl" in adj and "tq_xl" not in adj_lds
        # an LDS remap has two barriers per array (psi, lam); a lane remap none
        assert adj.count("__syncthreads();") == adj_lds.count("__syncthreads();") - 4 * len(lane)
        assert "tq_xl" not in J.gen_forward(st.words, "k")
    assert found >= 2
