"""Run-time compiled stage kernels: straight-line HIP code per stage program.

The interpreter kernels (``k_stage_fwd`` / ``k_stage_adj`` in
csrc/tqsv_stage.hip) decode every op of a stage program at run time.  This
module turns one stage program -- the same int32 words the planner builds in
``engine.encode_stage`` -- into HIP source in which every op is unrolled code
with its register bits, controls and matrix offsets as constants, compiles it
for gfx950 with hipRTC (``tqsv_jit_compile``) and launches it through
``tqsv_jit_stage_forward`` / ``tqsv_jit_stage_adjoint``.

What the generated code computes is what the interpreter computes, op for op
(the per-gate bmm chain of torchquantum/functional/gate_wrapper.py:96-140 and,
for the adjoint, its autograd backward :127-136).  What changes is how:

* the slot -> variable assignment is tracked while generating, so an
  exchange (X / CNOT on a register bit with no thread-bit control) is a pure
  renaming and costs no instruction;
* a control on thread bits selects the identity matrix per lane once per op
  (8 selects) instead of selecting every result (4 per amplitude pair);
* register-bit controls drop the pairs that fail them at generation time;
* the LDS addresses of every remap are hoisted out of the group loop.
"""
from __future__ import annotations

import atexit
import ctypes
import hashlib
import os
import queue
import threading
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .engine import (H_NOPS, H_M, H_T, H_NNL, H_OFF_OPS, H_NGP, H_OFF_GSLOT, H_FLAGS, H_POS, OP_MAT1, OP_MAT2,
                     OP_DIAG, OP_MATK, OP_REMAP, F_BATCHED, F_NOCTRL, F_REAL, F_XPERM, F_SIGN, F_MG, R_BITS,
                     FL_LOAD_SHIFT, FL_STORE_SHIFT)

NS = 1 << R_BITS
MIN_T = 6            # compiled stages need full 64-lane waves
KERNELS_PER_SOURCE = 4
# Extra compiler options per kind ("f" forward, "a" adjoint).  The adjoint
# kernels hold psi and lam (64 VGPRs of state) plus the matrix-gradient
# accumulators; LLVM's occupancy-driven scheduler brings them from 200-230
# to 160-190 VGPRs.  Measured on the 24-qubit gradient: 531 -> 500 ms.  The
# forward kernels gain nothing from it.
COMPILE_OPTS = {"f": "", "a": "-mllvm -amdgpu-sched-strategy=iterative-maxocc"}
COMPILE_OPTS["p"] = COMPILE_OPTS["a"]     # pruned adjoint stages (gen_adjoint with skip / no_store)
# Code-generation defaults (gen_forward / gen_adjoint arguments of the same
# names), chosen by timing the stages of the 24-qubit bsz-64 gradient on
# MI355X: buffer instructions for the state (no per-access 64-bit address
# arithmetic), the padded remap buffer (one address register per layout
# instead of one per slot: 99 instead of 142 VGPRs on an adjoint stage, so
# more waves per SIMD), next-group loads issued ahead in forward stages only
# (in adjoint stages their 64 registers cost more occupancy than the overlap
# gains), no register accumulation of gradient partials (same trade).
# mfma_f / mfma_a: the 2x2 complex products of single-wire ops without
# thread-bit controls (forward U, adjoint U^dagger on psi and lam) run on the
# matrix cores (tq_mm).  Forward stages: 18.7 -> 18.0 ms over five stages of
# the gradient workload.  Adjoint stages: half the vector instructions, but
# 176 VGPRs + 64 AGPRs against 106 (or spills when held to 128), and
# 42.1 -> 44.2 ms; they stay on the vector ALUs.
# lane_remap_f / _a: a remap that leaves the wave bits alone and only
# exchanges lane bits with register slots runs as cross-lane register moves
# (permlane swaps for lane bits 4 and 5, DPP for 0-3): no LDS traffic and no
# barrier.  Adjoint stages of the gradient workload (13-bit stages, six
# stages): 62.1 -> 60.8 ms; forward stages, which remap one array instead of
# two: 25.1 -> 25.4 ms, so they keep the LDS.
CODEGEN = {"buf": True, "prefetch_f": True, "prefetch_a": False, "lds": "pad4", "racc": 0,
           "mfma_f": True, "mfma_a": False, "lane_remap_f": False, "lane_remap_a": True}

PRELUDE = r"""
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ float2 cfmac(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}
// Buffer loads and stores of the state: a scalar resource (base address of a
// group), the lane's byte offset and a constant scalar offset per slot.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tq_rsrc(const float2* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(p), (short)0, -1, 0x00020000);
}
__device__ __forceinline__ float2 tq_ld(__amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
__device__ __forceinline__ void tq_st(__amdgpu_buffer_rsrc_t r, unsigned vo, int so, float2 x) {
  typedef unsigned tq_v2u __attribute__((__vector_size__(8)));
  const tq_v2u v = {__float_as_uint(x.x), __float_as_uint(x.y)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, vo, so, 0);
}
__device__ __forceinline__ float redot(float2 a, float2 b) { return fmaf(a.x, b.x, a.y * b.y); }
// A 2x2 complex matrix times this lane's amplitude pair on the matrix cores.
// v_mfma_f32_4x4x1_16b (16 blocks of 4 lanes): output register i of lane
// 4b + j accumulates A(lane 4b + i) * B(lane 4b + j) -- a layout checked on
// MI355X -- so with A = row (lane & 3) of the 4x4 real form of the matrix
// (one k-step per column) and B = component k of this lane's own pair
// (Re x0, Im x0, Re x1, Im x1), the lane receives its pair's product.  The
// four lanes of a block must share the matrix (no per-lane controls).
typedef float tq_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ tq_f4 tq_mm(float m0, float m1, float m2, float m3, float2 x0, float2 x1) {
  tq_f4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m0, x0.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m1, x0.y, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m2, x1.x, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m3, x1.y, acc, 0, 0, 0);
  return acc;
}
// Row (2 hi + odd) of the real form of U = [[u0, u1], [u2, u3]] acting on
// (Re x0, Im x0, Re x1, Im x1): rows Re y0, Im y0, Re y1, Im y1.
__device__ __forceinline__ tq_f4 tq_row_u(float2 u0, float2 u1, float2 u2, float2 u3, bool hi, bool odd) {
  const float2 p = hi ? u2 : u0, q = hi ? u3 : u1;
  tq_f4 r;
  r[0] = odd ? p.y : p.x;
  r[1] = odd ? p.x : -p.y;
  r[2] = odd ? q.y : q.x;
  r[3] = odd ? q.x : -q.y;
  return r;
}
// The same for U^dagger: y0 = conj(u0) x0 + conj(u2) x1, y1 = conj(u1) x0 + conj(u3) x1.
__device__ __forceinline__ tq_f4 tq_row_udag(float2 u0, float2 u1, float2 u2, float2 u3, bool hi, bool odd) {
  const float2 p = hi ? u1 : u0, q = hi ? u3 : u2;
  tq_f4 r;
  r[0] = odd ? -p.y : p.x;
  r[1] = odd ? p.x : p.y;
  r[2] = odd ? -q.y : q.x;
  r[3] = odd ? q.x : q.y;
  return r;
}
__device__ __forceinline__ float2 csel(bool c, float2 a, float2 b) {
  return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}
// Cross-lane sums without the LDS crossbar (gfx950): v_permlane32_swap
// exchanges the upper 32 lanes of a with the lower 32 lanes of b, so a' + b'
// holds, in lanes 0-31, a summed over lanes (l, l + 32) and, in lanes 32-63,
// the same for b.  v_permlane16_swap does this for the odd 16-lane rows of a
// and the even rows of b.  DPP (row_ror:8, row_half_mirror, quad_perm) then
// reduces within a 16-lane row.
__device__ __forceinline__ float sum32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float sum16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float dpp_ror8(float v) { return __builtin_amdgcn_update_dpp(0.f, v, 0x128, 0xf, 0xf, false); }
// all-reduce over the 8-lane halves of each row
__device__ __forceinline__ float allsum8(float v) {
  v += __builtin_amdgcn_update_dpp(0.f, v, 0x141, 0xf, 0xf, false);   // row_half_mirror: l <-> 7 - l
  v += __builtin_amdgcn_update_dpp(0.f, v, 0x4E, 0xf, 0xf, false);    // quad_perm [2,3,0,1]: l ^ 2
  v += __builtin_amdgcn_update_dpp(0.f, v, 0xB1, 0xf, 0xf, false);    // quad_perm [1,0,3,2]: l ^ 1
  return v;
}
// Exchange of lane bit b with a register slot bit.  (x0, x1) are the two
// slots that differ in that slot bit; afterwards lane bit b indexes what the
// slot bit did and vice versa: a lane with bit b set takes x1 of lane
// l - 2^b into x0, a lane with bit b clear takes x0 of lane l + 2^b into x1.
// b = 5 / 4: v_permlane32_swap / v_permlane16_swap exchange exactly these
// halves; b = 3 / 2: DPP row_shr / row_shl by 2^b with a bank mask keeping
// the lanes that stay; b = 1 / 0: quad_perm (l ^ 2^b) and a select.
struct tq_p2 { float2 a, b; };
__device__ __forceinline__ tq_p2 tq_xl5(float2 x0, float2 x1) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x0.x), __float_as_uint(x1.x), false, false);
  const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(x0.y), __float_as_uint(x1.y), false, false);
  return {make_float2(__uint_as_float(r[0]), __uint_as_float(q[0])), make_float2(__uint_as_float(r[1]), __uint_as_float(q[1]))};
}
__device__ __forceinline__ tq_p2 tq_xl4(float2 x0, float2 x1) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x0.x), __float_as_uint(x1.x), false, false);
  const auto q = __builtin_amdgcn_permlane16_swap(__float_as_uint(x0.y), __float_as_uint(x1.y), false, false);
  return {make_float2(__uint_as_float(r[0]), __uint_as_float(q[0])), make_float2(__uint_as_float(r[1]), __uint_as_float(q[1]))};
}
__device__ __forceinline__ tq_p2 tq_xl3(float2 x0, float2 x1) {
  return {make_float2(__builtin_amdgcn_update_dpp(x0.x, x1.x, 0x118, 0xf, 0xc, false),
                      __builtin_amdgcn_update_dpp(x0.y, x1.y, 0x118, 0xf, 0xc, false)),
          make_float2(__builtin_amdgcn_update_dpp(x1.x, x0.x, 0x108, 0xf, 0x3, false),
                      __builtin_amdgcn_update_dpp(x1.y, x0.y, 0x108, 0xf, 0x3, false))};
}
__device__ __forceinline__ tq_p2 tq_xl2(float2 x0, float2 x1) {
  return {make_float2(__builtin_amdgcn_update_dpp(x0.x, x1.x, 0x114, 0xf, 0xa, false),
                      __builtin_amdgcn_update_dpp(x0.y, x1.y, 0x114, 0xf, 0xa, false)),
          make_float2(__builtin_amdgcn_update_dpp(x1.x, x0.x, 0x104, 0xf, 0x5, false),
                      __builtin_amdgcn_update_dpp(x1.y, x0.y, 0x104, 0xf, 0x5, false))};
}
template <int ctrl>
__device__ __forceinline__ tq_p2 tq_xlq(float2 x0, float2 x1, bool hi) {
  const float2 d1 = make_float2(__builtin_amdgcn_update_dpp(0.f, x1.x, ctrl, 0xf, 0xf, false),
                                __builtin_amdgcn_update_dpp(0.f, x1.y, ctrl, 0xf, 0xf, false));
  const float2 d0 = make_float2(__builtin_amdgcn_update_dpp(0.f, x0.x, ctrl, 0xf, 0xf, false),
                                __builtin_amdgcn_update_dpp(0.f, x0.y, ctrl, 0xf, 0xf, false));
  return {hi ? d1 : x0, hi ? x1 : d0};
}
__device__ __forceinline__ tq_p2 tq_xl1(float2 x0, float2 x1, bool hi) { return tq_xlq<0x4E>(x0, x1, hi); }
__device__ __forceinline__ tq_p2 tq_xl0(float2 x0, float2 x1, bool hi) { return tq_xlq<0xB1>(x0, x1, hi); }
"""

NT_HELPERS = r"""
typedef float tq_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldnt(const float2* p) {
  const tq_v2f v = __builtin_nontemporal_load(reinterpret_cast<const tq_v2f*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stnt(float2* p, float2 v) {
  tq_v2f t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<tq_v2f*>(p));
}
"""

ONE = "make_float2(1.f, 0.f)"
ZERO = "make_float2(0.f, 0.f)"


def swz(l: int) -> int:
    """LDS index swizzle of the interpreter (XOR-linear)."""
    return l ^ ((l >> 5) & 31) ^ ((l >> 10) & 31)


PAD_SHIFT = {"pad4": 4, "pad5": 5}


def pad_index(l: int, mode: str) -> int:
    """Padded LDS index: one spare float2 after every 16 ("pad4") or 32
    ("pad5").  Additive over local indices with disjoint bits (pad(t | r) =
    pad(t) + pad(r)), so a lane's part is one register and each slot's part
    is the constant offset of the LDS instruction."""
    return l + (l >> PAD_SHIFT[mode])


def lds_words(m: int, mode: str) -> int:
    """float2 entries of the remap buffer of a 2^m-amplitude group."""
    return pad_index((1 << m) - 1, mode) + 1 if mode in PAD_SHIFT else 1 << m


class StageProgram:
    """Decoded stage words (format: csrc/tqsv_stage.hip)."""

    def __init__(self, words: np.ndarray):
        h = [int(x) for x in np.asarray(words, dtype=np.int64)]
        self.words = h
        self.n_ops, self.m, self.T, self.nnl = h[H_NOPS], h[H_M], h[H_T], h[H_NNL]
        self.n = self.m + self.nnl
        T, m, n = self.T, self.m, self.n
        self.tpos = h[H_POS:H_POS + T]
        self.rpos = h[H_POS + T:H_POS + m]
        self.nlpos = h[H_POS + m:H_POS + n]
        off = h[H_OFF_OPS]
        self.ops = [h[off + 8 * i:off + 8 * i + 8] for i in range(self.n_ops)]
        self.ngp = h[H_NGP]
        self.gslots = h[h[H_OFF_GSLOT]:h[H_OFF_GSLOT] + self.ngp]
        self.roff = self._slot_offsets(self.rpos)
        # load / store layouts (register local bits as a mask; 0 = canonical)
        flags = h[H_FLAGS]
        self.lmask = (flags >> FL_LOAD_SHIFT) & 0x1FFF
        self.smask = (flags >> FL_STORE_SHIFT) & 0x1FFF
        self.lay = {"ld": self._layout(self.lmask), "st": self._layout(self.smask)}

    @staticmethod
    def _slot_offsets(rp):
        return [sum(1 << rp[k] for k in range(R_BITS) if (j >> k) & 1) for j in range(NS)]

    def _layout(self, mask):
        """(thread-bit positions, slot offsets) of the layout whose register
        local bits are `mask`: slot bit k = the k-th register bit ascending,
        thread-index bit i = the i-th other local bit."""
        if mask == 0:
            return list(self.tpos), list(self.roff)
        L = list(self.tpos) + list(self.rpos)
        regs = [b for b in range(self.m) if (mask >> b) & 1]
        thr = [b for b in range(self.m) if not (mask >> b) & 1]
        return [L[b] for b in thr], self._slot_offsets([L[b] for b in regs])

    def lane_var(self, lay: str) -> str:
        """Name of the lane-offset variable of layout lay ("ld"/"st"); the
        store layout shares the load layout's when they are the same."""
        return "toffs" if lay == "st" and self.smask != self.lmask else "toff"

    def remap_layouts(self, w):
        rec = self.words[w[7]:w[7] + 2 * (self.T + R_BITS)]
        k = self.T + R_BITS
        return rec[:k], rec[k:]


def supported(words: np.ndarray) -> bool:
    """Whether a stage program can be compiled (full waves, no register
    controls on dense k-bit ops)."""
    sp = StageProgram(words)
    if sp.T < MIN_T:
        return False
    for w in sp.ops:
        if (w[0] & 15) == OP_MATK and (w[6] & 255):
            return False
    return True


class _Gen:
    def __init__(self, sp: StageProgram, lds_mode: str = "xor", mfma: bool = False, lane_remap: bool = False):
        self.sp = sp
        self.lane_remap = lane_remap   # wave-local remaps as cross-lane moves
        self.mfma = mfma          # 2x2 products of uncontrolled single-wire ops on the matrix cores
        self.used_mfma = False
        self.used_lds = False     # some remap goes through the LDS buffer
        self.lane_flags = set()   # lane bits (0, 1) whose flags lane remaps select on
        self.lines: List[str] = []
        self.k = 0
        self.ind = 2
        self.addr: Dict[tuple, str] = {}     # thread layout -> hoisted LDS index variable
        if lds_mode not in ("xor", "pad4", "pad5"):
            raise ValueError(f"unknown LDS layout {lds_mode!r}")
        self.lds_mode = lds_mode

    def line(self, s: str):
        self.lines.append("  " * self.ind + s)

    def detached(self, fn) -> List[str]:
        """Lines that fn() emits, returned instead of appended to the body
        (variable numbering continues, so names stay unique)."""
        saved, self.lines = self.lines, []
        fn()
        out, self.lines = self.lines, saved
        return [x.strip() for x in out]

    def var(self, expr: str, ty: str = "float2") -> str:
        name = f"v{self.k}"
        self.k += 1
        self.line(f"const {ty} {name} = {expr};")
        return name

    # ---- pieces shared by both directions
    def mat_base(self, w) -> str:
        if w[0] & F_BATCHED:
            return self.var(f"mats + {w[1]} + (long long)row * {w[2]}", "float2*")
        return self.var(f"mats + {w[1]}", "float2*")

    def thread_ok(self, w) -> Optional[str]:
        tm, tv = w[4], w[5]
        if tm == 0:
            return None
        return self.var(f"(tid & {tm}) == {tv}", "bool")

    @staticmethod
    def reg_ctrl(w):
        return w[6] & 255, (w[6] >> 8) & 255

    def select_identity(self, tk: Optional[str], u: List[str], d: int) -> List[str]:
        if tk is None:
            return u
        out = []
        for q, x in enumerate(u):
            r, c = divmod(q, d)
            out.append(self.var(f"csel({tk}, {x}, {ONE if r == c else ZERO})"))
        return out

    def lane_index(self, layout: Sequence[int]) -> str:
        """Hoisted LDS index of this thread's part of a layout."""
        key = tuple(layout[:self.sp.T])
        return self.addr.setdefault(key, f"ra{len(self.addr)}")

    def hoisted(self) -> List[str]:
        out = []
        for key, name in self.addr.items():
            if self.lds_mode in PAD_SHIFT:
                terms = [f"((tid & {1 << i}) ? {pad_index(int(key[i]), self.lds_mode)}u : 0u)"
                         for i in range(len(key))]
                out.append(f"  const unsigned {name} = {' + '.join(terms) if terms else '0u'};")
            else:
                terms = [f"((tid & {1 << i}) ? {swz(int(key[i]))}u : 0u)" for i in range(len(key))]
                out.append(f"  const unsigned {name} = {' ^ '.join(terms) if terms else '0u'};")
        return out

    def _slot(self, ra: str, lr: int) -> str:
        if self.lds_mode in PAD_SHIFT:
            return f"{ra} + {pad_index(lr, self.lds_mode)}u"
        return f"{ra} ^ {swz(lr)}u"

    def remap(self, a: List[str], frm: Sequence[int], to: Sequence[int], lds: str = "lds") -> List[str]:
        plan = lane_plan(self.sp.T, frm, to) if self.lane_remap else None
        if plan is not None:
            return self.remap_lanes(a, plan)
        self.used_lds = True
        T = self.sp.T
        ra, rb = self.lane_index(frm), self.lane_index(to)
        self.line("__syncthreads();")
        for j in range(NS):
            lr = 0
            for k in range(R_BITS):
                if (j >> k) & 1:
                    lr |= frm[T + k]
            self.line(f"{lds}[{self._slot(ra, lr)}] = {a[j]};")
        self.line("__syncthreads();")
        out = []
        for j in range(NS):
            lr = 0
            for k in range(R_BITS):
                if (j >> k) & 1:
                    lr |= to[T + k]
            out.append(self.var(f"{lds}[{self._slot(rb, lr)}]"))
        return out

    def remap_lanes(self, a: List[str], plan) -> List[str]:
        """A remap within each wave as lane-bit / slot-bit exchanges
        (tq_xl<b>), then a renaming of the registers."""
        steps, perm = plan
        cur = list(a)
        for b, sb in steps:
            if b < 2:
                self.lane_flags.add(b)
            for j0 in range(NS):
                if j0 & (1 << sb):
                    continue
                j1 = j0 | (1 << sb)
                extra = f", lb{b}" if b < 2 else ""
                r = self.var(f"tq_xl{b}({cur[j0]}, {cur[j1]}{extra})", "tq_p2")
                cur[j0], cur[j1] = f"{r}.a", f"{r}.b"
        return [cur[perm[j]] for j in range(NS)]

    # ---- forward ops (k_stage_fwd's fwd_op)
    def f_xperm(self, w, a: List[str]):
        K = w[3]
        rm, rv = self.reg_ctrl(w)
        tk = self.thread_ok(w)
        for j in range(NS):
            if j & (1 << K) or (j & rm) != rv:
                continue
            j1 = j | (1 << K)
            if tk is None:
                a[j], a[j1] = a[j1], a[j]
            else:
                x0, x1 = a[j], a[j1]
                a[j] = self.var(f"csel({tk}, {x1}, {x0})")
                a[j1] = self.var(f"csel({tk}, {x0}, {x1})")

    def f_mat1(self, w, a: List[str]):
        K = w[3]
        rm, rv = self.reg_ctrl(w)
        M = self.mat_base(w)
        if w[0] & F_REAL:          # uncontrolled real 2x2 (RY, H ...)
            r = [self.var(f"{M}[{i}].x", "float") for i in range(4)]
            for j in range(NS):
                if j & (1 << K):
                    continue
                j1 = j | (1 << K)
                x0, x1 = a[j], a[j1]
                a[j] = self.var(f"make_float2(fmaf({r[0]}, {x0}.x, {r[1]} * {x1}.x), "
                                f"fmaf({r[0]}, {x0}.y, {r[1]} * {x1}.y))")
                a[j1] = self.var(f"make_float2(fmaf({r[2]}, {x0}.x, {r[3]} * {x1}.x), "
                                 f"fmaf({r[2]}, {x0}.y, {r[3]} * {x1}.y))")
            return
        u = [self.var(f"{M}[{i}]") for i in range(4)]
        tk = self.thread_ok(w)
        u = self.select_identity(tk, u, 2)
        row = self.mm_row(u, adjoint=False) if self.mfma and tk is None else None
        for j in range(NS):
            if j & (1 << K) or (j & rm) != rv:
                continue
            j1 = j | (1 << K)
            x0, x1 = a[j], a[j1]
            if row is not None:
                a[j], a[j1] = self.mm_pair(row, x0, x1)
                continue
            a[j] = self.var(f"cfma({u[1]}, {x1}, cmul({u[0]}, {x0}))")
            a[j1] = self.var(f"cfma({u[3]}, {x1}, cmul({u[2]}, {x0}))")

    def f_sign(self, w, a: List[str]):
        rm, rv = self.reg_ctrl(w)
        tk = self.thread_ok(w)
        for j in range(NS):
            if (j & rm) != rv:
                continue
            x = a[j]
            if tk is None:
                a[j] = self.var(f"make_float2(-{x}.x, -{x}.y)")
            else:
                a[j] = self.var(f"make_float2({tk} ? -{x}.x : {x}.x, {tk} ? -{x}.y : {x}.y)")

    def diag_factors(self, w, M: str, tk: Optional[str]):
        """(fA, fB): the phase for amplitudes whose register-bit parity is even
        / odd, per lane (thread-bit parity folded in; identity where the
        thread-bit controls fail)."""
        tpm = w[3]
        p0, p1 = self.var(f"{M}[0]"), self.var(f"{M}[1]")
        if tpm:
            tp = self.var(f"(__popc(tid & {tpm}) & 1) != 0", "bool")
            fA = self.var(f"{tp} ? {p1} : {p0}")
            fB = self.var(f"{tp} ? {p0} : {p1}")
        else:
            tp, fA, fB = None, p0, p1
        if tk is not None:
            fA = self.var(f"csel({tk}, {fA}, {ONE})")
            fB = self.var(f"csel({tk}, {fB}, {ONE})")
        return tp, fA, fB

    def f_diag(self, w, a: List[str]):
        rm, rv = self.reg_ctrl(w)
        rpm = (w[6] >> 16) & 255
        M = self.mat_base(w)
        _, fA, fB = self.diag_factors(w, M, self.thread_ok(w))
        for j in range(NS):
            if (j & rm) != rv:
                continue
            f = fB if bin(j & rpm).count("1") & 1 else fA
            a[j] = self.var(f"cmul({f}, {a[j]})")

    @staticmethod
    def quads(KH: int, KL: int):
        for j in range(NS):
            if j & ((1 << KH) | (1 << KL)):
                continue
            yield j, [j, j | (1 << KL), j | (1 << KH), j | (1 << KH) | (1 << KL)]

    def f_mat2(self, w, a: List[str]):
        KH, KL = w[3] & 15, (w[3] >> 4) & 15
        rm, rv = self.reg_ctrl(w)
        M = self.mat_base(w)
        u = [self.var(f"{M}[{q}]") for q in range(16)]
        tk = self.thread_ok(w)
        for j, idx in self.quads(KH, KL):
            if (j & rm) != rv:
                continue
            x = [a[i] for i in idx]
            ys = []
            for r in range(4):
                e = f"cmul({u[4 * r]}, {x[0]})"
                for c in range(1, 4):
                    e = f"cfma({u[4 * r + c]}, {x[c]}, {e})"
                y = self.var(e)
                if tk is not None:
                    y = self.var(f"csel({tk}, {y}, {x[r]})")
                ys.append(y)
            for r in range(4):
                a[idx[r]] = ys[r]

    @staticmethod
    def matk_slots(KB: int, EX: int):
        bases = [0] if KB == 4 else [0, 1 << EX]
        for base in bases:
            slots = []
            for r in range(1 << KB):
                s, q = base, 0
                for k in range(R_BITS):
                    if KB == 3 and k == EX:
                        continue
                    if (r >> q) & 1:
                        s |= 1 << k
                    q += 1
                slots.append(s)
            yield slots

    def f_matk(self, w, a: List[str], adjoint: bool = False):
        KB = 4 if w[3] == 4 else 3
        EX = w[3] if KB == 3 else 0
        D = 1 << KB
        M = self.mat_base(w)
        tk = self.thread_ok(w)
        for slots in self.matk_slots(KB, EX):
            x = [a[s] for s in slots]
            ys = []
            for r in range(D):
                if adjoint:     # U^dagger: conj(M[c][r])
                    e = f"cmulc({M}[{r}], {x[0]})"
                    for c in range(1, D):
                        e = f"cfmac({M}[{c * D + r}], {x[c]}, {e})"
                else:
                    e = f"cmul({M}[{r * D}], {x[0]})"
                    for c in range(1, D):
                        e = f"cfma({M}[{r * D + c}], {x[c]}, {e})"
                y = self.var(e)
                if tk is not None:
                    y = self.var(f"csel({tk}, {y}, {x[r]})")
                ys.append(y)
            for r in range(D):
                a[slots[r]] = ys[r]

    def fwd_op(self, w, a: List[str]):
        t = w[0] & 15
        if t == OP_MAT1:
            if w[0] & F_XPERM:
                self.f_xperm(w, a)
            else:
                self.f_mat1(w, a)
        elif t == OP_DIAG:
            if w[0] & F_SIGN:
                self.f_sign(w, a)
            else:
                self.f_diag(w, a)
        elif t == OP_MAT2:
            self.f_mat2(w, a)
        elif t == OP_MATK:
            self.f_matk(w, a)
        else:
            raise ValueError(f"unknown stage op {t}")

    # ---- adjoint ops (k_stage_adj's bwd_op); each returns per-lane partials
    def mm_row(self, u: List[str], adjoint: bool) -> str:
        """This lane's row of the real 4x4 form of U (or U^dagger)."""
        self.used_mfma = True
        fn = "tq_row_udag" if adjoint else "tq_row_u"
        return self.var(f"{fn}({u[0]}, {u[1]}, {u[2]}, {u[3]}, mhi, modd)", "tq_f4")

    def mm_pair(self, row: str, x0: str, x1: str) -> Tuple[str, str]:
        """The lane's pair (x0, x1) times the matrix whose row is `row`."""
        v = self.var(f"tq_mm({row}[0], {row}[1], {row}[2], {row}[3], {x0}, {x1})", "tq_f4")
        return self.var(f"make_float2({v}[0], {v}[1])"), self.var(f"make_float2({v}[2], {v}[3])")

    def acc(self, cur: Optional[str], term: str) -> str:
        return self.var(term if cur is None else f"{cur} + {term}", "float")

    def lane_gate(self, tk: Optional[str], g: List[Optional[str]]) -> List[str]:
        out = []
        for x in g:
            if x is None:
                out.append("0.f")
            elif tk is None:
                out.append(x)
            else:
                out.append(self.var(f"{tk} ? {x} : 0.f", "float"))
        return out

    def b_mat1(self, w, p: List[str], l: List[str], P: int, update: bool = True) -> List[str]:
        K = w[3]
        rm, rv = self.reg_ctrl(w)
        M = self.mat_base(w)
        tk = self.thread_ok(w)
        u = self.select_identity(tk, [self.var(f"{M}[{i}]") for i in range(4)], 2)
        row = self.mm_row(u, adjoint=True) if self.mfma and tk is None else None
        Dm = [[f"{M}[{4 * (q + 1) + i}]" for i in range(4)] for q in range(P)]
        g: List[Optional[str]] = [None] * P
        for j in range(NS):
            if j & (1 << K) or (j & rm) != rv:
                continue
            j1 = j | (1 << K)
            x0, x1 = p[j], p[j1]
            if row is not None:
                y0, y1 = self.mm_pair(row, x0, x1)
            else:
                y0 = self.var(f"cfmac({u[2]}, {x1}, cmulc({u[0]}, {x0}))")
                y1 = self.var(f"cfmac({u[3]}, {x1}, cmulc({u[1]}, {x0}))")
            l0, l1 = l[j], l[j1]
            for q in range(P):
                D = Dm[q]
                v0 = self.var(f"cfma({D[1]}, {y1}, cmul({D[0]}, {y0}))")
                v1 = self.var(f"cfma({D[3]}, {y1}, cmul({D[2]}, {y0}))")
                g[q] = self.acc(g[q], f"(redot({l0}, {v0}) + redot({l1}, {v1}))")
            if update:
                if row is not None:
                    n0, n1 = self.mm_pair(row, l0, l1)
                else:
                    n0 = self.var(f"cfmac({u[2]}, {l1}, cmulc({u[0]}, {l0}))")
                    n1 = self.var(f"cfmac({u[3]}, {l1}, cmulc({u[1]}, {l0}))")
                p[j], p[j1], l[j], l[j1] = y0, y1, n0, n1
        return self.lane_gate(tk, g)

    def b_mat1_mg(self, w, p: List[str], l: List[str], update: bool = True) -> List[str]:
        """Hermitian-generator form (see k_mg_to_grad): the four sums
        S = Im<lam|s|x>, s = I, X, Y, Z, over x = psi before the un-apply,
        12 FMA per pair, then (update) psi <- U^dag psi, lam <- U^dag lam."""
        K = w[3]
        nc = bool(w[0] & F_NOCTRL)
        rm, rv = (0, 0) if nc else self.reg_ctrl(w)
        M = self.mat_base(w)
        tk = None if nc else self.thread_ok(w)
        u = self.select_identity(tk, [self.var(f"{M}[{i}]") for i in range(4)], 2)
        row = self.mm_row(u, adjoint=True) if self.mfma and tk is None and update else None
        a = b = c = d = None
        for j in range(NS):
            if j & (1 << K) or (j & rm) != rv:
                continue
            j1 = j | (1 << K)
            x0, x1 = p[j], p[j1]
            l0, l1 = l[j], l[j1]
            # a += Im conj(l0) x0, b += Im conj(l1) x1, c += conj(l0) x1, d += conj(l1) x0
            a = self.var(f"fmaf({l0}.x, {x0}.y, fmaf(-{l0}.y, {x0}.x, {a if a else '0.f'}))", "float")
            b = self.var(f"fmaf({l1}.x, {x1}.y, fmaf(-{l1}.y, {x1}.x, {b if b else '0.f'}))", "float")
            c = self.var(f"cfmac({l0}, {x1}, {c if c else ZERO})")
            d = self.var(f"cfmac({l1}, {x0}, {d if d else ZERO})")
            if not update:
                continue
            if row is not None:
                y0, y1 = self.mm_pair(row, x0, x1)
                n0, n1 = self.mm_pair(row, l0, l1)
            else:
                y0 = self.var(f"cfmac({u[2]}, {x1}, cmulc({u[0]}, {x0}))")
                y1 = self.var(f"cfmac({u[3]}, {x1}, cmulc({u[1]}, {x0}))")
                n0 = self.var(f"cfmac({u[2]}, {l1}, cmulc({u[0]}, {l0}))")
                n1 = self.var(f"cfmac({u[3]}, {l1}, cmulc({u[1]}, {l0}))")
            p[j], p[j1], l[j], l[j1] = y0, y1, n0, n1
        if a is None:
            return self.lane_gate(tk, [None] * 4)
        g = [f"{a} + {b}", f"{c}.y + {d}.y", f"{d}.x - {c}.x", f"{a} - {b}"]
        return self.lane_gate(tk, g)

    def b_mat2(self, w, p: List[str], l: List[str], P: int, update: bool = True) -> List[str]:
        KH, KL = w[3] & 15, (w[3] >> 4) & 15
        rm, rv = self.reg_ctrl(w)
        M = self.mat_base(w)
        tk = self.thread_ok(w)
        u = self.select_identity(tk, [self.var(f"{M}[{q}]") for q in range(16)], 4)
        g: List[Optional[str]] = [None] * P
        for j, idx in self.quads(KH, KL):
            if (j & rm) != rv:
                continue
            x = [p[i] for i in idx]
            lm = [l[i] for i in idx]
            y = []
            for r in range(4):
                e = f"cmulc({u[r]}, {x[0]})"
                for c in range(1, 4):
                    e = f"cfmac({u[4 * c + r]}, {x[c]}, {e})"
                y.append(self.var(e))
            for q in range(P):
                base = 16 * (q + 1)
                t = None
                for r in range(4):
                    e = f"cmul({M}[{base + 4 * r}], {y[0]})"
                    for c in range(1, 4):
                        e = f"cfma({M}[{base + 4 * r + c}], {y[c]}, {e})"
                    v = self.var(e)
                    t = self.acc(t, f"redot({lm[r]}, {v})")
                g[q] = self.acc(g[q], t)
            if not update:
                continue
            for r in range(4):
                e = f"cmulc({u[r]}, {lm[0]})"
                for c in range(1, 4):
                    e = f"cfmac({u[4 * c + r]}, {lm[c]}, {e})"
                l[idx[r]] = self.var(e)
                p[idx[r]] = y[r]
        return self.lane_gate(tk, g)

    def b_diag(self, w, p: List[str], l: List[str], P: int, update: bool = True) -> List[str]:
        rm, rv = self.reg_ctrl(w)
        rpm = (w[6] >> 16) & 255
        M = self.mat_base(w)
        tk = self.thread_ok(w)
        tp, fA, fB = self.diag_factors(w, M, tk)
        dA, dB = [], []
        for q in range(P):
            e0, e1 = f"{M}[{2 * (q + 1)}]", f"{M}[{2 * (q + 1) + 1}]"
            if tp is None:
                dA.append(self.var(e0))
                dB.append(self.var(e1))
            else:
                dA.append(self.var(f"{tp} ? {e1} : {e0}"))
                dB.append(self.var(f"{tp} ? {e0} : {e1}"))
        g: List[Optional[str]] = [None] * P
        for j in range(NS):
            if (j & rm) != rv:
                continue
            odd = bin(j & rpm).count("1") & 1
            f = fB if odd else fA
            y = self.var(f"cmulc({f}, {p[j]})")
            lj = l[j]
            for q in range(P):
                d = dB[q] if odd else dA[q]
                g[q] = self.acc(g[q], f"redot({lj}, cmul({d}, {y}))")
            if update:
                p[j] = y
                l[j] = self.var(f"cmulc({f}, {lj})")
        return self.lane_gate(tk, g)

    def bwd_op(self, w, p: List[str], l: List[str], update: bool = True) -> Optional[List[str]]:
        """One op of the reverse sweep: its gradient partials (None for ops
        without parameters) and, with update, psi and lam un-applied."""
        t = w[0] & 15
        fl = w[0]
        P = ((fl >> 4) & 15) if w[7] >= 0 else 0
        if not update and w[7] < 0:
            return None
        if t == OP_MAT1:
            if fl & F_XPERM:
                self.f_xperm(w, p)
                self.f_xperm(w, l)
                return None
            if fl & F_MG:
                return self.b_mat1_mg(w, p, l, update)
            return self.b_mat1(w, p, l, P, update)
        if t == OP_DIAG:
            if fl & F_SIGN:
                self.f_sign(w, p)
                self.f_sign(w, l)
                return None
            return self.b_diag(w, p, l, P, update)
        if t == OP_MAT2:
            return self.b_mat2(w, p, l, P, update)
        if t == OP_MATK:
            self.f_matk(w, p, adjoint=True)
            self.f_matk(w, l, adjoint=True)
            return None
        raise ValueError(f"unknown stage op {t}")

    def reduce(self, w, g: List[str]):
        """Wave totals of a gradient op's per-lane partials, added into the
        workgroup's LDS accumulators.  Lane l ends with the total of value
        l >> 3 (8 values), l >> 4 (4), l >> 5 (2) or of the only value --
        the assignment of the interpreter's halving exchange -- built from
        permlane swaps and DPP, with no LDS traffic and no selects."""
        Pp = 4 if (w[0] & F_MG) else ((w[0] >> 4) & 15)
        if Pp > 4:
            nv, sh, cond = 8, 3, "(lane & 7) == 0"
        elif Pp > 2:
            nv, sh, cond = 4, 4, f"(lane & 15) == 0 && (lane >> 4) < {Pp}"
        elif Pp == 2:
            nv, sh, cond = 2, 5, "(lane & 31) == 0"
        else:
            nv, sh, cond = 1, 6, "lane == 0"
        a = (list(g) + ["0.f"] * nv)[:nv]
        f = lambda e: self.var(e, "float")
        if nv == 8:
            w4 = [f(f"sum32({a[i]}, {a[i + 4]})") for i in range(4)]      # lanes < 32: value i, else i + 4
            u0 = f(f"sum16({w4[0]}, {w4[2]})")                          # rows: values 0, 2, 4, 6
            u1 = f(f"sum16({w4[1]}, {w4[3]})")                          # rows: values 1, 3, 5, 7
            s0 = f(f"{u0} + dpp_ror8({u0})")
            s1 = f(f"{u1} + dpp_ror8({u1})")
            z = f(f"(lane & 8) ? {s1} : {s0}")                       # half h of row r: value 2 r + h
            c = f(f"allsum8({z})")
        elif nv == 4:
            w0 = f(f"sum32({a[0]}, {a[2]})")
            w1 = f(f"sum32({a[1]}, {a[3]})")
            u = f(f"sum16({w0}, {w1})")                                  # row r: value r
            c = f(f"allsum8({u} + dpp_ror8({u}))")
        elif nv == 2:
            w0 = f(f"sum32({a[0]}, {a[1]})")                             # lanes < 32: value 0
            u = f(f"sum16({w0}, {w0})")
            c = f(f"allsum8({u} + dpp_ror8({u}))")
        else:
            w0 = f(f"sum32({a[0]}, {a[0]})")
            u = f(f"sum16({w0}, {w0})")
            c = f(f"allsum8({u} + dpp_ror8({u}))")
        self.line(f"if ({cond}) gacc[wrow + {w[7]} + (lane >> {sh})] += {c};")


def _split_addr(sp: StageProgram) -> bool:
    """Amplitude addresses as (wave-uniform 64-bit base + constant) plus one
    32-bit per-lane byte offset (toff * 8, one per layout), so that every
    load (store) of a group shares one offset register; needs toff * 8 < 2^32."""
    return max(sp.lay["ld"][0] + sp.lay["st"][0], default=0) <= 28


def _buf_ok(sp: StageProgram) -> bool:
    """Buffer instructions address base + 32-bit lane offset + 32-bit scalar
    offset: every in-group byte offset 8 * (toff + roff) must stay below 2^31,
    i.e. no local bit above 27."""
    return max(list(sp.tpos) + list(sp.rpos), default=0) <= 27


def _mem_mode(sp: StageProgram, nontemporal: bool, buf: bool) -> str:
    """How the state is addressed: "buf" (buffer loads/stores: one scalar
    resource per array and group, the lane offset toff8 and a scalar constant
    per slot, no per-access vector address arithmetic), "nt" (pointers with
    non-temporal hints) or "ptr" (pointers)."""
    if nontemporal:
        return "nt"
    return "buf" if buf and _buf_ok(sp) else "ptr"


def _group_vars(sp: StageProgram, gv: str, sfx: str, mode: str, arrays: Sequence[str]) -> List[str]:
    """Row, group offset and base index of group `gv` (names suffixed by
    `sfx`), and what addressing `mode` needs for each array."""
    terms = [f"((unsigned)(({gv} >> {i}) & 1ll) << {sp.nlpos[i]})" for i in range(sp.nnl)]
    out = [
        f"const unsigned long long row{sfx} = (unsigned long long)({gv}) >> {sp.nnl};",
        f"const unsigned goff{sfx} = {' | '.join(terms) if terms else '0u'};",
        f"const unsigned long long gbase{sfx} = (row{sfx} << {sp.n}) + goff{sfx};",
    ]
    if mode == "buf":
        out += [f"const __amdgpu_buffer_rsrc_t r_{a}{sfx} = tq_rsrc({a} + gbase{sfx});" for a in arrays]
    return out


def _ptr(sp: StageProgram, arr: str, j: int, const: bool, sfx: str, lay: str) -> str:
    """Address of amplitude slot j of this lane's group in array arr, in
    layout lay ("ld": the stage's load layout, "st": its store layout)."""
    c = "const " if const else ""
    tv, ro = sp.lane_var(lay), sp.lay[lay][1][j]
    if _split_addr(sp):
        return f"({c}float2*)(({c}char*)({arr} + gbase{sfx}) + {8 * ro}ull + {tv}8)"
    return f"({arr} + gbase{sfx} + {tv} + {ro})"


def _load(sp: StageProgram, arr: str, j: int, mode: str, sfx: str = "", lay: str = "ld") -> str:
    if mode == "buf":
        return f"tq_ld(r_{arr}{sfx}, {sp.lane_var(lay)}8, {8 * sp.lay[lay][1][j]})"
    p = _ptr(sp, arr, j, True, sfx, lay)
    return f"ldnt({p})" if mode == "nt" else f"*{p}"


def _store(sp: StageProgram, arr: str, j: int, val: str, mode: str, sfx: str = "", lay: str = "st") -> str:
    if mode == "buf":
        return f"tq_st(r_{arr}{sfx}, {sp.lane_var(lay)}8, {8 * sp.lay[lay][1][j]}, {val});"
    p = _ptr(sp, arr, j, False, sfx, lay)
    return f"stnt({p}, {val});" if mode == "nt" else f"*{p} = {val};"


def _toff(tpos: Sequence[int]) -> str:
    terms = [f"((((unsigned)tid >> {i}) & 1u) << {p})" for i, p in enumerate(tpos)]
    return " | ".join(terms) if terms else "0u"


def _lane_offsets(sp: StageProgram) -> List[str]:
    """Per-lane element and byte offsets of the load layout (toff, toff8)
    and, when it differs, of the store layout (toffs, toffs8)."""
    out = [f"  const unsigned toff = {_toff(sp.lay['ld'][0])};",
           "  const unsigned toff8 = toff << 3;"]
    if sp.lane_var("st") != "toff":
        out += [f"  const unsigned toffs = {_toff(sp.lay['st'][0])};",
                "  const unsigned toffs8 = toffs << 3;"]
    return out


def _mfma(kind: str, mfma: Optional[bool] = None) -> bool:
    kind = "a" if kind == "p" else kind
    return CODEGEN["mfma_" + kind] if mfma is None else bool(mfma)


def _lane(kind: str, lane_remap: Optional[bool] = None) -> bool:
    kind = "a" if kind == "p" else kind
    return CODEGEN["lane_remap_" + kind] if lane_remap is None else bool(lane_remap)


def _lane_bits(g: "_Gen") -> List[str]:
    """Lane-position flags the matrix-core rows (tq_row_u / tq_row_udag) and
    the lane remaps on lane bits 0 and 1 need."""
    out = (["  const bool mhi = (threadIdx.x & 2u) != 0u;", "  const bool modd = (threadIdx.x & 1u) != 0u;"]
           if g.used_mfma else [])
    return out + [f"  const bool lb{b} = (threadIdx.x & {1 << b}u) != 0u;" for b in sorted(g.lane_flags)]


def lane_plan(T: int, frm: Sequence[int], to: Sequence[int]):
    """How a remap from layout `frm` to layout `to` (local-bit values of the
    T thread bits, then the R_BITS slot bits) runs inside each wave: None
    unless every wave bit (thread bit >= 6) keeps its local bit and every lane
    bit that changes takes a local bit from a register slot and gives its own
    to one.  Otherwise ([(lane bit b, slot bit s), ...], perm): exchange lane
    bit b with slot bit s in turn (slot bit s indexing the slots of the layout
    at that point), then slot j of `to` is slot perm[j] of the result."""
    F, G = [int(x) for x in frm], [int(x) for x in to]
    if T < 6 or any(F[i] != G[i] for i in range(6, T)):
        return None
    reg, regs_to = F[T:T + R_BITS], set(G[T:T + R_BITS])
    steps = []
    for b in range(6):
        if F[b] == G[b]:
            continue
        if G[b] not in reg or F[b] not in regs_to:
            return None
        sb = reg.index(G[b])
        steps.append((b, sb))
        reg[sb] = F[b]
    if not steps:
        return None
    where = [reg.index(G[T + k]) for k in range(R_BITS)]
    perm = [sum(1 << where[k] for k in range(R_BITS) if (j >> k) & 1) for j in range(NS)]
    return steps, perm


def _codegen(buf: Optional[bool], prefetch: Optional[bool], kind: str, lds: Optional[str] = None,
             racc: Optional[int] = None):
    kind = "a" if kind == "p" else kind
    return (CODEGEN["buf"] if buf is None else buf,
            CODEGEN["prefetch_" + kind] if prefetch is None else prefetch,
            CODEGEN["lds"] if lds is None else lds,
            (CODEGEN["racc"] if racc is None else racc) if kind == "a" else 0)


def gen_forward(words: np.ndarray, name: str, nontemporal: bool = False, buf: Optional[bool] = None,
                prefetch: Optional[bool] = None, lds: Optional[str] = None, mfma: Optional[bool] = None,
                lane_remap: Optional[bool] = None) -> str:
    """HIP source of one forward stage kernel (k_stage_fwd semantics).
    nontemporal: the state is read and written with non-temporal hints;
    buf: buffer instructions for the state where the offsets allow;
    prefetch: each group's loads are issued one group ahead (software
    pipelining), so that they are in flight while the previous group
    computes; lds: the remap buffer's index map, "xor" (swizzled) or "pad4" /
    "pad5" (padded, additive).  None takes the module default (CODEGEN)."""
    sp = StageProgram(words)
    if sp.T < MIN_T:
        raise ValueError("compiled stages need T >= 6")
    buf, prefetch, lds, _ = _codegen(buf, prefetch, "f", lds)
    mode = _mem_mode(sp, nontemporal, buf)
    g = _Gen(sp, lds, _mfma("f", mfma), _lane("f", lane_remap))
    a = [f"i{j}" for j in range(NS)]
    for w in sp.ops:
        if (w[0] & 15) == OP_REMAP:
            frm, to = sp.remap_layouts(w)
            a = g.remap(a, frm, to)
        else:
            g.fwd_op(w, a)
    body = ["  " + s for s in g.lines]
    stores = [f"      {_store(sp, 'out', j, a[j], mode, lay='st')}" for j in range(NS)]
    nt = 1 << sp.T
    src = [f'extern "C" __global__ void __launch_bounds__({nt}) {name}(float2* __restrict__ out, '
           "const float2* __restrict__ in, const float2* __restrict__ mats, int init_zero, long long n_groups, "
           "long long gpw) {"]
    if g.used_lds:
        src.append(f"  __shared__ float2 lds[{lds_words(sp.m, lds)}];")
    src.append("  const int tid = threadIdx.x;")
    src += _lane_offsets(sp)
    src += _lane_bits(g)
    src += g.hoisted()
    src += ["  const long long g0 = (long long)blockIdx.x * gpw;",
            "  long long g1 = g0 + gpw;",
            "  if (g1 > n_groups) g1 = n_groups;"]
    # the first stage of a pass starts from |0..0> and reads nothing
    src += ["  if (init_zero) {",
            "    for (long long g = g0; g < g1; ++g) {"]
    src += ["      " + s for s in _group_vars(sp, "g", "", mode, ["out"])]
    src += [f"      const float2 i{j} = make_float2((goff + toff + {sp.lay['ld'][1][j]}u) == 0u ? 1.f : 0.f, 0.f);"
            for j in range(NS)]
    src += body + stores + ["    }", "    return;", "  }"]
    if prefetch:
        src += [f"  float2 {', '.join(f'c{j}' for j in range(NS))};",
                "  if (g0 < g1) {"]
        src += ["    " + s for s in _group_vars(sp, "g0", "n", mode, ["in"])]
        src += [f"    c{j} = {_load(sp, 'in', j, mode, 'n')};" for j in range(NS)]
        src += ["  }",
                "  for (long long g = g0; g < g1; ++g) {"]
        src += ["    " + s for s in _group_vars(sp, "g", "", mode, ["out"])]
        src += ["    const long long gn = g + 1 < g1 ? g + 1 : g;"]
        src += ["    " + s for s in _group_vars(sp, "gn", "n", mode, ["in"])]
        src += [f"    const float2 q{j} = {_load(sp, 'in', j, mode, 'n')};" for j in range(NS)]
        src += [f"    const float2 i{j} = c{j};" for j in range(NS)]
        src += body + stores
        src += [f"    c{j} = q{j};" for j in range(NS)]
        src += ["  }"]
    else:
        src += ["  for (long long g = g0; g < g1; ++g) {"]
        src += ["    " + s for s in _group_vars(sp, "g", "", mode, ["in", "out"])]
        src += [f"    const float2 i{j} = {_load(sp, 'in', j, mode)};" for j in range(NS)]
        src += body + stores + ["  }"]
    src.append("}")
    return "\n".join(src) + "\n"


def gen_adjoint(words: np.ndarray, name: str, nontemporal: bool = False, buf: Optional[bool] = None,
                prefetch: Optional[bool] = None, lds: Optional[str] = None, racc: Optional[int] = None,
                skip: Sequence[int] = (), no_store: bool = False, mfma: Optional[bool] = None,
                lane_remap: Optional[bool] = None) -> str:
    """HIP source of one adjoint stage kernel (k_stage_adj semantics);
    nontemporal, buf, prefetch and lds as for gen_forward (prefetch loads
    psi and lam of the next group).  racc: when the stage has at most this
    many gradient values, each lane keeps its partial sums in registers
    across the groups of a row and the cross-lane reduction runs once per
    row instead of once per group.

    Pruned sweeps (no gradient with respect to the input state): the ops at
    the program positions in `skip` are not un-applied (a trainable one still
    adds its gradient partials), and with no_store nothing is written back --
    the sweep ends with this stage, so the code after its last trainable op
    (in sweep order) is dropped too."""
    sp = StageProgram(words)
    if sp.T < MIN_T:
        raise ValueError("compiled stages need T >= 6")
    buf, prefetch, lds, racc = _codegen(buf, prefetch, "a", lds, racc)
    mode = _mem_mode(sp, nontemporal, buf)
    g = _Gen(sp, lds, _mfma("a", mfma), _lane("a", lane_remap))
    p = [f"p{j}" for j in range(NS)]
    l = [f"l{j}" for j in range(NS)]
    in_regs = 0 < sp.ngp <= racc
    acc_ops = []            # (op words, register accumulators of its partials)
    skip = frozenset(int(x) for x in skip)
    grad_pos = [i for i, w in enumerate(sp.ops) if (w[0] & 15) != OP_REMAP and w[7] >= 0]
    first_grad = min(grad_pos) if grad_pos else len(sp.ops)
    for pos in reversed(range(len(sp.ops))):
        w = sp.ops[pos]
        if no_store and pos < first_grad:
            break
        if (w[0] & 15) == OP_REMAP:
            frm, to = sp.remap_layouts(w)
            p = g.remap(p, to, frm)
            l = g.remap(l, to, frm)
            continue
        parts = g.bwd_op(w, p, l, update=pos not in skip)
        if w[7] >= 0 and parts is not None:
            if in_regs:
                names = [f"ga{len(acc_ops)}_{q}" for q in range(len(parts))]
                for nm, e in zip(names, parts):
                    if e != "0.f":
                        g.line(f"{nm} += {e};")
                acc_ops.append((w, names))
            else:
                g.reduce(w, parts)
    # register accumulators -> the workgroup's LDS accumulators (full waves)
    drain: List[str] = []
    for w, names in acc_ops:
        drain += g.detached(lambda w=w, names=names: g.reduce(w, names))
        drain += [f"{nm} = 0.f;" for nm in names]
    nt = 1 << sp.T
    ngp = sp.ngp
    nwave = max(1, nt >> 6)
    src = []
    # psi / lam: where the swept states are written; psi_in / lam_in: where
    # they are read (the same buffers in place, or the forward output and
    # the incoming gradient for the first stage of the sweep).  Gradient
    # partials: each wave adds its per-op totals into its own row of gacc
    # (no atomics); at every batch-row change the workgroup writes the sum
    # over waves, in wave order, to part[(blockIdx.x * rpw + row - row0) *
    # ngp + t], which a fixed-order reduction adds into the gradient.
    src.append(f'extern "C" __global__ void __launch_bounds__({nt}) {name}(float2* __restrict__ psi, '
               "float2* __restrict__ lam, const float2* psi_in, const float2* lam_in, "
               "const float2* __restrict__ mats, float* __restrict__ part, "
               "long long rpw, long long n_groups, long long gpw) {")
    if g.used_lds:
        src.append(f"  __shared__ float2 lds[{lds_words(sp.m, lds)}];")
    if ngp:
        src.append(f"  __shared__ float gacc[{nwave * ngp}];")
        src.append(f"  for (int t = threadIdx.x; t < {nwave * ngp}; t += {nt}) gacc[t] = 0.f;")
        src.append(f"  const int wrow = (threadIdx.x >> 6) * {ngp};")
    for _, names in acc_ops:
        src.append(f"  float {', '.join(nm + ' = 0.f' for nm in names)};")
    src.append("  const int tid = threadIdx.x;")
    src.append("  const int lane = tid & 63;")
    src += _lane_offsets(sp)
    src += _lane_bits(g)
    src += g.hoisted()
    src += ["  const long long g0 = (long long)blockIdx.x * gpw;",
            "  long long g1 = g0 + gpw;",
            "  if (g1 > n_groups) g1 = n_groups;",
            "  long long cur_row = -1;",
            f"  float* __restrict__ wpart = part + (long long)blockIdx.x * rpw * {max(ngp, 1)};",
            f"  const long long row0 = g0 >> {sp.nnl};",
            "  __syncthreads();"]
    if prefetch:
        src += [f"  float2 {', '.join(f'pc{j}' for j in range(NS))};",
                f"  float2 {', '.join(f'lc{j}' for j in range(NS))};",
                "  if (g0 < g1) {"]
        src += ["    " + s for s in _group_vars(sp, "g0", "n", mode, ["psi_in", "lam_in"])]
        src += [f"    pc{j} = {_load(sp, 'psi_in', j, mode, 'n', 'st')};" for j in range(NS)]
        src += [f"    lc{j} = {_load(sp, 'lam_in', j, mode, 'n', 'st')};" for j in range(NS)]
        src += ["  }"]
    src += ["  for (long long g = g0; g < g1; ++g) {"]
    src += ["    " + s for s in _group_vars(sp, "g", "", mode, ["psi", "lam"] if prefetch else
                                            ["psi", "lam", "psi_in", "lam_in"])]
    if ngp:
        src += ["    if ((long long)row != cur_row) {",
                "      if (cur_row >= 0) {"]
        src += ["        " + x for x in drain]
        src += ["        __syncthreads();",
                f"        for (int t = tid; t < {ngp}; t += {nt}) {{",
                "          float v = 0.f;",
                f"          for (int w = 0; w < {nwave}; ++w) {{ v += gacc[w * {ngp} + t]; gacc[w * {ngp} + t] = 0.f; }}",
                f"          wpart[(cur_row - row0) * {ngp} + t] = v;",
                "        }",
                "        __syncthreads();",
                "      }",
                "      cur_row = (long long)row;",
                "    }"]
    if prefetch:
        src += ["    const long long gn = g + 1 < g1 ? g + 1 : g;"]
        src += ["    " + s for s in _group_vars(sp, "gn", "n", mode, ["psi_in", "lam_in"])]
        src += [f"    const float2 pq{j} = {_load(sp, 'psi_in', j, mode, 'n', 'st')};" for j in range(NS)]
        src += [f"    const float2 lq{j} = {_load(sp, 'lam_in', j, mode, 'n', 'st')};" for j in range(NS)]
        src += [f"    const float2 p{j} = pc{j};" for j in range(NS)]
        src += [f"    const float2 l{j} = lc{j};" for j in range(NS)]
    else:
        src += [f"    const float2 p{j} = {_load(sp, 'psi_in', j, mode, '', 'st')};" for j in range(NS)]
        src += [f"    const float2 l{j} = {_load(sp, 'lam_in', j, mode, '', 'st')};" for j in range(NS)]
    src += g.lines
    if not no_store:
        src += [f"    {_store(sp, 'psi', j, p[j], mode, lay='ld')}" for j in range(NS)]
        src += [f"    {_store(sp, 'lam', j, l[j], mode, lay='ld')}" for j in range(NS)]
    if prefetch:
        src += [f"    pc{j} = pq{j};" for j in range(NS)]
        src += [f"    lc{j} = lq{j};" for j in range(NS)]
    src.append("  }")
    if ngp:
        src += ["  if (cur_row >= 0) {"]
        src += ["    " + x for x in drain]
        src += ["    __syncthreads();",
                f"    for (int t = tid; t < {ngp}; t += {nt}) {{",
                "      float v = 0.f;",
                f"      for (int w = 0; w < {nwave}; ++w) v += gacc[w * {ngp} + t];",
                f"      wpart[(cur_row - row0) * {ngp} + t] = v;",
                "    }",
                "  }"]
    src.append("}")
    return "\n".join(src) + "\n"


# ----------------------------------------------------------------------------
# compile + cache
# ----------------------------------------------------------------------------
def _spec(prog):
    """(words, extra gen_adjoint arguments) of a program entry: stage words,
    or for kind "p" a tuple (words, skip positions, no_store)."""
    if isinstance(prog, tuple):
        words, skip, no_store = prog
        return words, {"skip": tuple(sorted(int(x) for x in skip)), "no_store": bool(no_store)}
    return prog, {}


def kernel_name(kind: str, prog, nt: bool = False) -> str:
    words, extra = _spec(prog)
    hs = hashlib.sha1(np.ascontiguousarray(words, dtype=np.int32).tobytes())
    if extra:
        hs.update(repr(sorted(extra.items())).encode())
    hs.update(COMPILE_OPTS[kind].encode())
    hs.update(repr((_codegen(None, None, kind), _mfma(kind), _lane(kind))).encode())
    if nt:
        hs.update(b"nt")
    h = hs.hexdigest()[:20]
    return f"tqsv_j{kind}_{h}"


def _cache_dir() -> Optional[Path]:
    """Directory of the on-disk code-object cache (TQSV_JIT_CACHE; "0" or
    "off" disables it; default ~/.cache/torchquantum_amd/stage-kernels)."""
    d = os.environ.get("TQSV_JIT_CACHE")
    if d is not None and d.strip().lower() in ("", "0", "off", "none"):
        return None
    base = Path(d) if d else Path.home() / ".cache" / "torchquantum_amd" / "stage-kernels"
    try:
        base.mkdir(parents=True, exist_ok=True)
    except OSError:
        return None
    return base


def _lib_tag() -> str:
    """Identifies the library build (its compile options travel with it)."""
    from . import _native as N
    try:
        st = N.LIB_PATH.stat()
        return f"{st.st_size}:{st.st_mtime_ns}"
    except OSError:
        return "?"


def compile_source(src: str, opts: str = "") -> bytes:
    """hipRTC-compile a source (no GPU needed) with extra compiler options
    `opts`; returns the code object.  Code objects are cached on disk by a
    hash of the source, the options and the library build."""
    from . import _native as N
    cache = _cache_dir()
    path = None
    if cache is not None:
        h = hashlib.sha1((_lib_tag() + "\n" + opts + "\n" + src).encode()).hexdigest()
        path = cache / f"{h}.co"
        try:
            data = path.read_bytes()
            if data:
                return data
        except OSError:
            pass
    lib = N.lib()
    size = ctypes.c_int64(0)
    log = ctypes.create_string_buffer(1 << 16)
    srcb = src.encode()
    optb = opts.encode() if opts else None
    rc = lib.tqsv_jit_compile(srcb, optb, None, 0, ctypes.byref(size), log, len(log))
    if rc != 0:
        raise RuntimeError(f"stage kernel compile failed: {log.value.decode(errors='replace')[:4000]}")
    buf = ctypes.create_string_buffer(size.value)
    rc = lib.tqsv_jit_compile(srcb, optb, buf, size.value, ctypes.byref(size), None, 0)
    if rc != 0:
        raise RuntimeError(f"stage kernel compile failed: {lib.tqsv_last_error().decode(errors='replace')}")
    code = buf.raw[:size.value]
    if path is not None:
        try:
            tmp = path.with_suffix(f".{os.getpid()}.{threading.get_ident()}.tmp")
            tmp.write_bytes(code)
            os.replace(tmp, path)
        except OSError:
            pass
    return code


_lock = threading.Lock()
_shutdown = threading.Event()
_funcs: Dict[tuple, int] = {}          # (device index, kernel name) -> hipFunction_t
_modules: List[int] = []                # loaded modules (kept for the process lifetime)


def get_kernels(programs: Sequence[np.ndarray], kind: str, device_index: int, nt: bool = False) -> List[int]:
    """hipFunction_t handles (as ints) of the compiled stage kernels for the
    given programs; kind "f" = forward, "a" = adjoint, "p" = pruned adjoint
    (programs given as (words, skip positions, no_store)).  Compiles and
    loads (on the calling thread) what is not loaded yet.  hipRTC serialises
    compiles within a process, so sources are compiled one after another."""
    from . import _native as N

    def gen(prog, nm, nt):
        words, extra = _spec(prog)
        if kind == "f":
            return gen_forward(words, nm, nt)
        return gen_adjoint(words, nm, nt, **extra)

    names = [kernel_name(kind, w, nt) for w in programs]
    with _lock:
        missing = {}
        for nm, w in zip(names, programs):
            if (device_index, nm) not in _funcs and nm not in missing:
                missing[nm] = w
    if missing:
        items = list(missing.items())
        chunks = [items[i:i + KERNELS_PER_SOURCE] for i in range(0, len(items), KERNELS_PER_SOURCE)]
        codes = []
        verbose = os.environ.get("TQSV_JIT_LOG") == "1"
        for ci, ch in enumerate(chunks):
            if _shutdown.is_set():
                raise RuntimeError("stage kernel compile cancelled at interpreter exit")
            codes.append(compile_source(PRELUDE + (NT_HELPERS if nt else "") + "\n".join(gen(w, nm, nt) for nm, w in ch),
                                        COMPILE_OPTS[kind]))
            if verbose:
                import sys
                print(f"[stagejit] {kind}: {min((ci + 1) * KERNELS_PER_SOURCE, len(items))}/{len(items)} stage "
                      "kernels compiled", file=sys.stderr, flush=True)
        lib = N.lib()
        with _lock:
            for ch, code in zip(chunks, codes):
                if all((device_index, nm) in _funcs for nm, _ in ch):
                    continue
                cnames = (ctypes.c_char_p * len(ch))(*[nm.encode() for nm, _ in ch])
                fptrs = (ctypes.c_void_p * len(ch))()
                mod = ctypes.c_void_p()
                buf = ctypes.create_string_buffer(code, len(code))
                rc = lib.tqsv_jit_load(buf, len(code), len(ch), cnames, ctypes.byref(mod), fptrs)
                if rc != 0:
                    raise RuntimeError(f"stage kernel load failed: {lib.tqsv_last_error().decode(errors='replace')}")
                _modules.append(mod.value)
                for (nm, _), f in zip(ch, fptrs):
                    _funcs[(device_index, nm)] = f
    with _lock:
        return [_funcs[(device_index, nm)] for nm in names]


# ----------------------------------------------------------------------------
# background compiles: the interpreting kernels run until these are ready
# ----------------------------------------------------------------------------
class CompileJob:
    """One plan's stage kernels being compiled on the background thread."""

    def __init__(self, programs, kind, device_index, nt=False):
        self.programs, self.kind, self.device_index, self.nt = list(programs), kind, device_index, nt
        self.done = threading.Event()
        self.result: Optional[List[int]] = None
        self.error: Optional[BaseException] = None


_jobs: "queue.Queue[CompileJob]" = queue.Queue()
_pending: List[CompileJob] = []
_worker: Optional[threading.Thread] = None


def _work():
    import torch
    while True:
        job = _jobs.get()
        try:
            torch.cuda.set_device(job.device_index)
            job.result = get_kernels(job.programs, job.kind, job.device_index, job.nt)
        except BaseException as e:   # reported to the plan that asked
            job.error = e
        finally:
            job.done.set()
            with _lock:
                if job in _pending:
                    _pending.remove(job)


def submit(programs: Sequence[np.ndarray], kind: str, device_index: int, nt: bool = False) -> CompileJob:
    """Queue a compile on the background thread; returns at once."""
    global _worker
    job = CompileJob(programs, kind, device_index, nt)
    with _lock:
        _pending.append(job)
        if _worker is None or not _worker.is_alive():
            _worker = threading.Thread(target=_work, name="tqsv-stage-compile", daemon=True)
            _worker.start()
    _jobs.put(job)
    return job


def wait_all(timeout: Optional[float] = None) -> bool:
    """Block until every queued background compile has finished (True) or the
    timeout has passed (False)."""
    import time
    end = None if timeout is None else time.monotonic() + timeout
    while True:
        with _lock:
            jobs = list(_pending)
        if not jobs:
            return True
        for j in jobs:
            left = None if end is None else max(0.0, end - time.monotonic())
            if not j.done.wait(left):
                return False


def _at_exit():
    """Let a background compile stop at its next source boundary and finish
    before the HIP runtime is torn down."""
    _shutdown.set()
    wait_all(timeout=60)


atexit.register(_at_exit)
