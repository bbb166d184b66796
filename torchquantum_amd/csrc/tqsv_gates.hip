// Per-gate statevector kernels: one launch applies one gate to every batch
// row.  These replace the reference's apply_unitary_bmm
// (torchquantum/functional/gate_wrapper.py:96-140), which permutes the whole
// state, runs a K=2^k batched GEMM and permutes back (about two full-state
// passes per gate).  Here a gate is one coalesced pass that touches only the
// amplitudes the gate can change:
//   * dense 1q/2q:  read+write every amplitude once (controls shrink the set);
//   * swap:         only the |01>/|10> halves;
//   * diagonal:     only the amplitudes whose control bits match.
// Index math is 64-bit so that bsz * 2^n may exceed 2^31.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "tqsv_common.h"

namespace tqsv {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return E_LAUNCH;
  }
  return OK;
}

static int g_num_cu = 0;

int num_cu() {
  if (g_num_cu == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      g_num_cu = v;
    else
      g_num_cu = 256;
  }
  return g_num_cu;
}

// Grid size for a grid-stride loop over `work` units of `per_block` each.
static unsigned grid_for(uint64_t work, unsigned per_block) {
  uint64_t blocks = (work + per_block - 1) / per_block;
  uint64_t cap = (uint64_t)num_cu() * 16;
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  return (unsigned)blocks;
}

static bool build_insert(InsertSpec& s, uint64_t mask, int n) {
  s.nb = 0;
  for (int p = 0; p < n; ++p)
    if ((mask >> p) & 1ull) {
      if (s.nb >= 12) return false;
      s.pos[s.nb++] = p;
    }
  return true;
}

// ---------------------------------------------------------------------------
// basis-state initialisation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_init_basis(float4* __restrict__ st, uint64_t n_pairs, int n,
                                                   uint64_t index, int reps) {
  const uint64_t row_mask = (1ull << n) - 1ull;
  uint64_t i = tile_first(reps);
  for (int r = 0; r < reps; ++r, i += 256) {
    if (i >= n_pairs) return;
    const uint64_t a = 2 * i;  // amplitudes a, a+1 (same row: rows have even length)
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((a & row_mask) == (index & ~1ull)) {
      if (index & 1ull) v.z = 1.f; else v.x = 1.f;
    }
    st[i] = v;
  }
}

__global__ void k_init_basis_n0(float2* __restrict__ st, uint64_t bsz) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < bsz;
       i += (uint64_t)gridDim.x * blockDim.x)
    st[i] = make_float2(1.f, 0.f);
}

// ---------------------------------------------------------------------------
// 1-qubit (optionally controlled) dense gate
// ---------------------------------------------------------------------------
template <bool BATCHED>
__device__ __forceinline__ void load_m2(const float2* __restrict__ mat, int64_t bstride, uint64_t row,
                                        float2 (&m)[4]) {
  const float2* p = BATCHED ? mat + row * bstride : mat;
#pragma unroll
  for (int q = 0; q < 4; ++q) m[q] = p[q];
}

// target bit 0: the pair is adjacent -> one 16-byte access per pair
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_1q_t0(float4* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                              uint64_t cval, const float2* __restrict__ mat, int64_t bstride,
                                              int lg_items_row, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t i0 = insert_zeros(it, ins) | cval;
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, it >> lg_items_row, m);
    const float4 v = ld4<NT>(st + (i0 >> 1));
    const float2 a0 = make_float2(v.x, v.y), a1 = make_float2(v.z, v.w);
    const float2 b0 = cfma(m[1], a1, cmul(m[0], a0));
    const float2 b1 = cfma(m[3], a1, cmul(m[2], a0));
    st4<NT>(st + (i0 >> 1), make_float4(b0.x, b0.y, b1.x, b1.y));
  }
}

__device__ __forceinline__ void rot2(const float2 (&m)[4], const float4 u, const float4 w, float4& uo, float4& wo) {
  const float2 a0 = make_float2(u.x, u.y), a0n = make_float2(u.z, u.w);
  const float2 a1 = make_float2(w.x, w.y), a1n = make_float2(w.z, w.w);
  const float2 b0 = cfma(m[1], a1, cmul(m[0], a0));
  const float2 b1 = cfma(m[3], a1, cmul(m[2], a0));
  const float2 b0n = cfma(m[1], a1n, cmul(m[0], a0n));
  const float2 b1n = cfma(m[3], a1n, cmul(m[2], a0n));
  uo = make_float4(b0.x, b0.y, b0n.x, b0n.y);
  wo = make_float4(b1.x, b1.y, b1n.x, b1n.y);
}

// Target bit 1..8, bit 0 free, controls only on bits >= 9.  Items are float4
// (two adjacent amplitudes) with the control bits removed; a workgroup's 256
// items are one contiguous 4 KiB tile holding both halves of its pairs (the
// target is float4 bit tb4 = target - 1 <= 7).  Each lane reads and writes
// one float4 and trades it with the lane holding the partner: a wave shuffle
// for tb4 <= 5, LDS for tb4 = 6, 7.  Requires n_items % 256 == 0.
template <bool BATCHED, bool NT, bool LDSX>
__global__ __launch_bounds__(256) void k_1q_tile(float4* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                                uint64_t cval4, int tb4, const float2* __restrict__ mat,
                                                int64_t bstride, int lg_items_row, int reps) {
  __shared__ float4 xch[LDSX ? 256 : 1];
  const int S = 1 << tb4;
  const bool hi = (threadIdx.x >> tb4) & 1;
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;  // whole workgroups leave together (n_items % 256 == 0)
    const uint64_t f = insert_zeros(it, ins) | cval4;
    const float4 v = ld4<NT>(st + f);
    float4 p;
    if constexpr (LDSX) {
      xch[threadIdx.x] = v;
      __syncthreads();
      p = xch[threadIdx.x ^ S];
      __syncthreads();
    } else {
      p.x = __shfl_xor(v.x, S);
      p.y = __shfl_xor(v.y, S);
      p.z = __shfl_xor(v.z, S);
      p.w = __shfl_xor(v.w, S);
    }
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, it >> lg_items_row, m);
    // lane with target 0: m00 own + m01 partner; target 1: m11 own + m10 partner
    const float2 co = hi ? m[3] : m[0], cp = hi ? m[2] : m[1];
    const float2 o = cfma(cp, make_float2(p.x, p.y), cmul(co, make_float2(v.x, v.y)));
    const float2 on = cfma(cp, make_float2(p.z, p.w), cmul(co, make_float2(v.z, v.w)));
    st4<NT>(st + f, make_float4(o.x, o.y, on.x, on.y));
  }
}

// Target bit >= 1 and bit 0 free (any controls off bit 0): items are float4
// pairs (f0, f0 | S) with the target and control bits removed (ins4 and cval4
// in float4 bit positions, S = 1 << (target - 1)).
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_1q_pair(float4* __restrict__ st, uint64_t n_pairs, InsertSpec ins4,
                                                uint64_t cval4, uint64_t S, const float2* __restrict__ mat,
                                                int64_t bstride, int lg_pairs_row, int reps) {
  uint64_t q = tile_first(reps);
  for (int r = 0; r < reps; ++r, q += 256) {
    if (q >= n_pairs) return;
    const uint64_t f0 = insert_zeros(q, ins4) | cval4, f1 = f0 | S;
    const float4 u = ld4<NT>(st + f0), w = ld4<NT>(st + f1);
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, q >> lg_pairs_row, m);
    float4 uo, wo;
    rot2(m, u, w, uo, wo);
    st4<NT>(st + f0, uo);
    st4<NT>(st + f1, wo);
  }
}

// generic scalar path (bit 0 is a control bit)
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_1q_gen(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                               uint64_t cval, uint64_t tmask, const float2* __restrict__ mat,
                                               int64_t bstride, int lg_items_row, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t i0 = insert_zeros(it, ins) | cval;
    const uint64_t i1 = i0 | tmask;
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, it >> lg_items_row, m);
    const float2 a0 = st[i0], a1 = st[i1];
    st[i0] = cfma(m[1], a1, cmul(m[0], a0));
    st[i1] = cfma(m[3], a1, cmul(m[2], a0));
  }
}

// ---------------------------------------------------------------------------
// 2-qubit (optionally controlled) dense gate; matrix index = 2*b(hi) + b(lo)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mv4(const float2* __restrict__ p, const float2 (&a)[4], float2 (&b)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float2 acc = cmul(p[4 * r], a[0]);
#pragma unroll
    for (int c = 1; c < 4; ++c) acc = cfma(p[4 * r + c], a[c], acc);
    b[r] = acc;
  }
}

// bit 0 free: items are float4 quadruples (ins4/cval4/mhi4/mlo4 in float4 bit
// positions), two adjacent amplitudes per access
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_2q_v2(float4* __restrict__ st, uint64_t n_quads, InsertSpec ins4,
                                              uint64_t cval4, uint64_t mhi4, uint64_t mlo4,
                                              const float2* __restrict__ mat, int64_t bstride, int lg_quads_row,
                                              int reps) {
  uint64_t q = tile_first(reps);
  for (int r = 0; r < reps; ++r, q += 256) {
    if (q >= n_quads) return;
    const uint64_t base = insert_zeros(q, ins4) | cval4;
    const uint64_t idx[4] = {base, base | mlo4, base | mhi4, base | mhi4 | mlo4};
    const float2* p = BATCHED ? mat + (q >> lg_quads_row) * bstride : mat;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ld4<NT>(st + idx[k]);
    float2 a[4], b[4], an[4], bn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = make_float2(v[k].x, v[k].y);
      an[k] = make_float2(v[k].z, v[k].w);
    }
    mv4(p, a, b);
    mv4(p, an, bn);
#pragma unroll
    for (int k = 0; k < 4; ++k) st4<NT>(st + idx[k], make_float4(b[k].x, b[k].y, bn[k].x, bn[k].y));
  }
}

template <bool BATCHED>
__global__ __launch_bounds__(256) void k_2q_gen(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                               uint64_t cval, uint64_t mhi, uint64_t mlo,
                                               const float2* __restrict__ mat, int64_t bstride, int lg_items_row,
                                               int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t base = insert_zeros(it, ins) | cval;
    const uint64_t idx[4] = {base, base | mlo, base | mhi, base | mhi | mlo};
    const float2* p = BATCHED ? mat + (it >> lg_items_row) * bstride : mat;
    float2 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = st[idx[k]];
    mv4(p, a, b);
#pragma unroll
    for (int k = 0; k < 4; ++k) st[idx[k]] = b[k];
  }
}

// ---------------------------------------------------------------------------
// diagonal gate: amplitudes whose control bits match get phase[parity(j & pmask)]
// ---------------------------------------------------------------------------
// generic (bit 0 is a control): one amplitude per item
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_diag(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                             uint64_t cval, uint64_t pmask, const float2* __restrict__ ph,
                                             int64_t bstride, int lg_items_row, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t i = insert_zeros(it, ins) | cval;
    const float2* p = BATCHED ? ph + (it >> lg_items_row) * bstride : ph;
    const int par = __popcll(i & pmask) & 1;
    st[i] = cmul(par ? p[1] : p[0], st[i]);
  }
}

// items are float4 (two adjacent amplitudes, each with its own parity) with
// the control bits above bit 0 removed (ins4/cval4 in float4 bit positions;
// no controls = every float4 of the state).  c0: -1 = bit 0 is no control;
// 0 / 1 = bit 0 is a control with that value, so only that amplitude of the
// pair changes (CZ or CCZ on wire n-1 still moves 16-byte items).
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_diag4(float4* __restrict__ st, uint64_t n_items, InsertSpec ins4,
                                              uint64_t cval4, uint64_t pmask, const float2* __restrict__ ph,
                                              int64_t bstride, int lg_items_row, int reps, int c0) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t f = insert_zeros(it, ins4) | cval4;
    const float2* p = BATCHED ? ph + (it >> lg_items_row) * bstride : ph;
    const float2 f0 = p[0], f1 = p[1];
    const uint64_t i = 2 * f;
    const int par0 = __popcll(i & pmask) & 1;
    const int par1 = __popcll((i | 1ull) & pmask) & 1;
    const float4 v = ld4<NT>(st + f);
    float2 a = make_float2(v.x, v.y), b = make_float2(v.z, v.w);
    if (c0 != 1) a = cmul(par0 ? f1 : f0, a);
    if (c0 != 0) b = cmul(par1 ? f1 : f0, b);
    st4<NT>(st + f, make_float4(a.x, a.y, b.x, b.y));
  }
}

// Every float4 of the state (two adjacent amplitudes), each amplitude
// changed only where (index & cmask) == cval.  For controls that include
// float4 bit 0 (amplitude bit 1) the changed float4s are isolated 16-byte
// pieces: moving whole lines in and out costs less than the partial writes
// of k_diag4 (CZ on amplitude bits (1, 0) or (2, 1)).
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_diag4_lines(float4* __restrict__ st, uint64_t n_items, uint64_t cmask,
                                                    uint64_t cval, uint64_t pmask, const float2* __restrict__ ph,
                                                    int64_t bstride, int lg_items_row, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const float2* p = BATCHED ? ph + (it >> lg_items_row) * bstride : ph;
    const float2 f0 = p[0], f1 = p[1];
    const uint64_t i = 2 * it;   // row bits sit above bit n, outside cmask and pmask
    const float4 v = ld4<NT>(st + it);
    float2 a = make_float2(v.x, v.y), b = make_float2(v.z, v.w);
    if ((i & cmask) == cval) a = cmul((__popcll(i & pmask) & 1) ? f1 : f0, a);
    if (((i | 1ull) & cmask) == cval) b = cmul((__popcll((i | 1ull) & pmask) & 1) ? f1 : f0, b);
    st4<NT>(st + it, make_float4(a.x, a.y, b.x, b.y));
  }
}

// ---------------------------------------------------------------------------
// swap of bits a, b (optionally controlled): exchanges |..0..1..> <-> |..1..0..>
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_swap(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                             uint64_t cval, uint64_t ma, uint64_t mb, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t base = insert_zeros(it, ins) | cval;
    const uint64_t i = base | mb, j = base | ma;
    const float2 x = st[i], y = st[j];
    st[i] = y;
    st[j] = x;
  }
}

// bit 0 free: float4 items (positions in float4 bits)
template <bool NT>
__global__ __launch_bounds__(256) void k_swap4(float4* __restrict__ st, uint64_t n_items, InsertSpec ins4,
                                              uint64_t cval4, uint64_t ma4, uint64_t mb4, int reps) {
  uint64_t it = tile_first(reps);
  for (int r = 0; r < reps; ++r, it += 256) {
    if (it >= n_items) return;
    const uint64_t base = insert_zeros(it, ins4) | cval4;
    const uint64_t i = base | mb4, j = base | ma4;
    const float4 x = ld4<NT>(st + i), y = ld4<NT>(st + j);
    st4<NT>(st + i, y);
    st4<NT>(st + j, x);
  }
}

// ---------------------------------------------------------------------------
// dense k-qubit gate, k <= 10: one workgroup stages groups of 2^k amplitudes
// through LDS, each thread produces one output amplitude.
// ---------------------------------------------------------------------------
struct KSpec {
  int k;
  int bits[10];  // bits[0] = most significant bit of the matrix index
};

__device__ __forceinline__ uint64_t kq_offset(const KSpec& ks, int r) {
  uint64_t o = 0;
  for (int q = 0; q < ks.k; ++q)
    if ((r >> (ks.k - 1 - q)) & 1) o |= 1ull << ks.bits[q];
  return o;
}

template <bool BATCHED>
__global__ __launch_bounds__(256) void k_kq(float2* __restrict__ st, uint64_t n_groups, InsertSpec ins,
                                           uint64_t cval, KSpec ks, const float2* __restrict__ mat,
                                           int64_t bstride, int lg_groups_row) {
  extern __shared__ float2 lds[];
  const int d = 1 << ks.k;
  const int gpb = blockDim.x >= (unsigned)d ? blockDim.x / d : 1;  // groups per block pass
  for (uint64_t g0 = (uint64_t)blockIdx.x * gpb; g0 < n_groups; g0 += (uint64_t)gridDim.x * gpb) {
    for (int t = threadIdx.x; t < gpb * d; t += blockDim.x) {
      const uint64_t g = g0 + t / d;
      if (g < n_groups) lds[t] = st[(insert_zeros(g, ins) | cval) + kq_offset(ks, t % d)];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < gpb * d; t += blockDim.x) {
      const uint64_t g = g0 + t / d;
      if (g >= n_groups) continue;
      const int r = t % d;
      const float2* p = (BATCHED ? mat + (g >> lg_groups_row) * bstride : mat) + (uint64_t)r * d;
      const float2* x = lds + (t / d) * d;
      float2 acc = make_float2(0.f, 0.f);
      for (int c = 0; c < d; ++c) acc = cfma(p[c], x[c], acc);
      st[(insert_zeros(g, ins) | cval) + kq_offset(ks, r)] = acc;
    }
    __syncthreads();
  }
}

// Dense k-qubit gate, 3 <= k <= 6, with the matrix staged in LDS (rows padded
// to d + 1 elements so that the lanes of a wave, which read different rows,
// spread over the banks).  Each workgroup owns a contiguous range of groups
// (a multiple of its groups per pass, so with per-row matrices a pass never
// straddles two batch rows) and loads a matrix once per batch row it enters;
// the reads of the inner product then come from LDS instead of a strided
// walk of global memory per thread.
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_kq_lds(float2* __restrict__ st, uint64_t n_groups, InsertSpec ins,
                                               uint64_t cval, KSpec ks, const float2* __restrict__ mat,
                                               int64_t bstride, int lg_groups_row, uint64_t chunk) {
  extern __shared__ float2 lds[];
  const int d = 1 << ks.k, dp = d + 1;
  const int gpb = 256 / d;
  float2* us = lds;                 // d rows of dp
  float2* xs = lds + d * dp;        // gpb groups of d
  const int t = threadIdx.x, r = t % d, gl = t / d;
  const uint64_t gbeg = (uint64_t)blockIdx.x * chunk;
  const uint64_t gend = gbeg + chunk < n_groups ? gbeg + chunk : n_groups;
  const uint64_t off_r = kq_offset(ks, r);
  int64_t cur = -1;
  for (uint64_t g0 = gbeg; g0 < gend; g0 += gpb) {
    const int64_t row = BATCHED ? (int64_t)(g0 >> lg_groups_row) : 0;
    if (row != cur) {              // block-uniform
      __syncthreads();
      const float2* U = mat + (BATCHED ? row * bstride : 0);
      for (int e = t; e < d * d; e += 256) us[(e / d) * dp + (e % d)] = U[e];
      cur = row;
    }
    const uint64_t g = g0 + gl;
    const uint64_t a = (insert_zeros(g, ins) | cval) + off_r;
    if (g < gend) xs[t] = st[a];
    __syncthreads();
    if (g < gend) {
      const float2* up = us + r * dp;
      const float2* x = xs + gl * d;
      float2 acc = make_float2(0.f, 0.f);
      for (int c = 0; c < d; ++c) acc = cfma(up[c], x[c], acc);
      st[a] = acc;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// dense k-qubit gate on the matrix cores, k = 5, 6 (one matrix for all rows).
// Out = U In over groups of d = 2^k amplitudes is the real GEMM
//     [Or; Oi] = [[Ur, -Ui], [Ui, Ur]] [Ir; Ii]     (2d x 2d) x (2d x 64)
// for a tile of 64 groups, on v_mfma_f32_16x16x4f32.  A workgroup (4 waves)
// keeps its rows of the real matrix as MFMA A-fragments in registers for the
// whole launch, stages each tile through LDS as [2d][64] floats (re rows,
// then im rows), and streams tiles: 4 per workgroup.  The tile's amplitudes
// are loaded and stored in address order (the tile spans the k target bits
// and the 6 lowest free bits, so its offsets from the tile base are the same
// for every tile and are computed once per thread).
// ---------------------------------------------------------------------------
struct MSpec {
  int k;
  int bits[6];   // target bits, bits[0] = matrix MSB
  int sp[12];    // target bits and the 6 tile-column bits, ascending
  int jpos[6];   // the 6 tile-column bits (lowest free bits), ascending
};

constexpr int MQ_RS = 80;      // LDS row stride in floats (64 columns + 16 pad: rows 2 apart hit other banks)
constexpr int MQ_TILES = 4;    // tiles per workgroup

typedef float mq_v4 __attribute__((ext_vector_type(4)));

template <int K, bool PAIR>
__global__ __launch_bounds__(256) void k_kq_mfma(float2* __restrict__ st, uint64_t n_tiles, InsertSpec ins,
                                                uint64_t cval, MSpec ms, const float2* __restrict__ U) {
  constexpr int D = 1 << K;        // amplitudes per group
  constexpr int M2 = 2 * D;        // real rows / cols
  constexpr int MT = D / 32;       // 16-row m-tiles per wave
  constexpr int KS = M2 / 4;       // k-steps
  constexpr int EPT = D / 4;       // tile elements per thread (64 * D / 256)
  constexpr int NL = PAIR ? EPT / 2 : EPT;  // global accesses per thread (float4 when PAIR)
  __shared__ float X[M2 * MQ_RS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // A fragments: W[16 * (wave * MT + mt) + lane % 16][4 * ks + lane / 16]
  float wf[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int i = 16 * (wave * MT + mt) + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kk = 4 * ks + (lane >> 4);
      const float2 u = U[(i % D) * D + (kk % D)];
      const bool ri = i >= D, rk = kk >= D;
      wf[mt][ks] = (ri == rk) ? u.x : (ri ? u.y : -u.y);
    }
  }
  // this thread's tile elements (element index e: bit b at tile position
  // sp[b]); PAIR: elements 2 p and 2 p + 1 are adjacent amplitudes (sp[0] = 0)
  uint32_t rel[NL];
  int lds[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int idx = PAIR ? 2 * (tid + 256 * (e >> 1)) + (e & 1) : tid + 256 * e;
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < K + 6; ++b) r |= (uint32_t)((idx >> b) & 1) << ms.sp[b];
    int c = 0, j = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) c |= (int)((r >> ms.bits[q]) & 1u) << (K - 1 - q);
#pragma unroll
    for (int q = 0; q < 6; ++q) j |= (int)((r >> ms.jpos[q]) & 1u) << q;
    if (!PAIR || !(e & 1)) rel[PAIR ? e >> 1 : e] = r;
    lds[e] = c * MQ_RS + j;
  }
  float2 v[EPT];
  auto load_tile = [&](uint64_t tile) {
    const uint64_t base = insert_zeros(tile << 6, ins) | cval;
    if constexpr (PAIR) {
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const float4 a = *reinterpret_cast<const float4*>(st + base + rel[q]);
        v[2 * q] = make_float2(a.x, a.y);
        v[2 * q + 1] = make_float2(a.z, a.w);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NL; ++q) v[q] = st[base + rel[q]];
    }
  };
  uint64_t tile = (uint64_t)blockIdx.x * MQ_TILES;
  if (tile >= n_tiles) return;
  load_tile(tile);
  for (int t = 0; t < MQ_TILES; ++t, ++tile) {
    if (tile >= n_tiles) return;  // uniform per workgroup
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      X[lds[e]] = v[e].x;
      X[D * MQ_RS + lds[e]] = v[e].y;
    }
    __syncthreads();
    // the next tile's loads are in flight while the matrix cores run
    if (t + 1 < MQ_TILES && tile + 1 < n_tiles) load_tile(tile + 1);
    mq_v4 acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = (mq_v4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      float bf[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) bf[nt] = X[(4 * ks + (lane >> 4)) * MQ_RS + 16 * nt + (lane & 15)];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(wf[mt][ks], bf[nt], acc[mt][nt], 0, 0, 0);
    }
    __syncthreads();
    // D[i][j]: i = 16 m-tile + 4 (lane / 16) + v, j = 16 n-tile + lane % 16
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          X[(16 * (wave * MT + mt) + 4 * (lane >> 4) + r) * MQ_RS + 16 * nt + (lane & 15)] = acc[mt][nt][r];
    __syncthreads();
    const uint64_t base = insert_zeros(tile << 6, ins) | cval;
    if constexpr (PAIR) {
#pragma unroll
      for (int q = 0; q < NL; ++q) {
        const int l0 = lds[2 * q], l1 = lds[2 * q + 1];
        *reinterpret_cast<float4*>(st + base + rel[q]) =
            make_float4(X[l0], X[D * MQ_RS + l0], X[l1], X[D * MQ_RS + l1]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NL; ++q) st[base + rel[q]] = make_float2(X[lds[q]], X[D * MQ_RS + lds[q]]);
    }
    __syncthreads();
  }
}

}  // namespace tqsv

using namespace tqsv;

// ===========================================================================
// C ABI
// ===========================================================================
TQSV_API const char* tqsv_last_error(void) { return g_err; }

TQSV_API int tqsv_version(void) { return 1; }

TQSV_API int tqsv_num_cu(void) { return num_cu(); }

static int check_state(const void* st, int64_t bsz, int n) {
  if (!st) { set_error("state pointer is null"); return E_ARG; }
  if (bsz < 1) { set_error("bsz must be >= 1 (got %lld)", (long long)bsz); return E_ARG; }
  if (n < 0 || n > 40) { set_error("n_wires must be in [0, 40] (got %d)", n); return E_ARG; }
  return OK;
}

static int check_bit(int b, int n, const char* what) {
  if (b < 0 || b >= n) { set_error("%s bit %d out of range for %d qubits", what, b, n); return E_ARG; }
  return OK;
}

TQSV_API int tqsv_init_basis(hipStream_t s, void* state, int64_t bsz, int n, int64_t index) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if (index < 0 || (n < 63 && index >= (1ll << n))) { set_error("basis index out of range"); return E_ARG; }
  if (n == 0) {
    k_init_basis_n0<<<grid_for(bsz, 256), 256, 0, s>>>((float2*)state, (uint64_t)bsz);
  } else {
    const uint64_t pairs = ((uint64_t)bsz << n) / 2;
    const Tiles t = tiles_for(pairs);
    k_init_basis<<<t.grid, 256, 0, s>>>((float4*)state, pairs, n, (uint64_t)index, t.reps);
  }
  return check_launch("tqsv_init_basis");
}

// (BATCHED, NT) instantiation of a 256-thread tiled kernel
#define LAUNCH_BN(kern, b, nt, t, s, ...)                                  \
  do {                                                                     \
    if (b) {                                                               \
      if (nt) kern<true, true><<<(t).grid, 256, 0, s>>>(__VA_ARGS__);      \
      else kern<true, false><<<(t).grid, 256, 0, s>>>(__VA_ARGS__);        \
    } else {                                                               \
      if (nt) kern<false, true><<<(t).grid, 256, 0, s>>>(__VA_ARGS__);     \
      else kern<false, false><<<(t).grid, 256, 0, s>>>(__VA_ARGS__);       \
    }                                                                      \
  } while (0)

template <bool LDSX>
static void launch_1q_tile(bool b, bool nt, Tiles t, hipStream_t s, float4* st, uint64_t n_items,
                           const InsertSpec& ins, uint64_t cval4, int tb4, const float2* m, int64_t bs, int lg) {
  if (b) {
    if (nt) k_1q_tile<true, true, LDSX><<<t.grid, 256, 0, s>>>(st, n_items, ins, cval4, tb4, m, bs, lg, t.reps);
    else k_1q_tile<true, false, LDSX><<<t.grid, 256, 0, s>>>(st, n_items, ins, cval4, tb4, m, bs, lg, t.reps);
  } else {
    if (nt) k_1q_tile<false, true, LDSX><<<t.grid, 256, 0, s>>>(st, n_items, ins, cval4, tb4, m, bs, lg, t.reps);
    else k_1q_tile<false, false, LDSX><<<t.grid, 256, 0, s>>>(st, n_items, ins, cval4, tb4, m, bs, lg, t.reps);
  }
}

TQSV_API int tqsv_apply_1q(hipStream_t s, void* state, int64_t bsz, int n, int tbit, uint64_t cmask,
                           uint64_t cval, const void* mat, int64_t mat_bstride) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if ((rc = check_bit(tbit, n, "target"))) return rc;
  if (!mat) { set_error("matrix pointer is null"); return E_ARG; }
  const uint64_t tmask = 1ull << tbit;
  if ((cmask & tmask) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) {
    set_error("control mask overlaps target / out of range / value outside mask");
    return E_ARG;
  }
  const int c = __builtin_popcountll(cmask);
  const bool b = mat_bstride != 0;
  const bool nt = stream_nt(bsz, n);
  const float2* m = (const float2*)mat;
  InsertSpec ins;
  if (tbit == 0 || (cmask & 1ull)) {
    if (!build_insert(ins, cmask | tmask, n)) { set_error("too many control bits (max 11)"); return E_LIMIT; }
    const int lg = n - 1 - c;
    const uint64_t n_items = (uint64_t)bsz << lg;
    const Tiles t = tiles_for(n_items);
    if (tbit == 0) {
      LAUNCH_BN(k_1q_t0, b, nt, t, s, (float4*)state, n_items, ins, cval, m, mat_bstride, lg, t.reps);
    } else if (b) {
      k_1q_gen<true><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, tmask, m, mat_bstride, lg, t.reps);
    } else {
      k_1q_gen<false><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, tmask, m, 0, lg, t.reps);
    }
    return check_launch("tqsv_apply_1q");
  }
  // bit 0 free: float4 positions are amplitude positions - 1
  const int tb4 = tbit - 1;
  const uint64_t c4 = cmask >> 1, cval4 = cval >> 1;
  const int lg_items = n - 1 - c;  // float4 per row with the control bits removed
  if (tb4 <= 7 && (c4 & 0xFFull) == 0 && lg_items >= 8) {
    if (!build_insert(ins, c4, n - 1)) { set_error("too many control bits (max 11)"); return E_LIMIT; }
    const uint64_t n_items = (uint64_t)bsz << lg_items;
    const Tiles t = tiles_for(n_items);
    if (tb4 >= 6)
      launch_1q_tile<true>(b, nt, t, s, (float4*)state, n_items, ins, cval4, tb4, m, mat_bstride, lg_items);
    else
      launch_1q_tile<false>(b, nt, t, s, (float4*)state, n_items, ins, cval4, tb4, m, mat_bstride, lg_items);
  } else {
    if (!build_insert(ins, c4 | (1ull << tb4), n - 1)) { set_error("too many control bits (max 11)"); return E_LIMIT; }
    const int lg = n - 2 - c;
    const uint64_t n_pairs = (uint64_t)bsz << lg;
    const Tiles t = tiles_for(n_pairs);
    LAUNCH_BN(k_1q_pair, b, nt, t, s, (float4*)state, n_pairs, ins, cval4, 1ull << tb4, m, mat_bstride, lg, t.reps);
  }
  return check_launch("tqsv_apply_1q");
}

TQSV_API int tqsv_apply_2q(hipStream_t s, void* state, int64_t bsz, int n, int bit_hi, int bit_lo,
                           uint64_t cmask, uint64_t cval, const void* mat, int64_t mat_bstride) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if ((rc = check_bit(bit_hi, n, "hi")) || (rc = check_bit(bit_lo, n, "lo"))) return rc;
  if (bit_hi == bit_lo) { set_error("2-qubit gate on a repeated wire"); return E_ARG; }
  if (!mat) { set_error("matrix pointer is null"); return E_ARG; }
  const uint64_t mhi = 1ull << bit_hi, mlo = 1ull << bit_lo;
  if ((cmask & (mhi | mlo)) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) {
    set_error("control mask overlaps targets / out of range / value outside mask");
    return E_ARG;
  }
  const int c = __builtin_popcountll(cmask);
  const uint64_t all = cmask | mhi | mlo;
  const bool b = mat_bstride != 0;
  const float2* m = (const float2*)mat;
  InsertSpec ins;
  if (!(all & 1ull)) {
    if (!build_insert(ins, all >> 1, n - 1)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
    const int lg = n - 3 - c;
    const uint64_t n_quads = (uint64_t)bsz << lg;
    const Tiles t = tiles_for(n_quads);
    LAUNCH_BN(k_2q_v2, b, stream_nt(bsz, n), t, s, (float4*)state, n_quads, ins, cval >> 1, mhi >> 1, mlo >> 1, m,
              mat_bstride, lg, t.reps);
  } else {
    if (!build_insert(ins, all, n)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
    const int lg = n - 2 - c;
    const uint64_t n_items = (uint64_t)bsz << lg;
    const Tiles t = tiles_for(n_items);
    if (b) k_2q_gen<true><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, mhi, mlo, m, mat_bstride, lg, t.reps);
    else k_2q_gen<false><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, mhi, mlo, m, 0, lg, t.reps);
  }
  return check_launch("tqsv_apply_2q");
}

TQSV_API int tqsv_apply_diag(hipStream_t s, void* state, int64_t bsz, int n, uint64_t cmask, uint64_t cval,
                             uint64_t pmask, const void* phases, int64_t ph_bstride) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if (!phases) { set_error("phase pointer is null"); return E_ARG; }
  if ((n < 64 && ((cmask | pmask) >> n)) || (cval & ~cmask) || (cmask & pmask)) {
    set_error("diag masks out of range or overlapping");
    return E_ARG;
  }
  const int c = __builtin_popcountll(cmask);
  const bool b = ph_bstride != 0;
  const float2* p = (const float2*)phases;
  InsertSpec ins;
  if (n >= 1) {
    // float4 items; a control on bit 0 is applied inside the item
    const uint64_t cm4 = (cmask & ~1ull) >> 1;
    const int c0 = (cmask & 1ull) ? (int)(cval & 1ull) : -1;
    if (cm4 & 1ull) {
      // the touched float4s are isolated 16-byte pieces: whole-line pass
      const int lg = n - 1;
      const uint64_t n_items = (uint64_t)bsz << lg;
      const Tiles t = tiles_for(n_items);
      LAUNCH_BN(k_diag4_lines, b, stream_nt(bsz, n), t, s, (float4*)state, n_items, cmask, cval, pmask, p,
                ph_bstride, lg, t.reps);
      return check_launch("tqsv_apply_diag");
    }
    if (!build_insert(ins, cm4, n - 1)) { set_error("too many control bits (max 12)"); return E_LIMIT; }
    const int lg = n - 1 - __builtin_popcountll(cm4);
    const uint64_t n_items = (uint64_t)bsz << lg;
    const Tiles t = tiles_for(n_items);
    LAUNCH_BN(k_diag4, b, stream_nt(bsz, n), t, s, (float4*)state, n_items, ins, (cval & ~1ull) >> 1, pmask, p,
              ph_bstride, lg, t.reps, c0);
    return check_launch("tqsv_apply_diag");
  }
  if (!build_insert(ins, cmask, n)) { set_error("too many control bits (max 12)"); return E_LIMIT; }
  const int lg = n - c;
  const uint64_t n_items = (uint64_t)bsz << lg;
  const Tiles t = tiles_for(n_items);
  if (b) k_diag<true><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, pmask, p, ph_bstride, lg, t.reps);
  else k_diag<false><<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, pmask, p, 0, lg, t.reps);
  return check_launch("tqsv_apply_diag");
}

TQSV_API int tqsv_apply_swap(hipStream_t s, void* state, int64_t bsz, int n, int bit_a, int bit_b,
                             uint64_t cmask, uint64_t cval) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if ((rc = check_bit(bit_a, n, "swap a")) || (rc = check_bit(bit_b, n, "swap b"))) return rc;
  if (bit_a == bit_b) { set_error("swap on a repeated wire"); return E_ARG; }
  const uint64_t ma = 1ull << bit_a, mb = 1ull << bit_b;
  if ((cmask & (ma | mb)) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) {
    set_error("control mask overlaps swap bits / out of range");
    return E_ARG;
  }
  const int c = __builtin_popcountll(cmask);
  const uint64_t all = cmask | ma | mb;
  InsertSpec ins;
  if (!(all & 1ull)) {
    if (!build_insert(ins, all >> 1, n - 1)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
    const uint64_t n_items = (uint64_t)bsz << (n - 3 - c);
    const Tiles t = tiles_for(n_items);
    if (stream_nt(bsz, n))
      k_swap4<true><<<t.grid, 256, 0, s>>>((float4*)state, n_items, ins, cval >> 1, ma >> 1, mb >> 1, t.reps);
    else
      k_swap4<false><<<t.grid, 256, 0, s>>>((float4*)state, n_items, ins, cval >> 1, ma >> 1, mb >> 1, t.reps);
    return check_launch("tqsv_apply_swap");
  }
  if (!build_insert(ins, all, n)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
  const uint64_t n_items = (uint64_t)bsz << (n - 2 - c);
  const Tiles t = tiles_for(n_items);
  k_swap<<<t.grid, 256, 0, s>>>((float2*)state, n_items, ins, cval, ma, mb, t.reps);
  return check_launch("tqsv_apply_swap");
}

// bits: k host-side entries, bits[0] is the most significant bit of the
// matrix index (the reference's wires[0], gate_wrapper.py:120-127).
TQSV_API int tqsv_apply_kq(hipStream_t s, void* state, int64_t bsz, int n, int k, const int* bits,
                           uint64_t cmask, uint64_t cval, const void* mat, int64_t mat_bstride) {
  int rc = check_state(state, bsz, n);
  if (rc) return rc;
  if (k < 1 || k > 10 || k > n) { set_error("dense gate width k=%d outside [1, min(10, n)]", k); return E_LIMIT; }
  if (!bits || !mat) { set_error("null pointer argument"); return E_ARG; }
  KSpec ks;
  ks.k = k;
  uint64_t tm = 0;
  for (int q = 0; q < k; ++q) {
    if ((rc = check_bit(bits[q], n, "dense"))) return rc;
    if (tm & (1ull << bits[q])) { set_error("dense gate on a repeated wire"); return E_ARG; }
    tm |= 1ull << bits[q];
    ks.bits[q] = bits[q];
  }
  if ((cmask & tm) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) {
    set_error("control mask overlaps targets / out of range");
    return E_ARG;
  }
  InsertSpec ins;
  if (!build_insert(ins, cmask | tm, n)) { set_error("too many target+control bits (max 12)"); return E_LIMIT; }
  const int cc = __builtin_popcountll(cmask);
  if ((k == 5 || k == 6) && mat_bstride == 0 && n - k - cc >= 6) {
    // matrix cores: tiles of 64 groups whose columns are the 6 lowest free bits
    MSpec ms;
    ms.k = k;
    for (int q = 0; q < k; ++q) ms.bits[q] = bits[q];
    int nj = 0;
    for (int p = 0; p < n && nj < 6; ++p)
      if (!(((cmask | tm) >> p) & 1ull)) ms.jpos[nj++] = p;
    int ns = 0;
    for (int p = 0; p < n; ++p)
      if (((tm >> p) & 1ull) || (nj && p <= ms.jpos[5] && [&] { for (int q = 0; q < 6; ++q) if (ms.jpos[q] == p) return true; return false; }()))
        ms.sp[ns++] = p;
    const uint64_t n_tiles = (uint64_t)bsz << (n - k - cc - 6);
    const unsigned grid = (unsigned)((n_tiles + MQ_TILES - 1) / MQ_TILES);
    // bit 0 in the tile: element pairs are adjacent amplitudes (16-byte accesses)
    const bool pair = ms.sp[0] == 0;
    if (k == 5) {
      if (pair) k_kq_mfma<5, true><<<grid, 256, 0, s>>>((float2*)state, n_tiles, ins, cval, ms, (const float2*)mat);
      else k_kq_mfma<5, false><<<grid, 256, 0, s>>>((float2*)state, n_tiles, ins, cval, ms, (const float2*)mat);
    } else {
      if (pair) k_kq_mfma<6, true><<<grid, 256, 0, s>>>((float2*)state, n_tiles, ins, cval, ms, (const float2*)mat);
      else k_kq_mfma<6, false><<<grid, 256, 0, s>>>((float2*)state, n_tiles, ins, cval, ms, (const float2*)mat);
    }
    return check_launch("tqsv_apply_kq");
  }
  const int d = 1 << k;
  const int lg_groups_row = n - k - __builtin_popcountll(cmask);
  const uint64_t n_groups = (uint64_t)bsz << lg_groups_row;
  if (k >= 3 && k <= 6 && (mat_bstride == 0 || (1ll << lg_groups_row) >= 256 / d)) {
    // matrix in LDS; contiguous group ranges of a multiple of the groups per pass
    const uint64_t gpb = 256 / d;
    const uint64_t want = (uint64_t)num_cu() * 8;
    uint64_t chunk = (n_groups + want - 1) / want;
    chunk = ((chunk + gpb - 1) / gpb) * gpb;
    const unsigned grid = (unsigned)((n_groups + chunk - 1) / chunk);
    const size_t shm = sizeof(float2) * ((size_t)d * (d + 1) + (size_t)gpb * d);
    if (mat_bstride)
      k_kq_lds<true><<<grid, 256, shm, s>>>((float2*)state, n_groups, ins, cval, ks, (const float2*)mat, mat_bstride,
                                            lg_groups_row, chunk);
    else
      k_kq_lds<false><<<grid, 256, shm, s>>>((float2*)state, n_groups, ins, cval, ks, (const float2*)mat, 0,
                                             lg_groups_row, chunk);
    return check_launch("tqsv_apply_kq");
  }
  const unsigned threads = 256;
  const int gpb = threads >= (unsigned)d ? threads / d : 1;
  const size_t shm = sizeof(float2) * (size_t)gpb * d;
  const unsigned grid = grid_for(n_groups, gpb);
  if (mat_bstride)
    k_kq<true><<<grid, threads, shm, s>>>((float2*)state, n_groups, ins, cval, ks, (const float2*)mat,
                                          mat_bstride, lg_groups_row);
  else
    k_kq<false><<<grid, threads, shm, s>>>((float2*)state, n_groups, ins, cval, ks, (const float2*)mat, 0,
                                           lg_groups_row);
  return check_launch("tqsv_apply_kq");
}

// ---------------------------------------------------------------------------
// Per-gate plan executor: a run of gates, one launch each, issued from native
// code so the host loop costs a few microseconds per gate.  Record layout
// (TQSV_GATE_WORDS int64 words per gate):
//   [0] kind: 1 dense 1q, 2 dense 2q, 3 swap, 4 diagonal, 5 dense k-qubit
//   [1] k, the number of entries used in [2..11]
//   [2..11] bits: target(s) (bits[0] = matrix MSB) / swap bits / parity bits
//   [12] cmask  [13] cval
//   [14] offset of the gate's matrix (phases) in `mats`, in complex64 elements
//   [15] matrix stride between batch rows (complex elements, 0 = shared)
// ---------------------------------------------------------------------------
static const int kGateWords = 16;

TQSV_API int tqsv_gate_words(void) { return kGateWords; }

TQSV_API int tqsv_run_gates(hipStream_t s, void* state, int64_t bsz, int n, const int64_t* gates, int n_gates,
                            const void* mats) {
  if (n_gates < 0 || (n_gates > 0 && !gates)) { set_error("tqsv_run_gates: bad gate list"); return E_ARG; }
  const char* base = (const char*)mats;
  for (int g = 0; g < n_gates; ++g) {
    const int64_t* w = gates + (int64_t)g * kGateWords;
    const int k = (int)w[1];
    if (k < 0 || k > 10) { set_error("tqsv_run_gates: gate %d has k=%d outside [0, 10]", g, k); return E_ARG; }
    int bits[10];
    uint64_t pmask = 0;
    for (int q = 0; q < k; ++q) {
      bits[q] = (int)w[2 + q];
      if (bits[q] < 0 || bits[q] >= 64) { set_error("tqsv_run_gates: gate %d bit out of range", g); return E_ARG; }
      pmask |= 1ull << bits[q];
    }
    const uint64_t cmask = (uint64_t)w[12], cval = (uint64_t)w[13];
    const void* m = base ? base + 8 * w[14] : nullptr;
    const int64_t bs = w[15];
    const int kind = (int)w[0];
    int need = kind == 1 ? 1 : (kind == 2 || kind == 3) ? 2 : -1;
    if (need > 0 && k != need) { set_error("tqsv_run_gates: gate %d kind %d needs k=%d", g, kind, need); return E_ARG; }
    int rc;
    switch (kind) {
      case 1: rc = tqsv_apply_1q(s, state, bsz, n, bits[0], cmask, cval, m, bs); break;
      case 2: rc = tqsv_apply_2q(s, state, bsz, n, bits[0], bits[1], cmask, cval, m, bs); break;
      case 3: rc = tqsv_apply_swap(s, state, bsz, n, bits[0], bits[1], cmask, cval); break;
      case 4: rc = tqsv_apply_diag(s, state, bsz, n, cmask, cval, pmask, m, bs); break;
      case 5: rc = tqsv_apply_kq(s, state, bsz, n, k, bits, cmask, cval, m, bs); break;
      default: set_error("tqsv_run_gates: gate %d has unknown kind %d", g, kind); return E_ARG;
    }
    if (rc) {
      char msg[sizeof(g_err)];
      snprintf(msg, sizeof(msg), "%s", g_err);
      set_error("tqsv_run_gates: gate %d: %s", g, msg);
      return rc;
    }
  }
  return OK;
}
