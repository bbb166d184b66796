// Pauli expectation kernels (one read pass of the state per distinct flip mask).
//
// Replaces:
//   * expval / MeasureAll (torchquantum/measurement/measurements.py:282-343):
//     |psi|^2 then one full reduction per wire -> here one pass computes
//     <Z_w> for every wire: <Z_w> = sum_j p_j - 2 * sum_{j: bit(w)=1} p_j;
//   * expval_joint_analytical (measurements.py:232-279), which builds a dense
//     2^n x 2^n Kronecker Hamiltonian (O(4^n) memory) -> here a Pauli string
//     P = i^{nY} X^x Z^z is applied through its flip mask x and phase mask z:
//     <psi|P|psi> = i^{nY} sum_j (-1)^{popcount(j & z)} conj(psi_{j^x}) psi_j.
// Reductions: per-thread partial sums -> 64-lane shuffle tree -> LDS across
// waves -> one partial per (row, chunk, term) -> a second tiny kernel sums
// the chunks in float64.  No float atomics, so results are deterministic.
#include <algorithm>
#include <mutex>
#include <vector>

#include "tqsv_common.h"

namespace tqsv {
int num_cu();
}

namespace tqsv {

constexpr int EV_THREADS = 256;
constexpr int EV_MAXQ = 32;      // wires per <Z_w> launch (n <= 32)
constexpr int EV_MAXT = 16;      // terms per diagonal-term launch

// grid of 256-thread blocks for one wavefront per output
static unsigned sum_grid(int64_t outputs) { return (unsigned)((outputs * 64 + 255) / 256); }

// chunks per row so that the grid has enough blocks; chunk = power of two amps
static int ev_chunk_log2(int64_t bsz, int n) {
  // at least 2 amps per thread per chunk, at most the whole row
  int lg = n;
  const int min_lg = 9 + 2;  // 256 threads * 2 amps * 4 iterations
  const int64_t want_blocks = (int64_t)num_cu() * 8;
  while (lg > min_lg && (bsz << (n - lg)) < want_blocks) --lg;
  if (lg > n) lg = n;
  return lg;
}

__device__ __forceinline__ float block_sum(float v, float* red) {
  // red: >= EV_THREADS / 64 floats of LDS
  v = wave_sum(v, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;  // valid in thread 0
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Loads in flight per thread in the chunk walks below: one 16-byte load per
// trip leaves too few bytes in flight to cover HBM latency.
constexpr int EV_UNROLL = 4;
static_assert(EV_UNROLL == 4 && EV_THREADS == 256, "k_ez_part maps amplitude bits 9, 10 to the unroll slot");

// <Z_w> partials: part[(row * nchunks + chunk) * (n + 1) + q], q < n: sum of p
// over amplitudes with bit q set; q = n: total probability.
__global__ __launch_bounds__(EV_THREADS) void k_ez_part(const float4* __restrict__ st, int n, int lg_chunk,
                                                        float* __restrict__ part) {
  __shared__ float red[EV_THREADS / 64];
  const int nchunks = 1 << (n - lg_chunk);
  const uint64_t row = blockIdx.x >> (n - lg_chunk);
  const uint64_t chunk = blockIdx.x & (nchunks - 1);
  const uint64_t first = (row << n) + (chunk << lg_chunk);  // amplitude index
  const uint64_t npairs = (1ull << lg_chunk) >> 1;
  // Thread t reads float4 i = t + 256 k, i.e. amplitudes j = chunk_base +
  // 2 t + 512 k (+1), all parts on disjoint bits: bit 0 of j is the float4
  // half; bits 1..8 are bits of t (fixed per thread: those sums follow from
  // the thread's total at the end); bits 9, 10 are bits of the unroll slot u
  // (k = trip * 4 + u); bits 11 .. lg_chunk - 1 are bits of the trip
  // (wave-uniform, one add per trip); bits >= lg_chunk are bits of the
  // chunk (constant per block: taken from the total at the end).  A chunk
  // shorter than 512 amplitudes is a whole row (n <= 11), where k = 0 and
  // threads past its end read nothing.
  float S[EV_MAXQ];
#pragma unroll
  for (int q = 0; q < EV_MAXQ; ++q) S[q] = 0.f;
  float tot = 0.f, s0 = 0.f;
  for (uint64_t b = threadIdx.x, k = 0; b < npairs; b += EV_UNROLL * EV_THREADS, k += EV_UNROLL) {
    float4 v[EV_UNROLL];
#pragma unroll
    for (int u = 0; u < EV_UNROLL; ++u) {
      const uint64_t i = b + (uint64_t)u * EV_THREADS;
      v[u] = i < npairs ? st[(first >> 1) + i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float ts = 0.f;
#pragma unroll
    for (int u = 0; u < EV_UNROLL; ++u) {
      const float p0 = v[u].x * v[u].x + v[u].y * v[u].y;
      const float p1 = v[u].z * v[u].z + v[u].w * v[u].w;
      const float pp = p0 + p1;
      ts += pp;
      s0 += p1;
      if (u & 1) S[9] += pp;
      if (u & 2) S[10] += pp;
    }
    tot += ts;
    const uint32_t kk = (uint32_t)k;   // wave-uniform, a multiple of 4
#pragma unroll
    for (int q = 11; q < EV_MAXQ; ++q)
      if (q < lg_chunk && ((kk >> (q - 9)) & 1u)) S[q] += ts;
  }
  const uint64_t cb = chunk << lg_chunk;
#pragma unroll
  for (int q = 9; q < EV_MAXQ; ++q)
    if (q >= lg_chunk) S[q] = ((cb >> q) & 1u) ? tot : 0.f;
  S[0] = s0;
#pragma unroll
  for (int q = 1; q <= 8; ++q)
    if (q < n) S[q] = ((threadIdx.x >> (q - 1)) & 1u) ? tot : 0.f;
  float* out = part + (uint64_t)blockIdx.x * (n + 1);
#pragma unroll
  for (int q = 0; q < EV_MAXQ; ++q) {
    if (q < n) {
      const float s = block_sum(S[q], red);
      if (threadIdx.x == 0) out[q] = s;
    }
  }
  const float s = block_sum(tot, red);
  if (threadIdx.x == 0) out[n] = s;
}

// out[row * n + w] = total - 2 * S[n - 1 - w]; one wavefront per output, its
// lanes striding over the row's chunk partials.
__global__ __launch_bounds__(256) void k_ez_final(const float* __restrict__ part, int n, int nchunks, int64_t bsz,
                                                  float* __restrict__ out) {
  const int64_t idx = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;   // wave-uniform
  const int lane = threadIdx.x & 63;
  if (idx >= bsz * n) return;
  const int64_t row = idx / n;
  const int q = n - 1 - (int)(idx % n);
  double s = 0.0, t = 0.0;
  for (int c = lane; c < nchunks; c += 64) {
    const float* p = part + (row * nchunks + c) * (n + 1);
    s += p[q];
    t += p[n];
  }
  s = wave_sum_d(s);
  t = wave_sum_d(t);
  if (lane == 0) out[idx] = (float)(t - 2.0 * s);
}

// Diagonal Pauli strings (x = 0): sum_j p_j (-1)^{popcount(j & z_t)}, up to
// EV_MAXT terms per launch.  part[(row*nchunks+chunk)*nt + t].
struct ZMasks {
  int nt;
  uint32_t z[EV_MAXT];
};

// Amplitude j = chunk_base + 512 k + 2 t + h (disjoint bits, as in
// k_ez_part), so parity(j & z) = parity(h & z) ^ parity(2 t & z) ^
// parity((chunk_base + 512 k) & z): the h part picks p0 + p1 or p0 - p1 per
// term, the uniform part is a sign per trip, the thread part one sign at the
// end -- one multiply-add per term per float4.
__global__ __launch_bounds__(EV_THREADS) void k_ediag_part(const float4* __restrict__ st, int n, int lg_chunk,
                                                           ZMasks zm, float* __restrict__ part) {
  __shared__ float red[EV_THREADS / 64];
  const int nchunks = 1 << (n - lg_chunk);
  const uint64_t row = blockIdx.x >> (n - lg_chunk);
  const uint64_t chunk = blockIdx.x & (nchunks - 1);
  const uint64_t first = (row << n) + (chunk << lg_chunk);
  const uint64_t npairs = (1ull << lg_chunk) >> 1;
  float acc[EV_MAXT];
#pragma unroll
  for (int t = 0; t < EV_MAXT; ++t) acc[t] = 0.f;
  for (uint64_t b = threadIdx.x, k = 0; b < npairs; b += EV_UNROLL * EV_THREADS, k += EV_UNROLL) {
    float4 v[EV_UNROLL];
#pragma unroll
    for (int u = 0; u < EV_UNROLL; ++u) {
      const uint64_t i = b + (uint64_t)u * EV_THREADS;
      v[u] = i < npairs ? st[(first >> 1) + i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < EV_UNROLL; ++u) {
      const float p0 = v[u].x * v[u].x + v[u].y * v[u].y;
      const float p1 = v[u].z * v[u].z + v[u].w * v[u].w;
      const float ps = p0 + p1, pd = p0 - p1;
      const uint32_t jhi = (uint32_t)((chunk << lg_chunk) + 512ull * (k + u));   // wave-uniform
#pragma unroll
      for (int t = 0; t < EV_MAXT; ++t) {
        if (t < zm.nt) {
          const float sg = (__popc(jhi & zm.z[t]) & 1) ? -1.f : 1.f;
          acc[t] = fmaf(sg, (zm.z[t] & 1u) ? pd : ps, acc[t]);
        }
      }
    }
  }
  float* out = part + (uint64_t)blockIdx.x * zm.nt;
#pragma unroll
  for (int t = 0; t < EV_MAXT; ++t) {
    if (t < zm.nt) {
      const float tsg = (__popc((2u * threadIdx.x) & zm.z[t]) & 1) ? -1.f : 1.f;
      const float s = block_sum(tsg * acc[t], red);
      if (threadIdx.x == 0) out[t] = s;
    }
  }
}

// One off-diagonal string (x != 0): pairs (j, j ^ x) with j's top x-bit clear.
// Partial of sum_j s_j conj(psi_{j^x}) psi_j (complex; times i^{nY} on host).
__global__ __launch_bounds__(EV_THREADS) void k_ex_part(const float2* __restrict__ st, int n, int lg_chunk,
                                                        uint32_t x, uint32_t z, int top, float* __restrict__ part) {
  __shared__ float red[EV_THREADS / 64];
  // items: in-row indices with bit `top` clear; chunk over items
  const int lg_items = n - 1;
  const int lgc = lg_chunk > lg_items ? lg_items : lg_chunk;
  const int nchunks = 1 << (lg_items - lgc);
  const uint64_t row = blockIdx.x >> (lg_items - lgc);
  const uint64_t chunk = blockIdx.x & (nchunks - 1);
  const uint64_t cnt = 1ull << lgc;
  const float2* r = st + (row << n);
  float ar = 0.f, ai = 0.f;
  for (uint64_t i = threadIdx.x; i < cnt; i += blockDim.x) {
    const uint32_t it = (uint32_t)((chunk << lgc) + i);
    const uint32_t lo = it & ((1u << top) - 1u);
    const uint32_t j = ((it >> top) << (top + 1)) | lo;
    const uint32_t k = j ^ x;
    const float2 a = r[j], b = r[k];
    const float sj = (__popc(j & z) & 1) ? -1.f : 1.f;
    const float sk = (__popc(k & z) & 1) ? -1.f : 1.f;
    // sj * conj(b) * a + sk * conj(a) * b
    const float2 c1 = cmulc(b, a);
    const float2 c2 = cmulc(a, b);
    ar += sj * c1.x + sk * c2.x;
    ai += sj * c1.y + sk * c2.y;
  }
  float* out = part + (uint64_t)blockIdx.x * 2;
  const float sr = block_sum(ar, red);
  if (threadIdx.x == 0) out[0] = sr;
  const float si = block_sum(ai, red);
  if (threadIdx.x == 0) out[1] = si;
}

// Partials of Re <a_row|b_row>: block = (row, chunk of 2^lg amplitudes),
// part[row * nchunks + chunk].
__global__ __launch_bounds__(EV_THREADS) void k_rowdot_part(const float2* __restrict__ a,
                                                            const float2* __restrict__ b, int n, int lg_chunk,
                                                            float* __restrict__ part) {
  __shared__ float red[EV_THREADS / 64];
  const uint64_t row = blockIdx.x >> (n - lg_chunk);
  const uint64_t chunk = blockIdx.x & ((1u << (n - lg_chunk)) - 1u);
  const uint64_t base = (row << n) + (chunk << lg_chunk);
  float d = 0.f;
  for (uint64_t i = threadIdx.x; i < (1ull << lg_chunk); i += blockDim.x) {
    const float2 x = a[base + i], y = b[base + i];
    d = fmaf(x.x, y.x, fmaf(x.y, y.y, d));
  }
  const float t = block_sum(d, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sum per-chunk partials of one term: partial of (row, chunk) sits at
// part[(row * nchunks + chunk) * cstride] (complex: two floats re, im).
// out[row * ostride + ooff] = scale_re * Re(sum) - scale_im * Im(sum).
// One wavefront per row, its lanes striding over the chunks.
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part, int nchunks, int cstride,
                                                    int is_complex, int64_t bsz, float scale_re, float scale_im,
                                                    float* __restrict__ out, int64_t ostride, int ooff) {
  const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;   // wave-uniform
  const int lane = threadIdx.x & 63;
  if (row >= bsz) return;
  double sr = 0.0, si = 0.0;
  for (int c = lane; c < nchunks; c += 64) {
    const float* p = part + (row * nchunks + c) * cstride;
    sr += p[0];
    if (is_complex) si += p[1];
  }
  sr = wave_sum_d(sr);
  si = wave_sum_d(si);
  if (lane == 0) out[row * ostride + ooff] = (float)(scale_re * sr - scale_im * si);
}

// out_j (+)= sum_t coef[row, t] * (-1)^{popcount((j^x_t) & z_t)} * psi_{j ^ x_t}
struct PTerms {
  int nt;
  uint32_t x[EV_MAXT];
  uint32_t z[EV_MAXT];
  int col[EV_MAXT];   // term column in coef
};

__global__ __launch_bounds__(EV_THREADS) void k_pauli_apply(float2* __restrict__ out, const float2* __restrict__ in,
                                                            int n, uint64_t total, PTerms pt,
                                                            const float2* __restrict__ coef, int64_t coef_stride,
                                                            int accumulate) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = g >> n;
    const uint32_t j = (uint32_t)(g & ((1ull << n) - 1ull));
    const float2* r = in + (row << n);
    const float2* cf = coef + row * coef_stride;
    float2 acc = accumulate ? out[g] : make_float2(0.f, 0.f);
    for (int t = 0; t < pt.nt; ++t) {
      const uint32_t k = j ^ pt.x[t];
      const float2 v = r[k];
      const float2 c = cf[pt.col[t]];
      const float2 cv = cmul(c, v);
      const bool neg = __popc(k & pt.z[t]) & 1;
      acc.x += neg ? -cv.x : cv.x;
      acc.y += neg ? -cv.y : cv.y;
    }
    out[g] = acc;
  }
}


// ---- off-diagonal terms grouped by a local bit set -------------------------
// A "local set" L of m <= 12 bit positions (the 6 lowest always among them,
// for coalescing) splits each row into groups of 2^m amplitudes that differ
// only in L.  Every term whose flip mask x lies inside L maps a group onto
// itself, so one pass that stages each group in LDS serves all of them:
// per amplitude one LDS read per term instead of one pass over the state per
// term (expectations) or one uncached gather per term (their backward).
constexpr int PG_THREADS = 256;
constexpr int PG_MAXM = 12;
constexpr int PG_MAXT = 32;     // terms per local set (apply)
constexpr int PG_MAXT_EV = 16;  // terms per local set (expectations: two accumulators each)

struct PGroup {
  int m, nnl, nt, lo;        // local bits, non-local bits, terms, identity-mapped low bits
  int lpos[PG_MAXM];         // local bit i -> bit position (lpos[i] = i for i < lo)
  int nlpos[32];             // non-local positions, ascending
  uint32_t xl[PG_MAXT];      // flip mask in local numbering
  uint32_t zl[PG_MAXT];      // phase mask, local part in local numbering
  uint32_t zg[PG_MAXT];      // phase mask, non-local part (bit positions)
  int col[PG_MAXT];          // term column in coef / out
};

__device__ __forceinline__ uint32_t pg_scatter(const PGroup& g, uint32_t l) {
  uint32_t o = l & ((1u << g.lo) - 1u);
  for (int i = g.lo; i < g.m; ++i) o |= ((l >> i) & 1u) << g.lpos[i];
  return o;
}

__device__ __forceinline__ uint32_t pg_group_offset(const PGroup& g, uint64_t rest) {
  uint32_t o = 0;
  for (int i = 0; i < g.nnl; ++i) o |= (uint32_t)((rest >> i) & 1ull) << g.nlpos[i];
  return o;
}

// Stage one group (2^m amplitudes) into LDS; lane l of the group holds local
// indices l = tid + 256 r.  Returns nothing; buf is complete after the barrier.
__device__ __forceinline__ void pg_load(const PGroup& g, const float2* __restrict__ rin, float2* buf, uint32_t tsc,
                                        int tid, uint32_t size) {
  uint32_t r = 0;
  for (uint32_t l = tid; l < size; l += PG_THREADS, ++r) buf[l] = rin[tsc | pg_scatter(g, r << 8)];
}

// out (+)= sum_t coef[row, col_t] (-1)^{popcount((j^x_t) & z_t)} psi_{j ^ x_t}
// for the terms of one local set; one workgroup walks gpw consecutive groups.
// Term-outer order: each lane keeps its NR = 2^m / 256 outputs in registers,
// and a term's coefficient (with the group's non-local phase folded in) is
// wave-uniform, so a real or imaginary coefficient -- the usual case, i^{nY}
// times a real weight -- costs two FMA per amplitude.
//
// dotpart != null: the workgroup also adds Re conj(psi_l) out_l over its
// groups (all in one row) and writes the sum to dotpart[blockIdx.x]; on the
// last pass over out this is the workgroup's part of Re <psi|out>.
template <int NR>
__global__ __launch_bounds__(PG_THREADS) void k_pauli_group_apply(float2* __restrict__ out,
                                                                  const float2* __restrict__ in, int n,
                                                                  PGroup g, const float2* __restrict__ coef,
                                                                  int64_t coef_stride, int accumulate,
                                                                  int64_t n_groups, int64_t gpw,
                                                                  float* __restrict__ dotpart) {
  __shared__ float2 buf[NR * PG_THREADS];
  __shared__ float red[PG_THREADS / 64];
  float dsum = 0.f;
  const int tid = threadIdx.x;
  const uint32_t size = NR * PG_THREADS;
  // local index l = tid + 256 r: its bit positions are those of tid | those of r << 8
  const uint32_t tsc = pg_scatter(g, (uint32_t)tid);
  uint32_t rsc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) rsc[r] = pg_scatter(g, (uint32_t)r << 8);
  const int64_t g0 = (int64_t)blockIdx.x * gpw;
  const int64_t g1 = g0 + gpw < n_groups ? g0 + gpw : n_groups;
  const uint64_t rmask = (1ull << g.nnl) - 1ull;
  // this group's psi in registers (loaded one group ahead); it goes through
  // LDS for the flipped terms, diagonal terms read it from registers
  float2 a[NR];
  if (g0 < g1) {
    const float2* rin = in + (((uint64_t)g0 >> g.nnl) << n) + pg_group_offset(g, (uint64_t)g0 & rmask);
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = rin[tsc | rsc[r]];
  }
  for (int64_t grp = g0; grp < g1; ++grp) {
    const uint64_t row = (uint64_t)grp >> g.nnl;
    const uint32_t goff = pg_group_offset(g, (uint64_t)grp & rmask);
    float2* rout = out + (row << n) + goff;
    float2 acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = accumulate ? rout[tsc | rsc[r]] : make_float2(0.f, 0.f);
    __syncthreads();   // the previous group's LDS reads are done
#pragma unroll
    for (int r = 0; r < NR; ++r) buf[tid + r * PG_THREADS] = a[r];
    __syncthreads();
    float2 an[NR];
    if (grp + 1 < g1) {
      const int64_t gn = grp + 1;
      const float2* rin = in + (((uint64_t)gn >> g.nnl) << n) + pg_group_offset(g, (uint64_t)gn & rmask);
#pragma unroll
      for (int r = 0; r < NR; ++r) an[r] = rin[tsc | rsc[r]];
    }
    const float2* cf = coef + row * coef_stride;
    for (int t = 0; t < g.nt; ++t) {
      float2 c = cf[g.col[t]];
      if (__popc(goff & g.zg[t]) & 1) c = make_float2(-c.x, -c.y);
      // source amplitude k = l ^ x of local index l = tid | r << 8 splits
      // into disjoint fields: k = (tid ^ x_lo) | (r ^ x_hi) << 8, so its
      // phase sign is a per-thread part (folded into the coefficient once
      // per term) times a wave-uniform part per r
      const uint32_t xl = g.xl[t], zl = g.zl[t];
      const uint32_t tk = (uint32_t)tid ^ (xl & (PG_THREADS - 1));
      const uint32_t xh = xl >> 8, zh = zl >> 8;
      if (__popc(tk & zl & (PG_THREADS - 1)) & 1) c = make_float2(-c.x, -c.y);
      if (xl == 0u && c.y == 0.f) {   // diagonal, real coefficient: own amplitudes
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const float cr = (__popc((uint32_t)r & zh) & 1) ? -c.x : c.x;
          acc[r].x = fmaf(cr, a[r].x, acc[r].x);
          acc[r].y = fmaf(cr, a[r].y, acc[r].y);
        }
      } else if (c.y == 0.f) {        // real coefficient (wave-uniform branch)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const uint32_t rr = (uint32_t)r ^ xh;                   // wave-uniform
          const float2 v = buf[tk + rr * PG_THREADS];
          const float cr = (__popc(rr & zh) & 1) ? -c.x : c.x;
          acc[r].x = fmaf(cr, v.x, acc[r].x);
          acc[r].y = fmaf(cr, v.y, acc[r].y);
        }
      } else if (c.x == 0.f) {   // imaginary coefficient
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const uint32_t rr = (uint32_t)r ^ xh;
          const float2 v = buf[tk + rr * PG_THREADS];
          const float ci = (__popc(rr & zh) & 1) ? -c.y : c.y;
          acc[r].x = fmaf(-ci, v.y, acc[r].x);
          acc[r].y = fmaf(ci, v.x, acc[r].y);
        }
      } else {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const uint32_t rr = (uint32_t)r ^ xh;
          const float2 v = buf[tk + rr * PG_THREADS];
          const float2 cc = (__popc(rr & zh) & 1) ? make_float2(-c.x, -c.y) : c;
          acc[r] = cfma(cc, v, acc[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) rout[tsc | rsc[r]] = acc[r];
    if (dotpart) {
#pragma unroll
      for (int r = 0; r < NR; ++r) dsum = fmaf(a[r].x, acc[r].x, fmaf(a[r].y, acc[r].y, dsum));
    }
    if (grp + 1 < g1) {
#pragma unroll
      for (int r = 0; r < NR; ++r) a[r] = an[r];
    }
  }
  if (dotpart) {
    const float t = block_sum(dsum, red);
    if (tid == 0) dotpart[blockIdx.x] = t;
  }
}

// Partials of the terms of one local set over the gpw groups of one
// workgroup (all in one row).  With a = psi_j, b = psi_{j^x}:
// <P> = Re(i^{nY} S), S = sum_j (-1)^{popcount(j & z)} conj(b) a, so only
// Re S (nY even) or Im S (nY odd) is accumulated: two FMA per amplitude.
// Per-lane running sums live in LDS (one slot per term and lane), so the
// term loop stays a plain loop.  part[blockIdx.x * 2 nt + 2 t] = the partial.
template <int NR>
__global__ __launch_bounds__(PG_THREADS) void k_pauli_group_expval(const float2* __restrict__ in, int n,
                                                                   PGroup g, int64_t gpw,
                                                                   float* __restrict__ part) {
  __shared__ float2 buf[NR * PG_THREADS];
  __shared__ float lacc[PG_MAXT_EV][PG_THREADS];
  __shared__ float red[PG_THREADS / 64];
  const int tid = threadIdx.x;
  const uint32_t tsc = pg_scatter(g, (uint32_t)tid);
  uint32_t rsc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) rsc[r] = pg_scatter(g, (uint32_t)r << 8);
  const int64_t g0 = (int64_t)blockIdx.x * gpw;
  const uint64_t rmask = (1ull << g.nnl) - 1ull;
  for (int t = 0; t < g.nt; ++t) lacc[t][tid] = 0.f;
  for (int64_t grp = g0; grp < g0 + gpw; ++grp) {
    const uint32_t goff = pg_group_offset(g, (uint64_t)grp & rmask);
    const float2* rin = in + (((uint64_t)grp >> g.nnl) << n) + goff;
    __syncthreads();
    float2 a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      a[r] = rin[tsc | rsc[r]];
      buf[tid + r * PG_THREADS] = a[r];
    }
    __syncthreads();
    for (int t = 0; t < g.nt; ++t) {
      const uint32_t xl = g.xl[t], zl = g.zl[t];
      const bool odd = __popc(xl & zl) & 1;                  // nY odd: Im S
      // l = tid | r << 8 and l ^ x = (tid ^ x_lo) | (r ^ x_hi) << 8 (disjoint
      // fields): the sign of l splits into a per-thread part, applied once
      // per term, and a wave-uniform part per r
      const uint32_t tk = (uint32_t)tid ^ (xl & (PG_THREADS - 1));
      const uint32_t xh = xl >> 8, zh = zl >> 8;
      float s = 0.f;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const float2 b = buf[tk + ((uint32_t)r ^ xh) * PG_THREADS];
        const float d = odd ? fmaf(b.x, a[r].y, -b.y * a[r].x) : fmaf(b.x, a[r].x, b.y * a[r].y);
        s = fmaf((__popc((uint32_t)r & zh) & 1) ? -1.f : 1.f, d, s);
      }
      const bool neg = ((__popc(goff & g.zg[t]) ^ __popc((uint32_t)tid & zl & (PG_THREADS - 1))) & 1) != 0;
      lacc[t][tid] += neg ? -s : s;
    }
  }
  float* o = part + (uint64_t)blockIdx.x * 2 * g.nt;
  for (int t = 0; t < g.nt; ++t) {
    const float sr = block_sum(lacc[t][tid], red);
    if (tid == 0) o[2 * t] = sr;
  }
}

constexpr int PG_WG_PER_CU = 4;   // workgroups per CU of the grouped-term launches

// Workgroups per row of the grouped-term launches: a power of two dividing
// the 2^nnl groups of a row, about PG_WG_PER_CU per CU over the batch.
static int64_t pg_wg_per_row(int64_t bsz, int nnl) {
  const int64_t per_row = 1ll << nnl;
  int64_t wpr = ((int64_t)num_cu() * PG_WG_PER_CU) / bsz;
  if (wpr < 1) wpr = 1;
  if (wpr > per_row) wpr = per_row;
  while (per_row % wpr) --wpr;        // per_row is a power of two: wpr becomes one too
  return wpr;
}

// One pass over the terms of a local set; every workgroup stays in one row
// (wpr workgroups per row), so with dotpart each writes one row partial:
// dotpart[row * wpr + k].
template <int NR>
static int pg_apply_launch(hipStream_t s, float2* out, const float2* in, int n, const PGroup& g, const float2* coef,
                           int64_t coef_stride, int acc, int64_t bsz, float* dotpart) {
  const int64_t wpr = pg_wg_per_row(bsz, g.nnl);
  const int64_t gpw = (1ll << g.nnl) / wpr;
  const int64_t grid = bsz * wpr;
  if (grid > 0x7fffffffll) { set_error("tqsv_pauli_apply: batch too large"); return E_LIMIT; }
  k_pauli_group_apply<NR><<<(unsigned)grid, PG_THREADS, 0, s>>>(out, in, n, g, coef, coef_stride, acc,
                                                                bsz << g.nnl, gpw, dotpart);
  return check_launch("tqsv_pauli_apply/group");
}

template <int NR>
static int pg_expval_launch(hipStream_t s, const float2* in, int n, const PGroup& g, int64_t bsz, void* ws,
                            int64_t ws_bytes, int64_t& wpr) {
  const int64_t per_row = 1ll << g.nnl;
  wpr = pg_wg_per_row(bsz, g.nnl);
  const int64_t gpw = per_row / wpr;
  const int64_t grid = bsz * wpr;
  if (grid > 0x7fffffffll || ws_bytes < grid * 2 * g.nt * (int64_t)sizeof(float)) return 1;   // caller falls back
  k_pauli_group_expval<NR><<<(unsigned)grid, PG_THREADS, 0, s>>>(in, n, g, gpw, (float*)ws);
  return check_launch("tqsv_expval_pauli/group");
}

// Host side: split off-diagonal terms into local sets.  Each set keeps the
// lowest lo = min(6, n) bits plus up to m - lo higher bits that its terms flip;
// it is filled up to m bits with the lowest unused positions (n >= 8, so m >= 8:
// every lane of the 256-thread workgroup owns 2^m / 256 amplitudes).  Terms that flip
// more than m - lo bits above lo are left for the per-term kernels.
struct PGPlan {
  std::vector<PGroup> sets;
  std::vector<int> set_of_term;   // -1: not grouped
};

static PGPlan pg_plan(int n, int nterms, const uint64_t* xmask, const uint64_t* zmask, bool with_diag, int maxt) {
  PGPlan P;
  P.set_of_term.assign(nterms, -1);
  const int m = n < PG_MAXM ? n : PG_MAXM;
  const int lo = n < 6 ? n : 6;
  std::vector<uint32_t> hi_of(nterms, 0);
  for (int t = 0; t < nterms; ++t) hi_of[t] = (uint32_t)(xmask[t] >> lo) << lo;
  std::vector<int> todo;
  for (int t = 0; t < nterms; ++t)
    if ((xmask[t] || with_diag) && __builtin_popcount(hi_of[t]) <= m - lo) todo.push_back(t);
  while (!todo.empty()) {
    uint32_t H = 0;
    std::vector<int> members, rest;
    for (int t : todo) {
      const uint32_t h = H | hi_of[t];
      if (__builtin_popcount(h) <= m - lo && (int)members.size() < maxt) {
        H = h;
        members.push_back(t);
      } else {
        rest.push_back(t);
      }
    }
    PGroup g{};
    g.m = m;
    g.lo = lo;
    int nl = 0;
    for (int i = 0; i < lo; ++i) g.lpos[nl++] = i;
    uint32_t L = (1u << lo) - 1u;
    for (int b = lo; b < n && nl < m; ++b)
      if ((H >> b) & 1u) { g.lpos[nl++] = b; L |= 1u << b; }
    for (int b = lo; b < n && nl < m; ++b)
      if (!((L >> b) & 1u)) { g.lpos[nl++] = b; L |= 1u << b; }
    // local numbering must keep positions ascending past lo (scatter order is free,
    // but sorted keeps consecutive local indices close in memory)
    std::sort(g.lpos + lo, g.lpos + m);
    g.nnl = 0;
    for (int b = 0; b < n; ++b)
      if (!((L >> b) & 1u)) g.nlpos[g.nnl++] = b;
    auto to_local = [&](uint64_t mask) {
      uint32_t v = 0;
      for (int i = 0; i < m; ++i)
        if ((mask >> g.lpos[i]) & 1ull) v |= 1u << i;
      return v;
    };
    g.nt = (int)members.size();
    for (int q = 0; q < g.nt; ++q) {
      const int t = members[q];
      g.xl[q] = to_local(xmask[t]);
      g.zl[q] = to_local(zmask[t]);
      g.zg[q] = (uint32_t)(zmask[t] & ~(uint64_t)L);
      g.col[q] = t;
      P.set_of_term[t] = (int)P.sets.size();
    }
    P.sets.push_back(g);
    todo.swap(rest);
  }
  return P;
}

}  // namespace tqsv

using namespace tqsv;

static unsigned grid_cap(uint64_t work) {
  uint64_t b = (work + EV_THREADS - 1) / EV_THREADS;
  uint64_t cap = (uint64_t)num_cu() * 16;
  if (b > cap) b = cap;
  return b ? (unsigned)b : 1u;
}

// Bytes of float workspace the expectation entry points need.
// ---- allocation-free <Z_w> -------------------------------------------------
// tqsv_expval_z_all_noalloc runs the same two kernels as tqsv_expval_z_all,
// with the chunk partials in a fixed device array instead of a caller
// workspace, for callers that cannot allocate device memory.  Each GPU holds
// its own copy of the array until the process ends.  Every partial that
// k_ez_final reads is written by k_ez_part in the same call, so nothing needs
// clearing between calls.  Calls that overlap on one GPU share the array and
// race; callers serialise them (one stream, or their own ordering).
constexpr int64_t NOALLOC_SCRATCH_FLOATS = 1ll << 20;   // 4 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];
static std::mutex g_scratch_mu;
static float* g_scratch_dev[NOALLOC_MAX_DEVICES];      // device address of `scratch`, per device

// Device address of this GPU's `scratch`, looked up once per device (with
// that device current); host code uses only this address, never the symbol.
static int scratch_address(float** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("tqsv_expval_z_all_noalloc: no current HIP device (or device index >= %d)", NOALLOC_MAX_DEVICES);
    return E_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("tqsv_expval_z_all_noalloc: hipGetSymbolAddress failed on device %d: %s", dev, hipGetErrorString(e));
      return E_LAUNCH;
    }
    g_scratch_dev[dev] = (float*)p;
  }
  *out = g_scratch_dev[dev];
  return OK;
}

TQSV_API int64_t tqsv_expval_workspace_bytes(int64_t bsz, int n, int nterms) {
  if (bsz < 1 || n < 1 || n > 32) return -1;
  const int lg = ev_chunk_log2(bsz, n);
  const int64_t blocks = bsz << (n - lg);
  int64_t per = (int64_t)(n + 1);
  if (nterms * 2 > per) per = nterms * 2;
  if (EV_MAXT * 2 > per) per = EV_MAXT * 2;
  return blocks * per * (int64_t)sizeof(float) * 2;
}

// <Z_w> for all wires w: out[bsz, n] (wire order as in the reference: out[:, w]).
TQSV_API int tqsv_expval_z_all(hipStream_t s, const void* state, int64_t bsz, int n, float* out, void* workspace,
                               int64_t ws_bytes) {
  if (!state || !out || !workspace) { set_error("tqsv_expval_z_all: null pointer"); return E_ARG; }
  if (bsz < 1 || n < 1 || n > EV_MAXQ) { set_error("tqsv_expval_z_all: need 1 <= n <= %d", EV_MAXQ); return E_ARG; }
  if ((uintptr_t)state & 15) { set_error("tqsv_expval_z_all: state must be 16-byte aligned"); return E_ARG; }
  const int lg = ev_chunk_log2(bsz, n);
  const int64_t blocks = bsz << (n - lg);
  if (ws_bytes < blocks * (n + 1) * (int64_t)sizeof(float)) { set_error("tqsv_expval_z_all: workspace too small"); return E_ARG; }
  k_ez_part<<<(unsigned)blocks, EV_THREADS, 0, s>>>((const float4*)state, n, lg, (float*)workspace);
  int rc = check_launch("tqsv_expval_z_all/part");
  if (rc) return rc;
  const int64_t outs = bsz * n;
  k_ez_final<<<sum_grid(outs), 256, 0, s>>>((const float*)workspace, n, 1 << (n - lg), bsz, out);
  return check_launch("tqsv_expval_z_all/final");
}

// <Z_w> for every wire, as tqsv_expval_z_all, with no workspace argument and
// no allocation: the partials go to this GPU's fixed scratch array
// (NOALLOC_SCRATCH_FLOATS floats; larger shapes are rejected).
TQSV_API int tqsv_expval_z_all_noalloc(hipStream_t s, const void* state, int64_t bsz, int n, float* out) {
  if (!state || !out) { set_error("tqsv_expval_z_all_noalloc: null pointer"); return E_ARG; }
  if (bsz < 1 || n < 1 || n > EV_MAXQ) { set_error("tqsv_expval_z_all_noalloc: need 1 <= n <= %d", EV_MAXQ); return E_ARG; }
  if ((uintptr_t)state & 15) { set_error("tqsv_expval_z_all_noalloc: state must be 16-byte aligned"); return E_ARG; }
  const int lg = ev_chunk_log2(bsz, n);
  const int64_t blocks = bsz << (n - lg);
  const int64_t need = blocks * (n + 1);
  if (need > NOALLOC_SCRATCH_FLOATS) {
    set_error("tqsv_expval_z_all_noalloc: bsz %lld x %d qubits needs %lld floats of scratch, the maximum is %lld "
              "(use tqsv_expval_z_all with a workspace)", (long long)bsz, n, (long long)need,
              (long long)NOALLOC_SCRATCH_FLOATS);
    return E_LIMIT;
  }
  float* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  k_ez_part<<<(unsigned)blocks, EV_THREADS, 0, s>>>((const float4*)state, n, lg, ws);
  if ((rc = check_launch("tqsv_expval_z_all_noalloc/part"))) return rc;
  k_ez_final<<<sum_grid(bsz * n), 256, 0, s>>>(ws, n, 1 << (n - lg), bsz, out);
  return check_launch("tqsv_expval_z_all_noalloc/final");
}

// Re <psi|P_t|psi> for nterms Pauli strings given by bit masks (in-row bit
// positions; host arrays).  out[row * out_stride + t].
TQSV_API int tqsv_expval_pauli(hipStream_t s, const void* state, int64_t bsz, int n, int nterms,
                               const uint64_t* xmask, const uint64_t* zmask, float* out, int64_t out_stride,
                               void* workspace, int64_t ws_bytes) {
  if (!state || !out || !workspace || (nterms > 0 && (!xmask || !zmask))) {
    set_error("tqsv_expval_pauli: null pointer");
    return E_ARG;
  }
  if (bsz < 1 || n < 1 || n > 32 || nterms < 0) { set_error("tqsv_expval_pauli: need 1 <= n <= 32"); return E_ARG; }
  if ((uintptr_t)state & 15) { set_error("tqsv_expval_pauli: state must be 16-byte aligned"); return E_ARG; }
  const int lg = ev_chunk_log2(bsz, n);
  const int64_t nchunks = 1ll << (n - lg);
  const int64_t blocks = bsz * nchunks;
  // diagonal terms, batched EV_MAXT at a time (gather indices first)
  int idx[EV_MAXT];
  int cnt = 0;
  for (int t = 0; t <= nterms; ++t) {
    const bool flush = (t == nterms) || cnt == EV_MAXT;
    if (flush && cnt > 0) {
      if (ws_bytes < blocks * cnt * (int64_t)sizeof(float)) { set_error("tqsv_expval_pauli: workspace too small"); return E_ARG; }
      ZMasks zm;
      zm.nt = cnt;
      for (int q = 0; q < cnt; ++q) zm.z[q] = (uint32_t)zmask[idx[q]];
      k_ediag_part<<<(unsigned)blocks, EV_THREADS, 0, s>>>((const float4*)state, n, lg, zm, (float*)workspace);
      int rc = check_launch("tqsv_expval_pauli/diag");
      if (rc) return rc;
      // contiguous output indices? scatter one term at a time
      for (int q = 0; q < cnt; ++q) {
        // sum chunk partials of term q: treat part as [row, chunk, cnt]
        k_sum_chunks<<<sum_grid(bsz), 256, 0, s>>>(
            (const float*)workspace + q, (int)nchunks, cnt, 0, bsz, 1.f, 0.f, out, out_stride, idx[q]);
        if ((rc = check_launch("tqsv_expval_pauli/sum"))) return rc;
      }
      cnt = 0;
    }
    if (t == nterms) break;
    if (n < 64 && ((xmask[t] | zmask[t]) >> n)) { set_error("tqsv_expval_pauli: mask beyond n bits"); return E_ARG; }
    if (xmask[t] == 0) idx[cnt++] = t;
  }
  // off-diagonal terms: grouped by local set (one pass per set), the rest
  // (flip masks spanning too many high bits) one pass each
  std::vector<char> done(nterms, 0);
  if (n >= 8) {
    const PGPlan P = pg_plan(n, nterms, xmask, zmask, false, PG_MAXT_EV);
    for (const PGroup& g : P.sets) {
      int64_t wpr = 0;
      int rc;
      switch (g.m) {
        case 8: rc = pg_expval_launch<1>(s, (const float2*)state, n, g, bsz, workspace, ws_bytes, wpr); break;
        case 9: rc = pg_expval_launch<2>(s, (const float2*)state, n, g, bsz, workspace, ws_bytes, wpr); break;
        case 10: rc = pg_expval_launch<4>(s, (const float2*)state, n, g, bsz, workspace, ws_bytes, wpr); break;
        case 11: rc = pg_expval_launch<8>(s, (const float2*)state, n, g, bsz, workspace, ws_bytes, wpr); break;
        default: rc = pg_expval_launch<16>(s, (const float2*)state, n, g, bsz, workspace, ws_bytes, wpr); break;
      }
      if (rc == 1) continue;   // workspace too small: these terms take the per-term path
      if (rc) return rc;
      for (int q = 0; q < g.nt; ++q) {
        const int t = g.col[q];
        // partial = Re S (nY even) or Im S (nY odd); <P> = Re(i^{nY} S)
        const int ny = __builtin_popcount((uint32_t)(xmask[t] & zmask[t]));
        const float sc = (ny & 3) == 0 || (ny & 3) == 3 ? 1.f : -1.f;
        k_sum_chunks<<<sum_grid(bsz), 256, 0, s>>>((const float*)workspace + 2 * q, (int)wpr,
                                                                   2 * g.nt, 0, bsz, sc, 0.f, out, out_stride, t);
        if ((rc = check_launch("tqsv_expval_pauli/groupsum"))) return rc;
        done[t] = 1;
      }
    }
  }
  for (int t = 0; t < nterms; ++t) {
    if (xmask[t] == 0 || done[t]) continue;
    const uint32_t x = (uint32_t)xmask[t], z = (uint32_t)zmask[t];
    const int top = 31 - __builtin_clz(x);
    const int ny = __builtin_popcount(x & z);
    // i^{ny}: (re, im) scale
    float sre = 0.f, sim = 0.f;
    switch (ny & 3) { case 0: sre = 1.f; break; case 1: sim = 1.f; break; case 2: sre = -1.f; break; default: sim = -1.f; }
    const int lgi = n - 1;
    const int lgc = lg > lgi ? lgi : lg;
    const int64_t nch = 1ll << (lgi - lgc);
    const int64_t blk = bsz * nch;
    if (ws_bytes < blk * 2 * (int64_t)sizeof(float)) { set_error("tqsv_expval_pauli: workspace too small"); return E_ARG; }
    k_ex_part<<<(unsigned)blk, EV_THREADS, 0, s>>>((const float2*)state, n, lg, x, z, top, (float*)workspace);
    int rc = check_launch("tqsv_expval_pauli/x");
    if (rc) return rc;
    // Re(i^ny * S) = sre*Re(S) - sim*Im(S)
    k_sum_chunks<<<sum_grid(bsz), 256, 0, s>>>((const float*)workspace, (int)nch, 2, 1, bsz, sre,
                                                               sim, out, out_stride, t);
    if ((rc = check_launch("tqsv_expval_pauli/xsum"))) return rc;
  }
  return OK;
}

// out (+)= sum_t coef[row * coef_stride + t] P_t psi.  dotpart != null: the
// last pass also leaves the row partials of Re <psi|out> there and sets
// *dot_wpr to their count per row (dotpart[row * wpr + k]); 0 when the last
// pass was a per-term launch, which computes no partials.
static int pauli_apply_impl(hipStream_t s, void* out, const void* in, int64_t bsz, int n, int nterms,
                            const uint64_t* xmask, const uint64_t* zmask, const void* coef, int64_t coef_stride,
                            int accumulate, float* dotpart, int64_t* dot_wpr) {
  const uint64_t total = (uint64_t)bsz << n;
  int acc = accumulate ? 1 : 0;
  if (dot_wpr) *dot_wpr = 0;
  // terms grouped by local set: one LDS-staged pass per set
  std::vector<char> done(nterms, 0);
  std::vector<PGroup> sets;
  if (n >= 8) sets = pg_plan(n, nterms, xmask, zmask, true, PG_MAXT).sets;
  int n_rest = 0;
  for (const PGroup& g : sets)
    for (int q = 0; q < g.nt; ++q) done[g.col[q]] = 1;
  for (int t = 0; t < nterms; ++t) n_rest += !done[t];
  for (size_t i = 0; i < sets.size(); ++i) {
    const PGroup& g = sets[i];
    float2* o = (float2*)out;
    const float2* x = (const float2*)in;
    const float2* c = (const float2*)coef;
    float* dp = (dotpart && n_rest == 0 && i + 1 == sets.size()) ? dotpart : nullptr;
    int rc;
    switch (g.m) {
      case 8: rc = pg_apply_launch<1>(s, o, x, n, g, c, coef_stride, acc, bsz, dp); break;
      case 9: rc = pg_apply_launch<2>(s, o, x, n, g, c, coef_stride, acc, bsz, dp); break;
      case 10: rc = pg_apply_launch<4>(s, o, x, n, g, c, coef_stride, acc, bsz, dp); break;
      case 11: rc = pg_apply_launch<8>(s, o, x, n, g, c, coef_stride, acc, bsz, dp); break;
      default: rc = pg_apply_launch<16>(s, o, x, n, g, c, coef_stride, acc, bsz, dp); break;
    }
    if (rc) return rc;
    if (dp) *dot_wpr = pg_wg_per_row(bsz, g.nnl);
    acc = 1;
  }
  // the rest: gathered per amplitude, EV_MAXT terms per launch
  std::vector<int> rest;
  for (int t = 0; t < nterms; ++t)
    if (!done[t]) rest.push_back(t);
  for (size_t t0 = 0; t0 < rest.size(); t0 += EV_MAXT) {
    PTerms pt;
    pt.nt = (int)(rest.size() - t0 < (size_t)EV_MAXT ? rest.size() - t0 : (size_t)EV_MAXT);
    for (int q = 0; q < pt.nt; ++q) {
      pt.x[q] = (uint32_t)xmask[rest[t0 + q]];
      pt.z[q] = (uint32_t)zmask[rest[t0 + q]];
      pt.col[q] = rest[t0 + q];
    }
    k_pauli_apply<<<grid_cap(total), EV_THREADS, 0, s>>>((float2*)out, (const float2*)in, n, total, pt,
                                                         (const float2*)coef, coef_stride, acc);
    int rc = check_launch("tqsv_pauli_apply");
    if (rc) return rc;
    acc = 1;
  }
  return OK;
}

static int pauli_args_ok(const char* who, const void* out, const void* in, int64_t bsz, int n, int nterms,
                         const uint64_t* xmask, const uint64_t* zmask, const void* coef) {
  if (!out || !in || !coef || (nterms > 0 && (!xmask || !zmask))) { set_error("%s: null pointer", who); return E_ARG; }
  if (out == in) { set_error("%s: out must not alias in", who); return E_ARG; }
  if (bsz < 1 || n < 1 || n > 32 || nterms < 1) { set_error("%s: bad sizes", who); return E_ARG; }
  for (int t = 0; t < nterms; ++t)
    if ((xmask[t] | zmask[t]) >> n) { set_error("%s: mask beyond n bits", who); return E_ARG; }
  return OK;
}

// out (+)= sum_t coef[row, t] P_t psi   (coef: device complex64 [bsz, nterms],
// already including i^{nY}); used for d/d psi of Pauli expectations.
TQSV_API int tqsv_pauli_apply(hipStream_t s, void* out, const void* in, int64_t bsz, int n, int nterms,
                              const uint64_t* xmask, const uint64_t* zmask, const void* coef, int accumulate) {
  int rc = pauli_args_ok("tqsv_pauli_apply", out, in, bsz, n, nterms, xmask, zmask, coef);
  if (rc) return rc;
  return pauli_apply_impl(s, out, in, bsz, n, nterms, xmask, zmask, coef, nterms, accumulate, nullptr, nullptr);
}

// Floats of workspace tqsv_pauli_sum_grad needs (row partials of the dot).
TQSV_API int64_t tqsv_pauli_sum_grad_workspace_floats(int64_t bsz, int n) {
  if (bsz < 1 || n < 1 || n > 32) return -1;
  const int64_t fused = n >= 8 ? bsz * pg_wg_per_row(bsz, n - (n < PG_MAXM ? n : PG_MAXM)) : 0;
  const int64_t plain = bsz << (n - ev_chunk_log2(bsz, n));
  return fused > plain ? fused : plain;
}

// A Pauli sum and its state gradient in one go: out = sum_t coef[t] P_t psi
// (coef: device complex64 [nterms], the same for every row, including
// i^{nY}) and dot[row] = Re <psi_row|out_row> (device float [bsz]).  With
// coef = 2 c_t i^{nY_t}, dot / 2 is the row's sum_t c_t <P_t> and
// g_row * out is d(sum_row g_row value_row) / d conj(psi) -- what a
// differentiated Pauli-sum loss needs, for the cost of the backward's passes
// alone.  workspace: tqsv_pauli_sum_grad_workspace_floats floats.
TQSV_API int tqsv_pauli_sum_grad(hipStream_t s, void* out, const void* in, int64_t bsz, int n, int nterms,
                                 const uint64_t* xmask, const uint64_t* zmask, const void* coef, float* dot,
                                 float* workspace, int64_t workspace_floats) {
  int rc = pauli_args_ok("tqsv_pauli_sum_grad", out, in, bsz, n, nterms, xmask, zmask, coef);
  if (rc) return rc;
  if (!dot || !workspace) { set_error("tqsv_pauli_sum_grad: null pointer"); return E_ARG; }
  if ((uintptr_t)in & 15) { set_error("tqsv_pauli_sum_grad: state must be 16-byte aligned"); return E_ARG; }
  const int64_t need = tqsv_pauli_sum_grad_workspace_floats(bsz, n);
  if (workspace_floats < need) {
    set_error("tqsv_pauli_sum_grad: workspace of %lld floats, need %lld", (long long)workspace_floats,
              (long long)need);
    return E_ARG;
  }
  int64_t wpr = 0;
  rc = pauli_apply_impl(s, out, in, bsz, n, nterms, xmask, zmask, coef, 0, 0, workspace, &wpr);
  if (rc) return rc;
  if (wpr == 0) {      // the last pass computed no partials: one read of psi and out
    const int lg = ev_chunk_log2(bsz, n);
    k_rowdot_part<<<(unsigned)(bsz << (n - lg)), EV_THREADS, 0, s>>>((const float2*)in, (const float2*)out, n, lg,
                                                                      workspace);
    if ((rc = check_launch("tqsv_pauli_sum_grad/dot"))) return rc;
    wpr = 1ll << (n - lg);
  }
  k_sum_chunks<<<sum_grid(bsz), 256, 0, s>>>(workspace, (int)wpr, 1, 0, bsz, 1.f, 0.f, dot, 1, 0);
  return check_launch("tqsv_pauli_sum_grad/sum");
}
