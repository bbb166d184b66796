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
__global__ void k_init_basis(float4* __restrict__ st, uint64_t n_pairs, int n, uint64_t index) {
  const uint64_t row_mask = (1ull << n) - 1ull;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_pairs;
       i += (uint64_t)gridDim.x * blockDim.x) {
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
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_1q_t0(float4* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                              uint64_t cval, const float2* __restrict__ mat, int64_t bstride,
                                              int lg_items_row) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < n_items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i0 = insert_zeros(it, ins) | cval;
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, it >> lg_items_row, m);
    float4 v = st[i0 >> 1];
    const float2 a0 = make_float2(v.x, v.y), a1 = make_float2(v.z, v.w);
    const float2 b0 = cfma(m[1], a1, cmul(m[0], a0));
    const float2 b1 = cfma(m[3], a1, cmul(m[2], a0));
    st[i0 >> 1] = make_float4(b0.x, b0.y, b1.x, b1.y);
  }
}

// target bit > 0 and bit 0 free: two consecutive items share 16-byte accesses.
// Each thread keeps two such item pairs (grid-stride unroll) in flight; NT
// marks the accesses non-temporal for states larger than the Infinity Cache.
typedef float v4f __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ float4 ld4(const float4* p) {
  if constexpr (NT) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void st4(float4* p, float4 v) {
  if constexpr (NT) {
    v4f t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(p));
  } else {
    *p = v;
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

template <bool BATCHED, bool NT>
__global__ __launch_bounds__(256) void k_1q_v2(float4* __restrict__ st, uint64_t n_pairs, InsertSpec ins,
                                              uint64_t cval, uint64_t tmask, const float2* __restrict__ mat,
                                              int64_t bstride, int lg_items_row) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n_pairs; q += 2 * stride) {
    const uint64_t qb = q + stride;
    const bool two = qb < n_pairs;
    const uint64_t i0 = insert_zeros(2 * q, ins) | cval;
    const uint64_t i1 = i0 | tmask;
    const uint64_t j0 = two ? (insert_zeros(2 * qb, ins) | cval) : i0;
    const uint64_t j1 = j0 | tmask;
    const float4 u = ld4<NT>(st + (i0 >> 1));
    const float4 w = ld4<NT>(st + (i1 >> 1));
    float4 u2 = u, w2 = w;
    if (two) {
      u2 = ld4<NT>(st + (j0 >> 1));
      w2 = ld4<NT>(st + (j1 >> 1));
    }
    float2 m[4];
    load_m2<BATCHED>(mat, bstride, (2 * q) >> lg_items_row, m);
    float4 uo, wo;
    rot2(m, u, w, uo, wo);
    st4<NT>(st + (i0 >> 1), uo);
    st4<NT>(st + (i1 >> 1), wo);
    if (two) {
      if (BATCHED) load_m2<BATCHED>(mat, bstride, (2 * qb) >> lg_items_row, m);
      rot2(m, u2, w2, uo, wo);
      st4<NT>(st + (j0 >> 1), uo);
      st4<NT>(st + (j1 >> 1), wo);
    }
  }
}

// generic scalar path (bit 0 is a control bit)
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_1q_gen(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                               uint64_t cval, uint64_t tmask, const float2* __restrict__ mat,
                                               int64_t bstride, int lg_items_row) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < n_items;
       it += (uint64_t)gridDim.x * blockDim.x) {
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
template <bool BATCHED>
__device__ __forceinline__ void mv4(const float2* __restrict__ p, const float2 (&a)[4], float2 (&b)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float2 acc = cmul(p[4 * r], a[0]);
#pragma unroll
    for (int c = 1; c < 4; ++c) acc = cfma(p[4 * r + c], a[c], acc);
    b[r] = acc;
  }
}

template <bool BATCHED>
__global__ __launch_bounds__(256) void k_2q_v2(float4* __restrict__ st, uint64_t n_pairs, InsertSpec ins,
                                              uint64_t cval, uint64_t mhi, uint64_t mlo,
                                              const float2* __restrict__ mat, int64_t bstride, int lg_items_row) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n_pairs;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t it = 2 * q;
    const uint64_t base = insert_zeros(it, ins) | cval;
    const uint64_t idx[4] = {base, base | mlo, base | mhi, base | mhi | mlo};
    const float2* p = BATCHED ? mat + (it >> lg_items_row) * bstride : mat;
    float4 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = st[idx[r] >> 1];
    float2 a[4], b[4], an[4], bn[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = make_float2(v[r].x, v[r].y);
      an[r] = make_float2(v[r].z, v[r].w);
    }
    mv4<BATCHED>(p, a, b);
    mv4<BATCHED>(p, an, bn);
#pragma unroll
    for (int r = 0; r < 4; ++r) st[idx[r] >> 1] = make_float4(b[r].x, b[r].y, bn[r].x, bn[r].y);
  }
}

template <bool BATCHED>
__global__ __launch_bounds__(256) void k_2q_gen(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                               uint64_t cval, uint64_t mhi, uint64_t mlo,
                                               const float2* __restrict__ mat, int64_t bstride, int lg_items_row) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < n_items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t base = insert_zeros(it, ins) | cval;
    const uint64_t idx[4] = {base, base | mlo, base | mhi, base | mhi | mlo};
    const float2* p = BATCHED ? mat + (it >> lg_items_row) * bstride : mat;
    float2 a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = st[idx[r]];
    mv4<BATCHED>(p, a, b);
#pragma unroll
    for (int r = 0; r < 4; ++r) st[idx[r]] = b[r];
  }
}

// ---------------------------------------------------------------------------
// diagonal gate: amplitudes whose control bits match get phase[parity(j & pmask)]
// ---------------------------------------------------------------------------
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_diag(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                             uint64_t cval, uint64_t pmask, const float2* __restrict__ ph,
                                             int64_t bstride, int lg_items_row) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < n_items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = insert_zeros(it, ins) | cval;
    const float2* p = BATCHED ? ph + (it >> lg_items_row) * bstride : ph;
    const int par = __popcll(i & pmask) & 1;
    const float2 f = par ? p[1] : p[0];
    st[i] = cmul(f, st[i]);
  }
}

// no controls: two adjacent amplitudes per thread, 16-byte accesses
template <bool BATCHED>
__global__ __launch_bounds__(256) void k_diag_v2(float4* __restrict__ st, uint64_t n_pairs, uint64_t pmask,
                                                const float2* __restrict__ ph, int64_t bstride, int n) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n_pairs;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = 2 * q;
    const float2* p = BATCHED ? ph + (i >> n) * bstride : ph;
    const float2 f0 = p[0], f1 = p[1];
    float4 v = st[q];
    const int par0 = __popcll(i & pmask) & 1;
    const int par1 = __popcll((i | 1ull) & pmask) & 1;
    const float2 a = cmul(par0 ? f1 : f0, make_float2(v.x, v.y));
    const float2 b = cmul(par1 ? f1 : f0, make_float2(v.z, v.w));
    st[q] = make_float4(a.x, a.y, b.x, b.y);
  }
}

// ---------------------------------------------------------------------------
// swap of bits a, b (optionally controlled): exchanges |..0..1..> <-> |..1..0..>
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_swap(float2* __restrict__ st, uint64_t n_items, InsertSpec ins,
                                             uint64_t cval, uint64_t ma, uint64_t mb) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < n_items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t base = insert_zeros(it, ins) | cval;
    const uint64_t i = base | mb, j = base | ma;
    const float2 x = st[i], y = st[j];
    st[i] = y;
    st[j] = x;
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
    k_init_basis<<<grid_for(pairs, 256), 256, 0, s>>>((float4*)state, pairs, n, (uint64_t)index);
  }
  return check_launch("tqsv_init_basis");
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
  InsertSpec ins;
  if (!build_insert(ins, cmask | tmask, n)) { set_error("too many control bits (max 11)"); return E_LIMIT; }
  const int lg_items_row = n - 1 - __builtin_popcountll(cmask);
  const uint64_t n_items = (uint64_t)bsz << lg_items_row;
  const bool b = mat_bstride != 0;
  const float2* m = (const float2*)mat;
  if (tbit == 0) {
    if (b) k_1q_t0<true><<<grid_for(n_items, 256), 256, 0, s>>>((float4*)state, n_items, ins, cval, m, mat_bstride, lg_items_row);
    else k_1q_t0<false><<<grid_for(n_items, 256), 256, 0, s>>>((float4*)state, n_items, ins, cval, m, 0, lg_items_row);
  } else if (!(cmask & 1ull)) {
    const uint64_t n_pairs = n_items / 2;
    // states beyond the 256 MB Infinity Cache stream with non-temporal hints
    const bool nt = ((uint64_t)bsz << n) * sizeof(float2) > (256ull << 20);
    const unsigned g = grid_for((n_pairs + 1) / 2, 256);
    if (b) {
      if (nt) k_1q_v2<true, true><<<g, 256, 0, s>>>((float4*)state, n_pairs, ins, cval, tmask, m, mat_bstride, lg_items_row);
      else k_1q_v2<true, false><<<g, 256, 0, s>>>((float4*)state, n_pairs, ins, cval, tmask, m, mat_bstride, lg_items_row);
    } else {
      if (nt) k_1q_v2<false, true><<<g, 256, 0, s>>>((float4*)state, n_pairs, ins, cval, tmask, m, 0, lg_items_row);
      else k_1q_v2<false, false><<<g, 256, 0, s>>>((float4*)state, n_pairs, ins, cval, tmask, m, 0, lg_items_row);
    }
  } else {
    if (b) k_1q_gen<true><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, tmask, m, mat_bstride, lg_items_row);
    else k_1q_gen<false><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, tmask, m, 0, lg_items_row);
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
  InsertSpec ins;
  if (!build_insert(ins, cmask | mhi | mlo, n)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
  const int lg_items_row = n - 2 - __builtin_popcountll(cmask);
  const uint64_t n_items = (uint64_t)bsz << lg_items_row;
  const bool b = mat_bstride != 0;
  const float2* m = (const float2*)mat;
  if (!((cmask | mhi | mlo) & 1ull)) {
    const uint64_t n_pairs = n_items / 2;
    if (b) k_2q_v2<true><<<grid_for(n_pairs, 256), 256, 0, s>>>((float4*)state, n_pairs, ins, cval, mhi, mlo, m, mat_bstride, lg_items_row);
    else k_2q_v2<false><<<grid_for(n_pairs, 256), 256, 0, s>>>((float4*)state, n_pairs, ins, cval, mhi, mlo, m, 0, lg_items_row);
  } else {
    if (b) k_2q_gen<true><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, mhi, mlo, m, mat_bstride, lg_items_row);
    else k_2q_gen<false><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, mhi, mlo, m, 0, lg_items_row);
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
  const bool b = ph_bstride != 0;
  const float2* p = (const float2*)phases;
  if (cmask == 0 && n >= 1) {
    const uint64_t n_pairs = ((uint64_t)bsz << n) / 2;
    if (b) k_diag_v2<true><<<grid_for(n_pairs, 256), 256, 0, s>>>((float4*)state, n_pairs, pmask, p, ph_bstride, n);
    else k_diag_v2<false><<<grid_for(n_pairs, 256), 256, 0, s>>>((float4*)state, n_pairs, pmask, p, 0, n);
    return check_launch("tqsv_apply_diag");
  }
  InsertSpec ins;
  if (!build_insert(ins, cmask, n)) { set_error("too many control bits (max 12)"); return E_LIMIT; }
  const int lg_items_row = n - __builtin_popcountll(cmask);
  const uint64_t n_items = (uint64_t)bsz << lg_items_row;
  if (b) k_diag<true><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, pmask, p, ph_bstride, lg_items_row);
  else k_diag<false><<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, pmask, p, 0, lg_items_row);
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
  InsertSpec ins;
  if (!build_insert(ins, cmask | ma | mb, n)) { set_error("too many control bits (max 10)"); return E_LIMIT; }
  const uint64_t n_items = (uint64_t)bsz << (n - 2 - __builtin_popcountll(cmask));
  k_swap<<<grid_for(n_items, 256), 256, 0, s>>>((float2*)state, n_items, ins, cval, ma, mb);
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
  const int d = 1 << k;
  const int lg_groups_row = n - k - __builtin_popcountll(cmask);
  const uint64_t n_groups = (uint64_t)bsz << lg_groups_row;
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
