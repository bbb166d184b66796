// Per-gate adjoint-differentiation step kernels (the unfused gradient path).
//
// For gate k (walking the circuit backwards) one pass reads psi_k and
// lam_k and writes
//     psi_{k-1} = U^dag psi_k,   lam_{k-1} = U^dag lam_k,
// while accumulating, for every gate parameter p,
//     g_p += Re < lam_k | (dU/dtheta_p) psi_{k-1} >
// (lam = torch's gradient of the loss w.r.t. the state, i.e. dL/dRe + i dL/dIm).
// Gates without trainable parameters need no fused step: their U^dag is two
// ordinary per-gate launches (tqsv_gates.hip) on psi and lam.
//
// Gradients land in grad[row * grad_stride + gslot + p] (one row per batch
// row; the caller sums rows for parameters shared across the batch).  Each
// workgroup covers part of one row (grid.y = row), reduces its partials with
// a wave shuffle tree + LDS and writes them to its own slots of a caller-
// provided partial buffer; k_rowpart_reduce then adds the partials of each
// row in a fixed order (no float atomics: gradients are bitwise
// reproducible from run to run).
#include "tqsv_common.h"

namespace tqsv {
int num_cu();
}

namespace tqsv {

constexpr int AJ_THREADS = 256;

// The workgroup's P partial sums, written (not added) to its own 4 slots of
// the partial buffer: part[(row * gridDim.x + blockIdx.x) * 4 + q].
// k_rowpart_reduce adds them up in a fixed order afterwards, so gradients do
// not depend on the order in which workgroups finish.
__device__ __forceinline__ void block_partial(float (&g)[4], int P, float* __restrict__ dst) {
  __shared__ float red[4][AJ_THREADS / 64];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < P) {
      const float s = wave_sum(g[q], 64);
      if ((threadIdx.x & 63) == 0) red[q][w] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)P) {
    float s = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[threadIdx.x][i];
    dst[threadIdx.x] = s;
  }
}

// grad[row * gstride + gslot + q] += sum over x of part[(row * nx + x) * 4 + q],
// one workgroup per batch row: each thread sums a strided subset in order,
// then a fixed tree in LDS -- the same order on every run.
__global__ __launch_bounds__(AJ_THREADS) void k_rowpart_reduce(const float* __restrict__ part, int64_t nx, int P,
                                                               float* __restrict__ grad, int64_t gstride,
                                                               int gslot) {
  __shared__ float red[AJ_THREADS];
  const uint64_t row = blockIdx.x;
  for (int q = 0; q < P; ++q) {
    float s = 0.f;
    for (int64_t x = threadIdx.x; x < nx; x += AJ_THREADS) s += part[((int64_t)row * nx + x) * 4 + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = AJ_THREADS / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) grad[row * gstride + gslot + q] += red[0];
    __syncthreads();
  }
}

template <bool BATCHED>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_1q(float2* __restrict__ psi, float2* __restrict__ lam, int n,
                                                      uint64_t items_row, InsertSpec ins, uint64_t cval,
                                                      uint64_t tmask, const float2* __restrict__ mat,
                                                      int64_t bstride, int P, float* __restrict__ part) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? mat + row * bstride : mat;
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << n;
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items_row;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i0 = rbase + (insert_zeros(it, ins) | cval);
    const uint64_t i1 = i0 | tmask;
    const float2 x0 = psi[i0], x1 = psi[i1];
    const float2 l0 = lam[i0], l1 = lam[i1];
    const float2 y0 = cfmac(u2, x1, cmulc(u0, x0));
    const float2 y1 = cfmac(u3, x1, cmulc(u1, x0));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < P) {
        const float2* D = M + 4 * (q + 1);
        const float2 v0 = cfma(D[1], y1, cmul(D[0], y0));
        const float2 v1 = cfma(D[3], y1, cmul(D[2], y0));
        g[q] += redot(l0, v0) + redot(l1, v1);
      }
    }
    psi[i0] = y0;
    psi[i1] = y1;
    lam[i0] = cfmac(u2, l1, cmulc(u0, l0));
    lam[i1] = cfmac(u3, l1, cmulc(u1, l0));
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

template <bool BATCHED>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_2q(float2* __restrict__ psi, float2* __restrict__ lam, int n,
                                                      uint64_t items_row, InsertSpec ins, uint64_t cval,
                                                      uint64_t mhi, uint64_t mlo, const float2* __restrict__ mat,
                                                      int64_t bstride, int P, float* __restrict__ part) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? mat + row * bstride : mat;
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << n;
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items_row;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = rbase + (insert_zeros(it, ins) | cval);
    const uint64_t idx[4] = {b, b | mlo, b | mhi, b | mhi | mlo};
    float2 x[4], l[4], y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) { x[r] = psi[idx[r]]; l[r] = lam[idx[r]]; }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = cmulc(M[r], x[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cfmac(M[4 * c + r], x[c], acc);
      y[r] = acc;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < P) {
        const float2* D = M + 16 * (q + 1);
        float t = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float2 v = cmul(D[4 * r], y[0]);
#pragma unroll
          for (int c = 1; c < 4; ++c) v = cfma(D[4 * r + c], y[c], v);
          t += redot(l[r], v);
        }
        g[q] += t;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = cmulc(M[r], l[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cfmac(M[4 * c + r], l[c], acc);
      psi[idx[r]] = y[r];
      lam[idx[r]] = acc;
    }
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

template <bool BATCHED>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_diag(float2* __restrict__ psi, float2* __restrict__ lam, int n,
                                                        uint64_t items_row, InsertSpec ins, uint64_t cval,
                                                        uint64_t pmask, const float2* __restrict__ ph,
                                                        int64_t bstride, int P, float* __restrict__ part) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? ph + row * bstride : ph;
  const float2 p0 = M[0], p1 = M[1];
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << n;
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items_row;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = insert_zeros(it, ins) | cval;
    const uint64_t i = rbase + j;
    const int par = __popcll(j & pmask) & 1;
    const float2 f = par ? p1 : p0;
    const float2 y = cmulc(f, psi[i]);
    const float2 l = lam[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < P) {
        const float2 df = par ? M[2 * (q + 1) + 1] : M[2 * (q + 1)];
        g[q] += redot(l, cmul(df, y));
      }
    }
    psi[i] = y;
    lam[i] = cmulc(f, l);
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

// Bit 0 free of targets and controls: items are float4 (two adjacent
// amplitudes per access), and each workgroup owns a contiguous tile of
// 256 * reps items of one row (grid = (tiles per row, bsz)): the streaming
// shape of the forward per-gate kernels, with one gradient flush per tile.
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_1q4(float4* __restrict__ psi, float4* __restrict__ lam, int n,
                                                       uint64_t items_row, InsertSpec ins4, uint64_t cval4,
                                                       uint64_t S, const float2* __restrict__ mat, int64_t bstride,
                                                       int P, float* __restrict__ part, int reps) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? mat + row * bstride : mat;
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << (n - 1);
  uint64_t q = (uint64_t)blockIdx.x * AJ_THREADS * (uint64_t)reps + threadIdx.x;
  for (int r = 0; r < reps; ++r, q += AJ_THREADS) {
    if (q >= items_row) break;
    const uint64_t f0 = rbase + (insert_zeros(q, ins4) | cval4), f1 = f0 | S;
    const float4 xa = ld4<NT>(psi + f0), xb = ld4<NT>(psi + f1);
    const float4 la = ld4<NT>(lam + f0), lb = ld4<NT>(lam + f1);
    const float2 x0[2] = {make_float2(xa.x, xa.y), make_float2(xa.z, xa.w)};
    const float2 x1[2] = {make_float2(xb.x, xb.y), make_float2(xb.z, xb.w)};
    const float2 l0[2] = {make_float2(la.x, la.y), make_float2(la.z, la.w)};
    const float2 l1[2] = {make_float2(lb.x, lb.y), make_float2(lb.z, lb.w)};
    float2 y0[2], y1[2], m0[2], m1[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      y0[h] = cfmac(u2, x1[h], cmulc(u0, x0[h]));
      y1[h] = cfmac(u3, x1[h], cmulc(u1, x0[h]));
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < P) {
          const float2* D = M + 4 * (k + 1);
          const float2 v0 = cfma(D[1], y1[h], cmul(D[0], y0[h]));
          const float2 v1 = cfma(D[3], y1[h], cmul(D[2], y0[h]));
          g[k] += redot(l0[h], v0) + redot(l1[h], v1);
        }
      }
      m0[h] = cfmac(u2, l1[h], cmulc(u0, l0[h]));
      m1[h] = cfmac(u3, l1[h], cmulc(u1, l0[h]));
    }
    st4<NT>(psi + f0, make_float4(y0[0].x, y0[0].y, y0[1].x, y0[1].y));
    st4<NT>(psi + f1, make_float4(y1[0].x, y1[0].y, y1[1].x, y1[1].y));
    st4<NT>(lam + f0, make_float4(m0[0].x, m0[0].y, m0[1].x, m0[1].y));
    st4<NT>(lam + f1, make_float4(m1[0].x, m1[0].y, m1[1].x, m1[1].y));
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

template <bool BATCHED, bool NT>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_diag4(float4* __restrict__ psi, float4* __restrict__ lam, int n,
                                                         uint64_t items_row, InsertSpec ins4, uint64_t cval4,
                                                         uint64_t pmask, const float2* __restrict__ ph,
                                                         int64_t bstride, int P, float* __restrict__ part, int reps) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? ph + row * bstride : ph;
  const float2 p0 = M[0], p1 = M[1];
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << (n - 1);
  uint64_t q = (uint64_t)blockIdx.x * AJ_THREADS * (uint64_t)reps + threadIdx.x;
  for (int r = 0; r < reps; ++r, q += AJ_THREADS) {
    if (q >= items_row) break;
    const uint64_t j4 = insert_zeros(q, ins4) | cval4;  // in-row float4 index
    const float4 xv = ld4<NT>(psi + rbase + j4), lv = ld4<NT>(lam + rbase + j4);
    const float2 x[2] = {make_float2(xv.x, xv.y), make_float2(xv.z, xv.w)};
    const float2 l[2] = {make_float2(lv.x, lv.y), make_float2(lv.z, lv.w)};
    float2 y[2], m[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int par = __popcll((2 * j4 + h) & pmask) & 1;
      const float2 f = par ? p1 : p0;
      y[h] = cmulc(f, x[h]);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < P) {
          const float2 df = par ? M[2 * (k + 1) + 1] : M[2 * (k + 1)];
          g[k] += redot(l[h], cmul(df, y[h]));
        }
      }
      m[h] = cmulc(f, l[h]);
    }
    st4<NT>(psi + rbase + j4, make_float4(y[0].x, y[0].y, y[1].x, y[1].y));
    st4<NT>(lam + rbase + j4, make_float4(m[0].x, m[0].y, m[1].x, m[1].y));
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

// 2-qubit gate, bit 0 free: items are float4 quadruples (positions in float4 bits)
template <bool BATCHED, bool NT>
__global__ __launch_bounds__(AJ_THREADS) void k_adj_2q4(float4* __restrict__ psi, float4* __restrict__ lam, int n,
                                                       uint64_t items_row, InsertSpec ins4, uint64_t cval4,
                                                       uint64_t mhi4, uint64_t mlo4, const float2* __restrict__ mat,
                                                       int64_t bstride, int P, float* __restrict__ part, int reps) {
  const uint64_t row = blockIdx.y;
  const float2* M = BATCHED ? mat + row * bstride : mat;
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  const uint64_t rbase = row << (n - 1);
  uint64_t q = (uint64_t)blockIdx.x * AJ_THREADS * (uint64_t)reps + threadIdx.x;
  for (int rr = 0; rr < reps; ++rr, q += AJ_THREADS) {
    if (q >= items_row) break;
    const uint64_t b = rbase + (insert_zeros(q, ins4) | cval4);
    const uint64_t idx[4] = {b, b | mlo4, b | mhi4, b | mhi4 | mlo4};
    float4 xv[4], lv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      xv[r] = ld4<NT>(psi + idx[r]);
      lv[r] = ld4<NT>(lam + idx[r]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {   // the two amplitudes of each float4
      float2 x[4], l[4], y[4], m[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        x[r] = h ? make_float2(xv[r].z, xv[r].w) : make_float2(xv[r].x, xv[r].y);
        l[r] = h ? make_float2(lv[r].z, lv[r].w) : make_float2(lv[r].x, lv[r].y);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float2 acc = cmulc(M[r], x[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) acc = cfmac(M[4 * c + r], x[c], acc);
        y[r] = acc;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < P) {
          const float2* D = M + 16 * (k + 1);
          float t = 0.f;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float2 v = cmul(D[4 * r], y[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) v = cfma(D[4 * r + c], y[c], v);
            t += redot(l[r], v);
          }
          g[k] += t;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float2 acc = cmulc(M[r], l[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) acc = cfmac(M[4 * c + r], l[c], acc);
        m[r] = acc;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (h) {
          xv[r].z = y[r].x; xv[r].w = y[r].y;
          lv[r].z = m[r].x; lv[r].w = m[r].y;
        } else {
          xv[r].x = y[r].x; xv[r].y = y[r].y;
          lv[r].x = m[r].x; lv[r].y = m[r].y;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      st4<NT>(psi + idx[r], xv[r]);
      st4<NT>(lam + idx[r], lv[r]);
    }
  }
  block_partial(g, P, part + (row * gridDim.x + blockIdx.x) * 4);
}

// tiles per row for the float4 adjoint kernels: 4 items per thread, more when
// the grid would pass 2^22 workgroups per row
static dim3 adj_tiles(uint64_t items_row, int64_t bsz, int& reps) {
  reps = 4;
  uint64_t t = (items_row + (uint64_t)AJ_THREADS * reps - 1) / ((uint64_t)AJ_THREADS * reps);
  while (t > (1ull << 22)) {
    reps *= 2;
    t = (items_row + (uint64_t)AJ_THREADS * reps - 1) / ((uint64_t)AJ_THREADS * reps);
  }
  return dim3((unsigned)(t ? t : 1), (unsigned)bsz, 1);
}

static int build_insert_adj(InsertSpec& s, uint64_t mask, int n) {
  s.nb = 0;
  for (int p = 0; p < n; ++p)
    if ((mask >> p) & 1ull) {
      if (s.nb >= 12) return 0;
      s.pos[s.nb++] = p;
    }
  return 1;
}

static dim3 adj_grid(uint64_t items_row, int64_t bsz) {
  uint64_t want = ((uint64_t)num_cu() * 8 + bsz - 1) / bsz;
  uint64_t need = (items_row + AJ_THREADS - 1) / AJ_THREADS;
  uint64_t gx = want < need ? want : need;
  if (gx == 0) gx = 1;
  return dim3((unsigned)gx, (unsigned)bsz, 1);
}

}  // namespace tqsv

using namespace tqsv;

static int adj_common(void* psi, void* lam, int64_t bsz, int n, const void* mat, int P, float* grad,
                      float* work) {
  if (!psi || !lam || !mat || (P > 0 && (!grad || !work))) { set_error("adjoint: null pointer"); return E_ARG; }
  if (bsz < 1 || bsz > 65535 || n < 1 || n > 40) { set_error("adjoint: bad bsz/n (bsz <= 65535)"); return E_ARG; }
  if (P < 0 || P > 4) { set_error("adjoint: 0..4 parameters per gate (got %d)", P); return E_LIMIT; }
  return OK;
}

// The partial buffer must hold 4 floats per workgroup of the launch.
static int adj_work_check(const dim3& grid, int P, int64_t work_floats, const char* who) {
  if (P > 0 && (int64_t)grid.x * grid.y * 4 > work_floats) {
    set_error("%s: partial buffer of %lld floats, the launch needs %lld (tqsv_adjoint_work_floats)", who,
              (long long)work_floats, (long long)grid.x * grid.y * 4);
    return E_ARG;
  }
  return OK;
}

static int adj_finish(hipStream_t s, const dim3& grid, int P, const float* work, float* grad, int64_t gstride,
                      int gslot, const char* who) {
  int rc = check_launch(who);
  if (rc || P == 0) return rc;
  k_rowpart_reduce<<<grid.y, AJ_THREADS, 0, s>>>(work, grid.x, P, grad, gstride, gslot);
  return check_launch(who);
}

// mat: U (4 complex) followed by P derivative matrices dU/dtheta_p (4 each).
TQSV_API int tqsv_adjoint_1q(hipStream_t s, void* psi, void* lam, int64_t bsz, int n, int tbit, uint64_t cmask,
                             uint64_t cval, const void* mat, int64_t mat_bstride, int P, float* grad,
                             int64_t grad_stride, int gslot, float* work, int64_t work_floats) {
  int rc = adj_common(psi, lam, bsz, n, mat, P, grad, work);
  if (rc) return rc;
  if (tbit < 0 || tbit >= n) { set_error("adjoint_1q: target out of range"); return E_ARG; }
  const uint64_t tmask = 1ull << tbit;
  if ((cmask & tmask) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) { set_error("adjoint_1q: bad control mask"); return E_ARG; }
  InsertSpec ins;
  const int c = __builtin_popcountll(cmask);
  if (tbit >= 1 && !(cmask & 1ull)) {
    if (!build_insert_adj(ins, (cmask | tmask) >> 1, n - 1)) { set_error("adjoint_1q: too many controls"); return E_LIMIT; }
    const uint64_t items_row = 1ull << (n - 2 - c);
    int reps;
    const dim3 grid = adj_tiles(items_row, bsz, reps);
    if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_1q"))) return rc;
    const bool nt = stream_nt(2 * bsz, n);
    const uint64_t S = 1ull << (tbit - 1);
#define ADJ1Q4(B, NTV)                                                                                        \
  k_adj_1q4<B, NTV><<<grid, AJ_THREADS, 0, s>>>((float4*)psi, (float4*)lam, n, items_row, ins, cval >> 1, S, \
                                                (const float2*)mat, mat_bstride, P, work, reps)
    if (mat_bstride) { if (nt) ADJ1Q4(true, true); else ADJ1Q4(true, false); }
    else { if (nt) ADJ1Q4(false, true); else ADJ1Q4(false, false); }
#undef ADJ1Q4
    return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_1q");
  }
  if (!build_insert_adj(ins, cmask | tmask, n)) { set_error("adjoint_1q: too many controls"); return E_LIMIT; }
  const uint64_t items_row = 1ull << (n - 1 - c);
  const dim3 grid = adj_grid(items_row, bsz);
  if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_1q"))) return rc;
  if (mat_bstride)
    k_adj_1q<true><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, tmask,
                                               (const float2*)mat, mat_bstride, P, work);
  else
    k_adj_1q<false><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, tmask,
                                                (const float2*)mat, 0, P, work);
  return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_1q");
}

// mat: U (16 complex, index 2*b(hi)+b(lo)) followed by P derivative matrices.
TQSV_API int tqsv_adjoint_2q(hipStream_t s, void* psi, void* lam, int64_t bsz, int n, int bit_hi, int bit_lo,
                             uint64_t cmask, uint64_t cval, const void* mat, int64_t mat_bstride, int P,
                             float* grad, int64_t grad_stride, int gslot, float* work, int64_t work_floats) {
  int rc = adj_common(psi, lam, bsz, n, mat, P, grad, work);
  if (rc) return rc;
  if (bit_hi < 0 || bit_hi >= n || bit_lo < 0 || bit_lo >= n || bit_hi == bit_lo) {
    set_error("adjoint_2q: bad target bits");
    return E_ARG;
  }
  const uint64_t mhi = 1ull << bit_hi, mlo = 1ull << bit_lo;
  if ((cmask & (mhi | mlo)) || (n < 64 && (cmask >> n)) || (cval & ~cmask)) { set_error("adjoint_2q: bad control mask"); return E_ARG; }
  InsertSpec ins;
  const int c = __builtin_popcountll(cmask);
  const uint64_t all = cmask | mhi | mlo;
  if (!(all & 1ull)) {
    if (!build_insert_adj(ins, all >> 1, n - 1)) { set_error("adjoint_2q: too many controls"); return E_LIMIT; }
    const uint64_t items_row = 1ull << (n - 3 - c);
    int reps;
    const dim3 grid = adj_tiles(items_row, bsz, reps);
    if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_2q"))) return rc;
    const bool nt = stream_nt(2 * bsz, n);
#define ADJ2Q4(B, NTV)                                                                                        \
  k_adj_2q4<B, NTV><<<grid, AJ_THREADS, 0, s>>>((float4*)psi, (float4*)lam, n, items_row, ins, cval >> 1,    \
                                                mhi >> 1, mlo >> 1, (const float2*)mat, mat_bstride, P, work, \
                                                reps)
    if (mat_bstride) { if (nt) ADJ2Q4(true, true); else ADJ2Q4(true, false); }
    else { if (nt) ADJ2Q4(false, true); else ADJ2Q4(false, false); }
#undef ADJ2Q4
    return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_2q");
  }
  if (!build_insert_adj(ins, all, n)) { set_error("adjoint_2q: too many controls"); return E_LIMIT; }
  const uint64_t items_row = 1ull << (n - 2 - c);
  const dim3 grid = adj_grid(items_row, bsz);
  if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_2q"))) return rc;
  if (mat_bstride)
    k_adj_2q<true><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, mhi, mlo,
                                               (const float2*)mat, mat_bstride, P, work);
  else
    k_adj_2q<false><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, mhi, mlo,
                                                (const float2*)mat, 0, P, work);
  return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_2q");
}

// phases: (phi_even, phi_odd) followed by P derivative pairs.
TQSV_API int tqsv_adjoint_diag(hipStream_t s, void* psi, void* lam, int64_t bsz, int n, uint64_t cmask,
                               uint64_t cval, uint64_t pmask, const void* phases, int64_t ph_bstride, int P,
                               float* grad, int64_t grad_stride, int gslot, float* work, int64_t work_floats) {
  int rc = adj_common(psi, lam, bsz, n, phases, P, grad, work);
  if (rc) return rc;
  if ((n < 64 && ((cmask | pmask) >> n)) || (cval & ~cmask) || (cmask & pmask)) { set_error("adjoint_diag: bad masks"); return E_ARG; }
  InsertSpec ins;
  const int c = __builtin_popcountll(cmask);
  if (!(cmask & 1ull)) {
    if (!build_insert_adj(ins, cmask >> 1, n - 1)) { set_error("adjoint_diag: too many controls"); return E_LIMIT; }
    const uint64_t items_row = 1ull << (n - 1 - c);
    int reps;
    const dim3 grid = adj_tiles(items_row, bsz, reps);
    if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_diag"))) return rc;
    const bool nt = stream_nt(2 * bsz, n);
#define ADJD4(B, NTV)                                                                                           \
  k_adj_diag4<B, NTV><<<grid, AJ_THREADS, 0, s>>>((float4*)psi, (float4*)lam, n, items_row, ins, cval >> 1,    \
                                                  pmask, (const float2*)phases, ph_bstride, P, work, reps)
    if (ph_bstride) { if (nt) ADJD4(true, true); else ADJD4(true, false); }
    else { if (nt) ADJD4(false, true); else ADJD4(false, false); }
#undef ADJD4
    return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_diag");
  }
  if (!build_insert_adj(ins, cmask, n)) { set_error("adjoint_diag: too many controls"); return E_LIMIT; }
  const uint64_t items_row = 1ull << (n - c);
  const dim3 grid = adj_grid(items_row, bsz);
  if ((rc = adj_work_check(grid, P, work_floats, "tqsv_adjoint_diag"))) return rc;
  if (ph_bstride)
    k_adj_diag<true><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, pmask,
                                                 (const float2*)phases, ph_bstride, P, work);
  else
    k_adj_diag<false><<<grid, AJ_THREADS, 0, s>>>((float2*)psi, (float2*)lam, n, items_row, ins, cval, pmask,
                                                  (const float2*)phases, 0, P, work);
  return adj_finish(s, grid, P, work, grad, grad_stride, gslot, "tqsv_adjoint_diag");
}

// Floats of partial buffer that any tqsv_adjoint_1q / _2q / _diag launch on a
// [bsz, 2^n] state needs (4 per workgroup; the float4 tile kernels have the
// most workgroups per row).
TQSV_API int64_t tqsv_adjoint_work_floats(int64_t bsz, int n) {
  if (bsz < 1 || n < 1 || n > 40) return 0;
  int reps;
  const dim3 t = adj_tiles(n >= 2 ? 1ull << (n - 1) : 1ull, bsz, reps);
  const dim3 g = adj_grid(1ull << n, bsz);
  const int64_t gx = t.x > g.x ? t.x : g.x;
  return gx * bsz * 4;
}
