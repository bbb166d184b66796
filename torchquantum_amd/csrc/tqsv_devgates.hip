// Gate matrices built on the GPU from device-resident parameters.
//
// The reference builds every gate matrix with torch ops on the parameter
// tensor's device (rx.py:19-46, ry.py:50-67, rz.py:150-171, u3.py:52-77,
// u2 / phaseshift in functional/*.py) and, for a [bsz, P] parameter (an
// encoder feeding per-sample inputs, encoding/encodings.py:91-102), a
// [bsz, 2, 2] batch of matrices (gate_wrapper.py:373-376).  The engine
// otherwise evaluates matrices on the host in float64; for per-row
// parameters that live on the GPU this kernel writes them straight into the
// matrix pool instead: no device-to-host copy of the inputs and no host work
// per row.  One thread per (gate, row) writes U and, when the gate is
// trainable, dU/dtheta_p for each of its parameters, in the pool layout the
// stage and per-gate kernels read:
//   dense 2x2 ops:  [U (4), D_0 (4), ..., D_{P-1} (4)] complex per row;
//   diagonal ops:   [U[a][a], U[b][b], D_0[a][a], D_0[b][b], ...] per row.
#include "tqsv_common.h"

namespace tqsv {

enum DevGateCode : int32_t {
  DG_RX = 1,
  DG_RY = 2,
  DG_RZ = 3,
  DG_PHASE = 4,   // phaseshift / u1: diag(1, e^{i theta})
  DG_U2 = 5,
  DG_U3 = 6,
};

enum DevGateFlags : int32_t {
  DGF_INVERSE = 1,   // U -> U^dagger (and D -> D^dagger)
  DGF_CONJ = 2,      // U -> conj(U) (density-matrix ket/bra copy), after the inverse
  DGF_DERIV = 4,     // write the derivative blocks
  DGF_DIAG = 8,      // diagonal layout: entries (a, a) and (b, b) only
};

struct DevGate {
  const float* params;   // row r, parameter q at params[r * rstride + q * cstride]
  int64_t rstride, cstride;
  int64_t dst;           // pool offset (complex elements) of row 0's block
  int64_t bstride;       // complex elements between rows' blocks
  int64_t adj;           // pool offset of row 0's U^dagger block (-1: none)
  int64_t abstride;
  int32_t code, flags, rows, diag_ab;   // diag_ab: a | (b << 1)
};

__device__ __forceinline__ float2 cexpi(float t) {
  float s, c;
  sincosf(t, &s, &c);
  return make_float2(c, s);
}
__device__ __forceinline__ float2 cscale(float2 a, float k) { return make_float2(a.x * k, a.y * k); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }   // i * a

// m[0..3] = row-major 2x2; np = parameters; d[q][0..3] = dU/dtheta_q
__device__ void dev_gate_2x2(int code, const float* th, float2 (&m)[4], float2 (&d)[3][4], int& np) {
  const float2 z = make_float2(0.f, 0.f);
  switch (code) {
    case DG_RX: {
      float s, c;
      sincosf(0.5f * th[0], &s, &c);
      m[0] = make_float2(c, 0.f); m[1] = make_float2(0.f, -s); m[2] = make_float2(0.f, -s); m[3] = make_float2(c, 0.f);
      d[0][0] = make_float2(-0.5f * s, 0.f); d[0][1] = make_float2(0.f, -0.5f * c);
      d[0][2] = make_float2(0.f, -0.5f * c); d[0][3] = make_float2(-0.5f * s, 0.f);
      np = 1;
      break;
    }
    case DG_RY: {
      float s, c;
      sincosf(0.5f * th[0], &s, &c);
      m[0] = make_float2(c, 0.f); m[1] = make_float2(-s, 0.f); m[2] = make_float2(s, 0.f); m[3] = make_float2(c, 0.f);
      d[0][0] = make_float2(-0.5f * s, 0.f); d[0][1] = make_float2(-0.5f * c, 0.f);
      d[0][2] = make_float2(0.5f * c, 0.f); d[0][3] = make_float2(-0.5f * s, 0.f);
      np = 1;
      break;
    }
    case DG_RZ: {
      const float2 e0 = cexpi(-0.5f * th[0]), e1 = cexpi(0.5f * th[0]);
      m[0] = e0; m[1] = z; m[2] = z; m[3] = e1;
      d[0][0] = cscale(mul_i(e0), -0.5f); d[0][1] = z; d[0][2] = z; d[0][3] = cscale(mul_i(e1), 0.5f);
      np = 1;
      break;
    }
    case DG_PHASE: {
      const float2 e = cexpi(th[0]);
      m[0] = make_float2(1.f, 0.f); m[1] = z; m[2] = z; m[3] = e;
      d[0][0] = z; d[0][1] = z; d[0][2] = z; d[0][3] = mul_i(e);
      np = 1;
      break;
    }
    case DG_U2: {   // u3(pi/2, phi, lam)
      const float r = 0.70710678118654752f;
      const float2 ep = cexpi(th[0]), el = cexpi(th[1]), epl = cexpi(th[0] + th[1]);
      m[0] = make_float2(r, 0.f); m[1] = cscale(el, -r); m[2] = cscale(ep, r); m[3] = cscale(epl, r);
      d[0][0] = z; d[0][1] = z; d[0][2] = cscale(mul_i(ep), r); d[0][3] = cscale(mul_i(epl), r);
      d[1][0] = z; d[1][1] = cscale(mul_i(el), -r); d[1][2] = z; d[1][3] = cscale(mul_i(epl), r);
      np = 2;
      break;
    }
    default: {   // DG_U3: [[c, -e^{i lam} s], [e^{i phi} s, e^{i(phi+lam)} c]]
      float s, c;
      sincosf(0.5f * th[0], &s, &c);
      const float2 ep = cexpi(th[1]), el = cexpi(th[2]), epl = cexpi(th[1] + th[2]);
      m[0] = make_float2(c, 0.f); m[1] = cscale(el, -s); m[2] = cscale(ep, s); m[3] = cscale(epl, c);
      d[0][0] = make_float2(-0.5f * s, 0.f); d[0][1] = cscale(el, -0.5f * c);
      d[0][2] = cscale(ep, 0.5f * c); d[0][3] = cscale(epl, -0.5f * s);
      d[1][0] = z; d[1][1] = z; d[1][2] = mul_i(m[2]); d[1][3] = mul_i(m[3]);
      d[2][0] = z; d[2][1] = mul_i(m[1]); d[2][2] = z; d[2][3] = mul_i(m[3]);
      np = 3;
      break;
    }
  }
}

__device__ __forceinline__ void dev_transform(float2 (&a)[4], int flags) {
  if (flags & DGF_INVERSE) {   // conjugate transpose
    const float2 t = a[1];
    a[1] = make_float2(a[2].x, -a[2].y);
    a[2] = make_float2(t.x, -t.y);
    a[0].y = -a[0].y;
    a[3].y = -a[3].y;
  }
  if (flags & DGF_CONJ) {
    for (int e = 0; e < 4; ++e) a[e].y = -a[e].y;
  }
}

// grid: x over rows (256 per block), y over items
__global__ void k_build_dev_gates(float2* __restrict__ pool, const DevGate* __restrict__ items) {
  const DevGate& gi = items[blockIdx.y];
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= gi.rows) return;
  const DevGate g = gi;
  float th[3] = {0.f, 0.f, 0.f};
  const int want = (g.code == DG_U3) ? 3 : (g.code == DG_U2 ? 2 : 1);
  for (int q = 0; q < want; ++q) th[q] = g.params[r * g.rstride + q * g.cstride];
  float2 m[4], d[3][4];
  int np = 0;
  dev_gate_2x2(g.code, th, m, d, np);
  dev_transform(m, g.flags);
  float2* out = pool + g.dst + r * g.bstride;
  const bool diag = (g.flags & DGF_DIAG) != 0;
  const int ia = (g.diag_ab & 1) * 3, ib = ((g.diag_ab >> 1) & 1) * 3;   // (a, a) / (b, b) in a 2x2
  if (diag) {
    out[0] = m[ia];
    out[1] = m[ib];
  } else {
    for (int e = 0; e < 4; ++e) out[e] = m[e];
  }
  if (g.flags & DGF_DERIV) {
    for (int q = 0; q < np; ++q) {
      dev_transform(d[q], g.flags);
      if (diag) {
        out[2 + 2 * q] = d[q][ia];
        out[3 + 2 * q] = d[q][ib];
      } else {
        for (int e = 0; e < 4; ++e) out[4 + 4 * q + e] = d[q][e];
      }
    }
  }
  if (g.adj >= 0) {   // U^dagger of the (transformed) matrix
    float2* a = pool + g.adj + r * g.abstride;
    if (diag) {
      a[0] = make_float2(m[ia].x, -m[ia].y);
      a[1] = make_float2(m[ib].x, -m[ib].y);
    } else {
      a[0] = make_float2(m[0].x, -m[0].y);
      a[1] = make_float2(m[2].x, -m[2].y);
      a[2] = make_float2(m[1].x, -m[1].y);
      a[3] = make_float2(m[3].x, -m[3].y);
    }
  }
}

}  // namespace tqsv

using namespace tqsv;

TQSV_API int tqsv_dev_gate_bytes(void) { return (int)sizeof(DevGate); }

// Write the pool blocks of gates whose parameters live on the GPU.  items_dev:
// DevGate array in device memory; max_rows: the largest row count of an item.
TQSV_API int tqsv_build_dev_gates(hipStream_t s, void* pool, const void* items_dev, int n_items, int64_t max_rows) {
  if (!pool || !items_dev || n_items < 0 || n_items > 65535 || max_rows < 0) {
    set_error("tqsv_build_dev_gates: bad arguments");
    return E_ARG;
  }
  if (n_items == 0 || max_rows == 0) return OK;
  const dim3 grid((unsigned)((max_rows + 255) / 256), (unsigned)n_items);
  k_build_dev_gates<<<grid, 256, 0, s>>>((float2*)pool, (const DevGate*)items_dev);
  return check_launch("tqsv_build_dev_gates");
}
