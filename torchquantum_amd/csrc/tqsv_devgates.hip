// Gate matrices built on the GPU from device-resident parameters.
//
// The reference builds every gate matrix with torch ops on the parameter
// tensor's device (rx.py:19-46, ry.py:50-67, rz.py:150-171, u3.py:52-77,
// u2 / rot / phaseshift in functional/*.py): for a [bsz, P] parameter (an
// encoder feeding per-sample inputs, encoding/encodings.py:91-102) a
// [bsz, 2, 2] batch (gate_wrapper.py:373-376), for a trainable [1, P]
// Operator parameter one matrix.  The engine otherwise evaluates matrices on
// the host in float64, which needs the parameter values on the host (a
// device-to-host copy whenever they changed, e.g. after every optimizer
// step).  For single-wire ops whose parameters live on the GPU this kernel
// writes the matrix-pool blocks instead: one thread per (op, row) builds the
// op -- one gate, or a merged run of single-wire gates (the product of its
// members, with the product rule for each trainable member's derivatives) --
// in the layout the stage and per-gate kernels read:
//   dense 2x2 ops:  [U (4), D_0 (4), ..., D_{P-1} (4)] complex per row;
//   diagonal ops:   [U[a][a], U[b][b], D_0[a][a], D_0[b][b], ...] per row;
//   and U^dagger (or the conjugated diagonal) at the op's adjoint block.
#include "tqsv_common.h"

namespace tqsv {

enum DevGateCode : int32_t {
  DG_FIXED = 0,   // a fixed matrix given in the member record
  DG_RX = 1,
  DG_RY = 2,
  DG_RZ = 3,
  DG_PHASE = 4,   // phaseshift / u1: diag(1, e^{i theta})
  DG_U2 = 5,
  DG_U3 = 6,
  DG_ROT = 7,     // RZ(omega) RY(theta) RZ(phi)
};

enum DevGateFlags : int32_t {
  DGF_INVERSE = 1,   // member: U -> U^dagger (and D -> D^dagger)
  DGF_CONJ = 2,      // member: U -> conj(U) (density-matrix ket/bra copy), after the inverse
  DGF_DERIV = 4,     // member: its parameters are trainable; op: write the derivative blocks
  DGF_DIAG = 8,      // op: diagonal layout, entries (a, a) and (b, b); a, b in bits 8 and 9
};

constexpr int DG_MAX_MEMBERS = 8;
constexpr int DG_MAX_DERIV = 4;

struct DevMember {
  const float* params;   // row r, parameter q at params[r * rstride + q * cstride] (rstride 0: shared)
  int64_t rstride, cstride;
  int32_t code, flags;
  float fixed[8];        // DG_FIXED: row-major 2x2, (re, im) pairs, already transformed
};

struct DevOp {
  int64_t dst;           // pool offset (complex elements) of row 0's block
  int64_t bstride;       // complex elements between rows' blocks (0: shared)
  int64_t adj;           // pool offset of row 0's adjoint block (-1: none)
  int64_t abstride;
  int32_t rows, m0, nm, flags;   // rows; members m0 .. m0 + nm - 1 in application order
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
    case DG_ROT: {   // RZ(w) RY(t) RZ(p) = [[e^{-i(p+w)/2} c, -e^{i(p-w)/2} s], [e^{-i(p-w)/2} s, e^{i(p+w)/2} c]]
      float s, c;
      sincosf(0.5f * th[1], &s, &c);
      const float hp = 0.5f * (th[0] + th[2]), hm = 0.5f * (th[0] - th[2]);
      const float2 a = cexpi(-hp), b = cexpi(hm), bb = cexpi(-hm), aa = cexpi(hp);
      m[0] = cscale(a, c); m[1] = cscale(b, -s); m[2] = cscale(bb, s); m[3] = cscale(aa, c);
      // d/dphi: entries scale by -i/2, +i/2, -i/2, +i/2; d/domega: -i/2, -i/2, +i/2, +i/2
      d[0][0] = cscale(mul_i(m[0]), -0.5f); d[0][1] = cscale(mul_i(m[1]), 0.5f);
      d[0][2] = cscale(mul_i(m[2]), -0.5f); d[0][3] = cscale(mul_i(m[3]), 0.5f);
      d[1][0] = cscale(a, -0.5f * s); d[1][1] = cscale(b, -0.5f * c);
      d[1][2] = cscale(bb, 0.5f * c); d[1][3] = cscale(aa, -0.5f * s);
      d[2][0] = cscale(mul_i(m[0]), -0.5f); d[2][1] = cscale(mul_i(m[1]), -0.5f);
      d[2][2] = cscale(mul_i(m[2]), 0.5f); d[2][3] = cscale(mul_i(m[3]), 0.5f);
      np = 3;
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

__device__ __forceinline__ void mat2_mul(const float2 (&a)[4], const float2 (&b)[4], float2 (&c)[4]) {
  // c = a b (row-major 2x2)
  for (int r = 0; r < 2; ++r)
    for (int k = 0; k < 2; ++k) c[2 * r + k] = cfma(a[2 * r + 1], b[2 + k], cmul(a[2 * r], b[k]));
}

__device__ __forceinline__ void mat2_copy(float2 (&dst)[4], const float2 (&src)[4]) {
  for (int e = 0; e < 4; ++e) dst[e] = src[e];
}

// grid: x over rows (256 per block), y over ops
__global__ void k_build_dev_gates(float2* __restrict__ pool, const DevOp* __restrict__ ops,
                                  const DevMember* __restrict__ members) {
  const DevOp op = ops[blockIdx.y];
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= op.rows) return;
  const int nm = op.nm < DG_MAX_MEMBERS ? op.nm : DG_MAX_MEMBERS;
  float2 M[DG_MAX_MEMBERS][4];
  float2 D[DG_MAX_DERIV][4];
  int dmem[DG_MAX_DERIV];
  int nd = 0;
  for (int mi = 0; mi < nm; ++mi) {
    const DevMember mb = members[op.m0 + mi];
    if (mb.code == DG_FIXED) {
      for (int e = 0; e < 4; ++e) M[mi][e] = make_float2(mb.fixed[2 * e], mb.fixed[2 * e + 1]);
      continue;
    }
    const float* pr = mb.params + (mb.rstride ? r * mb.rstride : 0);
    float th[3] = {0.f, 0.f, 0.f};
    const int want = (mb.code == DG_U3 || mb.code == DG_ROT) ? 3 : (mb.code == DG_U2 ? 2 : 1);
    for (int q = 0; q < want; ++q) th[q] = pr[q * mb.cstride];
    float2 dm[3][4];
    int np = 0;
    dev_gate_2x2(mb.code, th, M[mi], dm, np);
    dev_transform(M[mi], mb.flags);
    if ((mb.flags & DGF_DERIV) && (op.flags & DGF_DERIV)) {
      for (int q = 0; q < np && nd < DG_MAX_DERIV; ++q) {
        dev_transform(dm[q], mb.flags);
        mat2_copy(D[nd], dm[q]);
        dmem[nd++] = mi;
      }
    }
  }
  // U = M[nm-1] ... M[0]
  float2 U[4], T[4];
  mat2_copy(U, M[0]);
  for (int mi = 1; mi < nm; ++mi) {
    mat2_mul(M[mi], U, T);
    mat2_copy(U, T);
  }
  // dU for a trainable member m: (M[nm-1] ... M[m+1]) D (M[m-1] ... M[0])
  for (int q = 0; q < nd; ++q) {
    const int m = dmem[q];
    float2 A[4];
    mat2_copy(A, D[q]);
    for (int mi = m - 1; mi >= 0; --mi) {     // right factors: A <- A M[mi], innermost first
      mat2_mul(A, M[mi], T);
      mat2_copy(A, T);
    }
    for (int mi = m + 1; mi < nm; ++mi) {     // left factors: A <- M[mi] A
      mat2_mul(M[mi], A, T);
      mat2_copy(A, T);
    }
    mat2_copy(D[q], A);
  }
  float2* out = pool + op.dst + r * op.bstride;
  const bool diag = (op.flags & DGF_DIAG) != 0;
  const int ia = ((op.flags >> 8) & 1) * 3, ib = ((op.flags >> 9) & 1) * 3;   // (a, a) / (b, b) of a 2x2
  if (diag) {
    out[0] = U[ia];
    out[1] = U[ib];
    for (int q = 0; q < nd; ++q) {
      out[2 + 2 * q] = D[q][ia];
      out[3 + 2 * q] = D[q][ib];
    }
  } else {
    for (int e = 0; e < 4; ++e) out[e] = U[e];
    for (int q = 0; q < nd; ++q)
      for (int e = 0; e < 4; ++e) out[4 + 4 * q + e] = D[q][e];
  }
  if (op.adj >= 0) {   // U^dagger
    float2* a = pool + op.adj + r * op.abstride;
    if (diag) {
      a[0] = make_float2(U[ia].x, -U[ia].y);
      a[1] = make_float2(U[ib].x, -U[ib].y);
    } else {
      a[0] = make_float2(U[0].x, -U[0].y);
      a[1] = make_float2(U[2].x, -U[2].y);
      a[2] = make_float2(U[1].x, -U[1].y);
      a[3] = make_float2(U[3].x, -U[3].y);
    }
  }
}

}  // namespace tqsv

using namespace tqsv;

TQSV_API int tqsv_dev_gate_bytes(void) { return (int)sizeof(DevOp) | ((int)sizeof(DevMember) << 16); }

// Write the pool blocks of single-wire ops whose parameters live on the GPU.
// ops_dev / members_dev: DevOp and DevMember arrays in device memory;
// max_rows: the largest row count of an op.
TQSV_API int tqsv_build_dev_gates(hipStream_t s, void* pool, const void* ops_dev, const void* members_dev, int n_ops,
                                  int64_t max_rows) {
  if (!pool || !ops_dev || !members_dev || n_ops < 0 || n_ops > 65535 || max_rows < 0) {
    set_error("tqsv_build_dev_gates: bad arguments");
    return E_ARG;
  }
  if (n_ops == 0 || max_rows == 0) return OK;
  const dim3 grid((unsigned)((max_rows + 255) / 256), (unsigned)n_ops);
  k_build_dev_gates<<<grid, 256, 0, s>>>((float2*)pool, (const DevOp*)ops_dev, (const DevMember*)members_dev);
  return check_launch("tqsv_build_dev_gates");
}
