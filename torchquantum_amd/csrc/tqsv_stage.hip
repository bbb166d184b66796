// Fused "stage" kernels: one launch applies a whole run of gates to every
// batch row in a single read + write of the state.
//
// The reference applies gates one at a time (gate_wrapper.py:96-140); its own
// fusion ("static mode", torchquantum/graph/graphs.py:384-532, 836-860)
// multiplies a k-wire block into one 2^k x 2^k matrix and then still applies
// it with the per-gate bmm.  Here instead:
//
//   * a stage owns m "local" bits of the in-row index (m <= 13); a workgroup
//     processes one group of 2^m amplitudes (all values of the local bits,
//     the other bits fixed) at a time, in a persistent loop over groups;
//   * each thread keeps NS = 16 amplitudes in VGPRs: R = 4 local bits are
//     "register bits" (the slot index), the other T = m - 4 bits are the
//     thread index.  Gates whose mixing targets are register bits run with no
//     communication at all; control bits and diagonal (phase) gates may sit
//     on any local bit;
//   * a REMAP op re-assigns which local bits are register bits by one
//     LDS write + read of the group (XOR-swizzled to spread banks);
//   * the canonical layout (register bits = the top R local bits) puts the
//     lowest local bits on the thread index, so lanes read consecutive
//     addresses.  A stage whose first ops need other register bits, none of
//     them among the lowest 4 local bits, loads in that layout instead of
//     loading canonically and remapping; likewise it stores in the layout of
//     its last ops.  Header word H_FLAGS: bits 0-1 the op features, bits
//     2-14 the register local bits (a mask) of the load layout, bits 15-27
//     those of the store layout, 0 meaning canonical.  Lanes 16 apart still
//     cover whole 128-byte lines in either.
//
// The adjoint kernel walks the same program backwards on two states at once
// (psi: the forward output being un-computed, lam: d loss / d conj(psi) in
// torch's convention), applying U^dagger to both and accumulating, for every
// parameter p of a gate, Re <lam | dU/dtheta_p psi_prev> -- the adjoint
// differentiation method, with O(1) extra states instead of autograd's saved
// per-gate tensors (the reference's backward, gate_wrapper.py:127-136).
#include "tqsv_common.h"

namespace tqsv {
int num_cu();
}

namespace tqsv {

constexpr int R = 4;
constexpr int NS = 1 << R;
constexpr int MAX_GP = 1024;  // grad params per stage (LDS accumulators)

enum StageOp { OP_MAT1 = 1, OP_MAT2 = 2, OP_DIAG = 3, OP_MATK = 4, OP_REMAP = 5 };

// stage header words
enum Hdr {
  H_NOPS = 0, H_M = 1, H_T = 2, H_NNL = 3, H_OFF_OPS = 4, H_NGP = 5, H_OFF_GSLOT = 6, H_FLAGS = 7, H_POS = 8
};

__device__ __forceinline__ int swz(int l) { return l ^ ((l >> 5) & 31) ^ ((l >> 10) & 31); }

// ------------------------------------------------------------------ forward
template <int K>
__device__ __forceinline__ void f_mat1(float2 (&a)[NS], const float2* __restrict__ M, bool tok, int rm, int rv) {
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    const float2 x0 = a[j], x1 = a[j1];
    const float2 y0 = cfma(u1, x1, cmul(u0, x0));
    const float2 y1 = cfma(u3, x1, cmul(u2, x0));
    a[j] = csel(ok, y0, x0);
    a[j1] = csel(ok, y1, x1);
  }
}

template <int KH, int KL>
__device__ __forceinline__ void f_mat2(float2 (&a)[NS], const float2* __restrict__ M, bool tok, int rm, int rv) {
  float2 u[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) u[q] = M[q];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & ((1 << KH) | (1 << KL))) continue;
    const int idx[4] = {j, j | (1 << KL), j | (1 << KH), j | (1 << KH) | (1 << KL)};
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    float2 x[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = a[idx[r]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = cmul(u[4 * r], x[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cfma(u[4 * r + c], x[c], acc);
      a[idx[r]] = csel(ok, acc, x[r]);
    }
  }
}

// dense op on KB = 3 or 4 register bits; EX = the excluded register bit for
// KB = 3 (matrix index bit q <-> q-th remaining register bit, ascending).
template <int KB, int EX>
__device__ __forceinline__ int matk_slot(int base, int r) {
  int s = base;
  int q = 0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (KB == 3 && k == EX) continue;
    if ((r >> q) & 1) s |= 1 << k;
    ++q;
  }
  return s;
}

template <int KB, int EX>
__device__ __forceinline__ void f_matk(float2 (&a)[NS], const float2* __restrict__ M, bool tok) {
  constexpr int D = 1 << KB;
#pragma unroll
  for (int base = 0; base < NS; ++base) {
    if (KB == 4 && base != 0) continue;
    if (KB == 3 && (base != 0 && base != (1 << EX))) continue;
    float2 x[D];
#pragma unroll
    for (int r = 0; r < D; ++r) x[r] = a[matk_slot<KB, EX>(base, r)];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = cmul(M[r * D], x[0]);
#pragma unroll
      for (int c = 1; c < D; ++c) acc = cfma(M[r * D + c], x[c], acc);
      a[matk_slot<KB, EX>(base, r)] = csel(tok, acc, x[r]);
    }
  }
}

__device__ __forceinline__ void f_diag(float2 (&a)[NS], const float2* __restrict__ M, bool tok, int rm, int rv,
                                       int tpar, int rpm) {
  const float2 p0 = M[0], p1 = M[1];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    const int par = (tpar + __popc(j & rpm)) & 1;
    const float2 f = par ? p1 : p0;
    a[j] = csel(ok, cmul(f, a[j]), a[j]);
  }
}

// ------------------------------------------------------------------ adjoint
// psi <- U^dag psi ; g_p += Re<lam | D_p psi> ; lam <- U^dag lam
template <int K>
__device__ __forceinline__ void b_mat1(float2 (&p)[NS], float2 (&l)[NS], const float2* __restrict__ M, int P,
                                       bool tok, int rm, int rv, float (&g)[4]) {
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    const float2 x0 = p[j], x1 = p[j1];
    const float2 y0 = cfmac(u2, x1, cmulc(u0, x0));
    const float2 y1 = cfmac(u3, x1, cmulc(u1, x0));
    const float2 l0 = l[j], l1 = l[j1];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < P) {
        const float2* D = M + 4 * (q + 1);
        const float2 v0 = cfma(D[1], y1, cmul(D[0], y0));
        const float2 v1 = cfma(D[3], y1, cmul(D[2], y0));
        const float t = redot(l0, v0) + redot(l1, v1);
        g[q] += ok ? t : 0.f;
      }
    }
    const float2 m0 = cfmac(u2, l1, cmulc(u0, l0));
    const float2 m1 = cfmac(u3, l1, cmulc(u1, l0));
    p[j] = csel(ok, y0, x0);
    p[j1] = csel(ok, y1, x1);
    l[j] = csel(ok, m0, l0);
    l[j1] = csel(ok, m1, l1);
  }
}

// Single-wire op with trainable parameters, Hermitian-generator form.
// For a unitary U(theta), H_p = i (dU/dtheta_p) U^dag is Hermitian, so with
// x = psi before the un-apply (the op's output) and lam the incoming
// gradient, dL/dtheta_p = Re <lam| dU_p |U^dag x> = Im <lam| H_p |x>.  Writing
// H_p = a I + b X + c Y + d Z (real a..d), every parameter's gradient is a
// combination of the four sums S = Im <lam|s|x>, s = I, X, Y, Z, which this
// op accumulates (12 FMA per amplitude pair); tqsv_mg_to_grad folds them
// with the coefficients of H_p afterwards.  Then psi <- U^dag psi and
// lam <- U^dag lam.  NC: the op has no controls (every pair is updated).
template <int K, bool NC>
__device__ __forceinline__ void b_mat1_mg(float2 (&p)[NS], float2 (&l)[NS], const float2* __restrict__ M, bool tok,
                                          int rm, int rv, float (&g)[8]) {
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
  // a = sum Im conj(l0) x0, b = sum Im conj(l1) x1, c = sum conj(l0) x1, d = sum conj(l1) x0
  float a = 0.f, b = 0.f;
  float2 c = make_float2(0.f, 0.f), d = c;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    if (!NC && (j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = NC || tok;
    const float2 x0 = p[j], x1 = p[j1];
    const float2 l0 = l[j], l1 = l[j1];
    const float2 z = make_float2(0.f, 0.f);
    const float2 g0 = ok ? l0 : z, g1 = ok ? l1 : z;
    a = fmaf(g0.x, x0.y, fmaf(-g0.y, x0.x, a));
    b = fmaf(g1.x, x1.y, fmaf(-g1.y, x1.x, b));
    c = cfmac(g0, x1, c);
    d = cfmac(g1, x0, d);
    const float2 y0 = cfmac(u2, x1, cmulc(u0, x0));
    const float2 y1 = cfmac(u3, x1, cmulc(u1, x0));
    const float2 n0 = cfmac(u2, l1, cmulc(u0, l0));
    const float2 n1 = cfmac(u3, l1, cmulc(u1, l0));
    if (NC) {
      p[j] = y0;
      p[j1] = y1;
      l[j] = n0;
      l[j1] = n1;
    } else {
      p[j] = csel(ok, y0, x0);
      p[j1] = csel(ok, y1, x1);
      l[j] = csel(ok, n0, l0);
      l[j1] = csel(ok, n1, l1);
    }
  }
  // S_I = Im<l|x>, S_X = Im<l|X|x>, S_Y = Im<l|Y|x> = Re conj(l1) x0 - Re conj(l0) x1, S_Z = Im<l|Z|x>
  g[0] += a + b;
  g[1] += c.y + d.y;
  g[2] += d.x - c.x;
  g[3] += a - b;
}

template <int KH, int KL>
__device__ __forceinline__ void b_mat2(float2 (&p)[NS], float2 (&l)[NS], const float2* __restrict__ M, int P,
                                       bool tok, int rm, int rv, float (&g)[4]) {
  float2 u[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) u[q] = M[q];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & ((1 << KH) | (1 << KL))) continue;
    const int idx[4] = {j, j | (1 << KL), j | (1 << KH), j | (1 << KH) | (1 << KL)};
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    float2 x[4], y[4], lm[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) { x[r] = p[idx[r]]; lm[r] = l[idx[r]]; }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = cmulc(u[r], x[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cfmac(u[4 * c + r], x[c], acc);
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
          t += redot(lm[r], v);
        }
        g[q] += ok ? t : 0.f;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = cmulc(u[r], lm[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cfmac(u[4 * c + r], lm[c], acc);
      p[idx[r]] = csel(ok, y[r], x[r]);
      l[idx[r]] = csel(ok, acc, lm[r]);
    }
  }
}

template <int KB, int EX>
__device__ __forceinline__ void b_matk1(float2 (&a)[NS], const float2* __restrict__ M, bool tok) {
  constexpr int D = 1 << KB;
#pragma unroll
  for (int base = 0; base < NS; ++base) {
    if (KB == 4 && base != 0) continue;
    if (KB == 3 && (base != 0 && base != (1 << EX))) continue;
    float2 x[D];
#pragma unroll
    for (int r = 0; r < D; ++r) x[r] = a[matk_slot<KB, EX>(base, r)];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = cmulc(M[r], x[0]);
#pragma unroll
      for (int c = 1; c < D; ++c) acc = cfmac(M[c * D + r], x[c], acc);
      a[matk_slot<KB, EX>(base, r)] = csel(tok, acc, x[r]);
    }
  }
}

__device__ __forceinline__ void b_diag(float2 (&p)[NS], float2 (&l)[NS], const float2* __restrict__ M, int P,
                                       bool tok, int rm, int rv, int tpar, int rpm, float (&g)[4]) {
  const float2 p0 = M[0], p1 = M[1];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    const int par = (tpar + __popc(j & rpm)) & 1;
    const float2 f = par ? p1 : p0;
    const float2 y = cmulc(f, p[j]);
    const float2 lj = l[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < P) {
        const float2 df = par ? M[2 * (q + 1) + 1] : M[2 * (q + 1)];
        g[q] += ok ? redot(lj, cmul(df, y)) : 0.f;
      }
    }
    p[j] = csel(ok, y, p[j]);
    l[j] = csel(ok, cmulc(f, lj), lj);
  }
}

// ------------------------------------------------------------------ fast paths
// Op flag bits (word 0): the compiler sets them from the gate structure.
enum OpFlag {
  F_BATCHED = 1 << 8, F_NOCTRL = 1 << 9, F_REAL = 1 << 10, F_XPERM = 1 << 11, F_SIGN = 1 << 12,
  F_MG = 1 << 13   // adjoint: accumulate the four Hermitian-generator sums Im<lam|s|psi>, s = I, X, Y, Z
};

// no controls: every pair is updated, no selects
template <int K>
__device__ __forceinline__ void f_mat1_nc(float2 (&a)[NS], const float2* __restrict__ M) {
  const float2 u0 = M[0], u1 = M[1], u2 = M[2], u3 = M[3];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    const float2 x0 = a[j], x1 = a[j1];
    a[j] = cfma(u1, x1, cmul(u0, x0));
    a[j1] = cfma(u3, x1, cmul(u2, x0));
  }
}

// real 2x2 (RY, H, ...): 8 FMA per pair
template <int K>
__device__ __forceinline__ void f_mat1_real_nc(float2 (&a)[NS], const float2* __restrict__ M) {
  const float r0 = M[0].x, r1 = M[1].x, r2 = M[2].x, r3 = M[3].x;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    const float2 x0 = a[j], x1 = a[j1];
    a[j] = make_float2(fmaf(r0, x0.x, r1 * x1.x), fmaf(r0, x0.y, r1 * x1.y));
    a[j1] = make_float2(fmaf(r2, x0.x, r3 * x1.x), fmaf(r2, x0.y, r3 * x1.y));
  }
}

// X on a register bit under controls (CNOT, Toffoli, ...): a pure exchange
template <int K>
__device__ __forceinline__ void f_xperm(float2 (&a)[NS], bool tok, int rm, int rv) {
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j & (1 << K)) continue;
    const int j1 = j | (1 << K);
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    const float2 x0 = a[j], x1 = a[j1];
    a[j] = csel(ok, x1, x0);
    a[j1] = csel(ok, x0, x1);
  }
}

// sign flip where the controls match (Z, CZ, CCZ): phase -1, no parity
__device__ __forceinline__ void f_sign(float2 (&a)[NS], bool tok, int rm, int rv) {
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if ((j & rm) != rv) continue;  // register controls: a wave-uniform skip
    const bool ok = tok;
    a[j] = make_float2(ok ? -a[j].x : a[j].x, ok ? -a[j].y : a[j].y);
  }
}

// ------------------------------------------------------------------ remap
__device__ __forceinline__ int layout_thread(const int* __restrict__ tl, int T, int tid) {
  int lt = 0;
  for (int i = 0; i < T; ++i) lt |= ((tid >> i) & 1) ? tl[i] : 0;
  return lt;
}

__device__ __forceinline__ void remap(float2 (&a)[NS], float2* __restrict__ lds, const int* __restrict__ from,
                                      const int* __restrict__ to, int T, int tid, bool active = true) {
  // record: tl[T], rl[R] (local-index value of each thread / register bit).
  // swz is linear over XOR, so address = swz(thread part) ^ swz(slot part),
  // the slot part being wave-uniform.
  int st0 = swz(layout_thread(from, T, tid));
  __syncthreads();
  if (active) {
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    int lr = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) lr |= ((j >> k) & 1) ? from[T + k] : 0;
    asm volatile("" : "+v"(st0));
    lds[st0 ^ swz(lr)] = a[j];
  }
  }
  __syncthreads();
  if (!active) return;
  int st1 = swz(layout_thread(to, T, tid));
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    int lr = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) lr |= ((j >> k) & 1) ? to[T + k] : 0;
    asm volatile("" : "+v"(st1));
    a[j] = lds[st1 ^ swz(lr)];
  }
}

// ------------------------------------------------------------------ dispatch
// FEAT: bit 0 = the stage contains 2-bit dense ops, bit 1 = 3/4-bit dense ops.
// Stages without them run a leaner instantiation (fewer live registers).
constexpr int FEAT_MAT2 = 1, FEAT_MATK = 2;

template <int FEAT>
__device__ __forceinline__ void fwd_op(float2 (&a)[NS], const int* __restrict__ w, const float2* __restrict__ M,
                                       int tid) {
  const int type = w[0] & 15;
  const int fl = w[0];
  if (type == OP_MAT1) {
    if (fl & F_XPERM) {
      const bool tok = (tid & w[4]) == w[5];
      const int rm = w[6] & 255, rv = (w[6] >> 8) & 255;
      switch (w[3]) {
        case 0: f_xperm<0>(a, tok, rm, rv); break;
        case 1: f_xperm<1>(a, tok, rm, rv); break;
        case 2: f_xperm<2>(a, tok, rm, rv); break;
        default: f_xperm<3>(a, tok, rm, rv); break;
      }
    } else if (fl & F_NOCTRL) {
      if (fl & F_REAL) {
        switch (w[3]) {
          case 0: f_mat1_real_nc<0>(a, M); break;
          case 1: f_mat1_real_nc<1>(a, M); break;
          case 2: f_mat1_real_nc<2>(a, M); break;
          default: f_mat1_real_nc<3>(a, M); break;
        }
      } else {
        switch (w[3]) {
          case 0: f_mat1_nc<0>(a, M); break;
          case 1: f_mat1_nc<1>(a, M); break;
          case 2: f_mat1_nc<2>(a, M); break;
          default: f_mat1_nc<3>(a, M); break;
        }
      }
    } else {
      const bool tok = (tid & w[4]) == w[5];
      const int rm = w[6] & 255, rv = (w[6] >> 8) & 255;
      switch (w[3]) {
        case 0: f_mat1<0>(a, M, tok, rm, rv); break;
        case 1: f_mat1<1>(a, M, tok, rm, rv); break;
        case 2: f_mat1<2>(a, M, tok, rm, rv); break;
        default: f_mat1<3>(a, M, tok, rm, rv); break;
      }
    }
  } else if (type == OP_DIAG) {
    const bool tok = (tid & w[4]) == w[5];
    const int rm = w[6] & 255, rv = (w[6] >> 8) & 255;
    if (fl & F_SIGN) {
      f_sign(a, tok, rm, rv);
    } else {
      const int tpar = __popc(tid & w[3]);
      f_diag(a, M, tok, rm, rv, tpar, (w[6] >> 16) & 255);
    }
  } else if ((FEAT & FEAT_MAT2) && type == OP_MAT2) {
    const bool tok = (tid & w[4]) == w[5];
    const int rm = w[6] & 255, rv = (w[6] >> 8) & 255;
    switch (w[3]) {  // kh | kl << 4, kh > kl
      case 0x01: f_mat2<1, 0>(a, M, tok, rm, rv); break;
      case 0x02: f_mat2<2, 0>(a, M, tok, rm, rv); break;
      case 0x03: f_mat2<3, 0>(a, M, tok, rm, rv); break;
      case 0x12: f_mat2<2, 1>(a, M, tok, rm, rv); break;
      case 0x13: f_mat2<3, 1>(a, M, tok, rm, rv); break;
      default: f_mat2<3, 2>(a, M, tok, rm, rv); break;
    }
  } else if ((FEAT & FEAT_MATK) && type == OP_MATK) {
    const bool tok = (tid & w[4]) == w[5];
    switch (w[3]) {  // 4 = all register bits; 0..3 = k=3 excluding that bit
      case 0: f_matk<3, 0>(a, M, tok); break;
      case 1: f_matk<3, 1>(a, M, tok); break;
      case 2: f_matk<3, 2>(a, M, tok); break;
      case 3: f_matk<3, 3>(a, M, tok); break;
      default: f_matk<4, 0>(a, M, tok); break;
    }
  }
}

template <int FEAT>
__device__ __forceinline__ void bwd_op(float2 (&p)[NS], float2 (&l)[NS], const int* __restrict__ w,
                                       const float2* __restrict__ M, int tid, float (&g8)[8]) {
  float (&g)[4] = *reinterpret_cast<float (*)[4]>(&g8[0]);
  const int type = w[0] & 15;
  const int fl = w[0];
  const int P = (w[7] >= 0) ? ((w[0] >> 4) & 15) : 0;
  const bool tok = (tid & w[4]) == w[5];
  const int rm = w[6] & 255, rv = (w[6] >> 8) & 255;
  if (type == OP_MAT1) {
    if (fl & F_XPERM) {   // X is its own adjoint and carries no parameters
      switch (w[3]) {
        case 0: f_xperm<0>(p, tok, rm, rv); f_xperm<0>(l, tok, rm, rv); break;
        case 1: f_xperm<1>(p, tok, rm, rv); f_xperm<1>(l, tok, rm, rv); break;
        case 2: f_xperm<2>(p, tok, rm, rv); f_xperm<2>(l, tok, rm, rv); break;
        default: f_xperm<3>(p, tok, rm, rv); f_xperm<3>(l, tok, rm, rv); break;
      }
      return;
    }
    if (fl & F_MG) {
      if (fl & F_NOCTRL) {
        switch (w[3]) {
          case 0: b_mat1_mg<0, true>(p, l, M, tok, rm, rv, g8); break;
          case 1: b_mat1_mg<1, true>(p, l, M, tok, rm, rv, g8); break;
          case 2: b_mat1_mg<2, true>(p, l, M, tok, rm, rv, g8); break;
          default: b_mat1_mg<3, true>(p, l, M, tok, rm, rv, g8); break;
        }
      } else {
        switch (w[3]) {
          case 0: b_mat1_mg<0, false>(p, l, M, tok, rm, rv, g8); break;
          case 1: b_mat1_mg<1, false>(p, l, M, tok, rm, rv, g8); break;
          case 2: b_mat1_mg<2, false>(p, l, M, tok, rm, rv, g8); break;
          default: b_mat1_mg<3, false>(p, l, M, tok, rm, rv, g8); break;
        }
      }
      return;
    }
    switch (w[3]) {
      case 0: b_mat1<0>(p, l, M, P, tok, rm, rv, g); break;
      case 1: b_mat1<1>(p, l, M, P, tok, rm, rv, g); break;
      case 2: b_mat1<2>(p, l, M, P, tok, rm, rv, g); break;
      default: b_mat1<3>(p, l, M, P, tok, rm, rv, g); break;
    }
  } else if (type == OP_DIAG) {
    if (fl & F_SIGN) {
      f_sign(p, tok, rm, rv);
      f_sign(l, tok, rm, rv);
      return;
    }
    const int tpar = __popc(tid & w[3]);
    b_diag(p, l, M, P, tok, rm, rv, tpar, (w[6] >> 16) & 255, g);
  } else if ((FEAT & FEAT_MAT2) && type == OP_MAT2) {
    switch (w[3]) {
      case 0x01: b_mat2<1, 0>(p, l, M, P, tok, rm, rv, g); break;
      case 0x02: b_mat2<2, 0>(p, l, M, P, tok, rm, rv, g); break;
      case 0x03: b_mat2<3, 0>(p, l, M, P, tok, rm, rv, g); break;
      case 0x12: b_mat2<2, 1>(p, l, M, P, tok, rm, rv, g); break;
      case 0x13: b_mat2<3, 1>(p, l, M, P, tok, rm, rv, g); break;
      default: b_mat2<3, 2>(p, l, M, P, tok, rm, rv, g); break;
    }
  } else if ((FEAT & FEAT_MATK) && type == OP_MATK) {
    switch (w[3]) {
      case 0: b_matk1<3, 0>(p, M, tok); b_matk1<3, 0>(l, M, tok); break;
      case 1: b_matk1<3, 1>(p, M, tok); b_matk1<3, 1>(l, M, tok); break;
      case 2: b_matk1<3, 2>(p, M, tok); b_matk1<3, 2>(l, M, tok); break;
      case 3: b_matk1<3, 3>(p, M, tok); b_matk1<3, 3>(l, M, tok); break;
      default: b_matk1<4, 0>(p, M, tok); b_matk1<4, 0>(l, M, tok); break;
    }
  }
}

// ------------------------------------------------------------------ kernels
struct StageArgs {
  const int* prog;        // this stage's words (device)
  const float2* mats;     // matrix pool (device)
  int n;                  // qubits
  int init_zero;          // forward: 1 = start from |0..0> instead of reading psi_in
  int64_t n_groups;       // bsz << (n - m)
  int64_t groups_per_wg;
  float* part;            // adjoint: per-workgroup gradient partials [grid][rpw][n_gp]
  int64_t rpw;            // adjoint: rows per workgroup slot count of `part`
};

constexpr int FL_LOAD_SHIFT = 2, FL_STORE_SHIFT = 15, FL_MASK = 0x1FFF;

// In-group offsets of a layout: L[b] is the in-row position of local bit b
// (the header's positions from H_POS: thread bits, then register bits, of
// the canonical layout); `mask` the layout's register local bits (0: the
// canonical ones, local bits T..m-1).  Slot bit k is the k-th register bit
// in ascending order, thread-index bit i the i-th other local bit.
__device__ __forceinline__ void layout_offsets(const int* __restrict__ L, int m, int T, uint32_t mask, int tid,
                                               uint32_t& toff, uint32_t (&roff)[NS]) {
  if (mask == 0) mask = ((1u << m) - 1u) & ~((1u << T) - 1u);
  uint32_t rp[R];
  uint32_t mm = mask;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    rp[k] = (uint32_t)L[__builtin_ctz(mm)];
    mm &= mm - 1u;
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) o |= (uint32_t)((j >> k) & 1) << rp[k];
    roff[j] = o;
  }
  toff = 0;
  uint32_t tm = ((1u << m) - 1u) & ~mask;
  for (int i = 0; tm != 0u; ++i) {
    toff |= (uint32_t)((tid >> i) & 1) << L[__builtin_ctz(tm)];
    tm &= tm - 1u;
  }
}

__device__ __forceinline__ uint32_t group_offset(const int* __restrict__ nlpos, int nnl, uint64_t rest) {
  uint32_t o = 0;
  for (int i = 0; i < nnl; ++i) o |= (uint32_t)((rest >> i) & 1ull) << nlpos[i];
  return o;
}

// The program and matrix pool are separate __restrict__ const parameters so
// the compiler may fetch their (wave-uniform) words with scalar loads: the
// kernel's stores provably cannot clobber them.
template <int FEAT>
__global__ __launch_bounds__(512) void k_stage_fwd(float2* __restrict__ out, const float2* __restrict__ in,
                                                   const int* __restrict__ prog, const float2* __restrict__ mats,
                                                   StageArgs A) {
  extern __shared__ float2 lds[];
  const int* __restrict__ h = prog;
  const int n_ops = h[H_NOPS], T = h[H_T], nnl = h[H_NNL];
  const int* __restrict__ tpos = h + H_POS;
  const int* __restrict__ rpos = tpos + T;
  const int* __restrict__ nlpos = rpos + R;
  const int* __restrict__ ops = h + h[H_OFF_OPS];
  const int tid = threadIdx.x;
  const int m = h[H_M];
  uint32_t toff, roff[NS], soff, sroff[NS];     // load / store layout
  layout_offsets(tpos, m, T, (h[H_FLAGS] >> FL_LOAD_SHIFT) & FL_MASK, tid, toff, roff);
  layout_offsets(tpos, m, T, (h[H_FLAGS] >> FL_STORE_SHIFT) & FL_MASK, tid, soff, sroff);
  const int64_t g0 = (int64_t)blockIdx.x * A.groups_per_wg;
  int64_t g1 = g0 + A.groups_per_wg;
  if (g1 > A.n_groups) g1 = A.n_groups;
  const uint64_t rest_mask = (1ull << nnl) - 1ull;
  for (int64_t g = g0; g < g1; ++g) {
    const uint64_t row = (uint64_t)g >> nnl;
    const uint32_t goff = group_offset(nlpos, nnl, (uint64_t)g & rest_mask);
    const uint64_t gbase = (row << A.n) + goff;
    float2 a[NS];
    if (A.init_zero) {
#pragma unroll
      for (int j = 0; j < NS; ++j) a[j] = make_float2((goff + toff + roff[j]) == 0 ? 1.f : 0.f, 0.f);
    } else {
      const float2* src = in + gbase + toff;
#pragma unroll
      for (int j = 0; j < NS; ++j) a[j] = src[roff[j]];
    }
    for (int o = 0; o < n_ops; ++o) {
      const int* __restrict__ w = ops + 8 * o;
      if ((w[0] & 15) == OP_REMAP) {
        const int* rec = h + w[7];
        remap(a, lds, rec, rec + T + R, T, tid);
        continue;
      }
      const float2* M = mats + w[1] + ((w[0] & F_BATCHED) ? (int64_t)row * w[2] : 0);
      fwd_op<FEAT>(a, w, M, tid);
    }
    float2* dst = out + gbase + soff;
    asm volatile("" : "+v"(dst));
#pragma unroll
    for (int j = 0; j < NS; ++j) dst[sroff[j]] = a[j];
  }
}

// Gradient partials: each wave adds its per-op totals into its own LDS row
// (gacc[wave][n_gp], after the remap buffer; no atomics), and at every row
// change the workgroup writes the sum over waves, in wave order, to its slot
// of A.part: part[(blockIdx.x * rpw + row - first row) * n_gp + t].
// k_stagepart_reduce adds those up per row in workgroup order, so gradients
// are the same on every run.
template <int FEAT>
__global__ __launch_bounds__(512) void k_stage_adj(float2* __restrict__ psi, float2* __restrict__ lam,
                                                   const int* __restrict__ prog, const float2* __restrict__ mats,
                                                   StageArgs A) {
  extern __shared__ float2 lds[];
  const int* __restrict__ h = prog;
  const int n_ops = h[H_NOPS], T = h[H_T], nnl = h[H_NNL], n_gp = h[H_NGP];
  const int nwave = blockDim.x >= 64 ? (int)(blockDim.x >> 6) : 1;
  float* __restrict__ gacc = reinterpret_cast<float*>(lds + (1 << h[H_M]));
  const int* __restrict__ tpos = h + H_POS;
  const int* __restrict__ rpos = tpos + T;
  const int* __restrict__ nlpos = rpos + R;
  const int* __restrict__ ops = h + h[H_OFF_OPS];
  const int* __restrict__ gslot = h + h[H_OFF_GSLOT];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wwidth = blockDim.x < 64 ? (int)blockDim.x : 64;
  float* __restrict__ wacc = gacc + (tid >> 6) * n_gp;   // this wave's row
  for (int t = tid; t < nwave * n_gp; t += blockDim.x) gacc[t] = 0.f;
  // the program runs backwards: read in its store layout, write in its load layout
  uint32_t toff, roff[NS], soff, sroff[NS];
  layout_offsets(tpos, h[H_M], T, (h[H_FLAGS] >> FL_STORE_SHIFT) & FL_MASK, tid, toff, roff);
  layout_offsets(tpos, h[H_M], T, (h[H_FLAGS] >> FL_LOAD_SHIFT) & FL_MASK, tid, soff, sroff);
  const int64_t g0 = (int64_t)blockIdx.x * A.groups_per_wg;
  int64_t g1 = g0 + A.groups_per_wg;
  if (g1 > A.n_groups) g1 = A.n_groups;
  const uint64_t rest_mask = (1ull << nnl) - 1ull;
  const int64_t row0 = g0 >> nnl;
  float* __restrict__ wpart = A.part + (int64_t)blockIdx.x * A.rpw * n_gp;
  int64_t cur_row = -1;
  __syncthreads();
  for (int64_t g = g0; g < g1; ++g) {
    const uint64_t row = (uint64_t)g >> nnl;
    if (n_gp > 0 && (int64_t)row != cur_row) {
      if (cur_row >= 0) {
        __syncthreads();
        for (int t = tid; t < n_gp; t += blockDim.x) {
          float v = 0.f;
          for (int w = 0; w < nwave; ++w) {
            v += gacc[w * n_gp + t];
            gacc[w * n_gp + t] = 0.f;
          }
          wpart[(cur_row - row0) * n_gp + t] = v;
        }
        __syncthreads();
      }
      cur_row = (int64_t)row;
    }
    const uint32_t goff = group_offset(nlpos, nnl, (uint64_t)g & rest_mask);
    const uint64_t base = (row << A.n) + goff + toff;
    const uint64_t sbase = (row << A.n) + goff + soff;
    float2 p[NS], l[NS];
    {
      const float2* sp = psi + base;
#pragma unroll
      for (int j = 0; j < NS; ++j) p[j] = sp[roff[j]];
      const float2* sl = lam + base;
      asm volatile("" : "+v"(sl));
#pragma unroll
      for (int j = 0; j < NS; ++j) l[j] = sl[roff[j]];
    }
    for (int o = n_ops - 1; o >= 0; --o) {
      const int* __restrict__ w = ops + 8 * o;
      if ((w[0] & 15) == OP_REMAP) {
        const int* rec = h + w[7];
        remap(p, lds, rec + T + R, rec, T, tid);
        remap(l, lds, rec + T + R, rec, T, tid);
        continue;
      }
      const float2* M = mats + w[1] + ((w[0] & F_BATCHED) ? (int64_t)row * w[2] : 0);
      float gq[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bwd_op<FEAT>(p, l, w, M, tid, gq);
      if (w[7] >= 0) {
        const int P = (w[0] & F_MG) ? 4 : ((w[0] >> 4) & 15);
        if (wwidth == 64) {
          // halving exchange: every lane ends with one value's wave total
          if (P > 4) {
            const float c = wave_reduce_t<8>(gq, lane);
            if ((lane & 7) == 0) wacc[w[7] + (lane >> 3)] += c;
          } else if (P > 2) {
            const float v4[4] = {gq[0], gq[1], gq[2], gq[3]};
            const float c = wave_reduce_t<4>(v4, lane);
            if ((lane & 15) == 0 && (lane >> 4) < P) wacc[w[7] + (lane >> 4)] += c;
          } else if (P == 2) {
            const float v2[2] = {gq[0], gq[1]};
            const float c = wave_reduce_t<2>(v2, lane);
            if ((lane & 31) == 0) wacc[w[7] + (lane >> 5)] += c;
          } else {
            const float v1[1] = {gq[0]};
            const float c = wave_reduce_t<1>(v1, lane);
            if (lane == 0) wacc[w[7]] += c;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if (q < P) {
              const float s = wave_sum(gq[q], wwidth);
              if (lane == 0) wacc[w[7] + q] += s;
            }
          }
        }
      }
    }
    {
      float2* dp = psi + sbase;
      asm volatile("" : "+v"(dp));
#pragma unroll
      for (int j = 0; j < NS; ++j) dp[sroff[j]] = p[j];
      float2* dl = lam + sbase;
      asm volatile("" : "+v"(dl));
#pragma unroll
      for (int j = 0; j < NS; ++j) dl[sroff[j]] = l[j];
    }
  }
  if (n_gp > 0 && cur_row >= 0) {
    __syncthreads();
    for (int t = tid; t < n_gp; t += blockDim.x) {
      float v = 0.f;
      for (int w = 0; w < nwave; ++w) v += gacc[w * n_gp + t];
      wpart[(cur_row - row0) * n_gp + t] = v;
    }
  }
}

// grad[row * gstride + gslot[t]] += the partials of every workgroup that
// covered `row`, added in workgroup order (one thread per (row, t)).
// Workgroup w covers groups [w * gpw, (w + 1) * gpw) and wrote its partials of
// row r to part[(w * rpw + r - (w * gpw >> nnl)) * n_gp + t].
__global__ void k_stagepart_reduce(const float* __restrict__ part, const int* __restrict__ prog, int64_t bsz,
                                   int64_t gpw, int64_t grid, int64_t rpw, float* __restrict__ grad,
                                   int64_t gstride) {
  const int n_gp = prog[H_NGP], nnl = prog[H_NNL];
  const int* __restrict__ gslot = prog + prog[H_OFF_GSLOT];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bsz * n_gp) return;
  const int64_t row = i / n_gp;
  const int t = (int)(i - row * n_gp);
  const int64_t w0 = (row << nnl) / gpw;
  int64_t w1 = (((row + 1) << nnl) - 1) / gpw;
  if (w1 > grid - 1) w1 = grid - 1;
  float s = 0.f;
  for (int64_t w = w0; w <= w1; ++w) s += part[(w * rpw + row - ((w * gpw) >> nnl)) * n_gp + t];
  grad[row * gstride + gslot[t]] += s;
}

// ------------------------------------------------------------------ batch
// Many circuits of different widths in one launch: each item is one stage
// program applied to a range of groups of one circuit's state (bsz = 1).
// The workgroup has BATCH_THREADS threads; an item with T thread bits uses
// the first 2^T of them, the rest only join the LDS barriers.
constexpr int BATCH_THREADS = 256;

struct BatchItem {
  int64_t state_off;   // amplitude offset of the circuit's state
  int32_t prog_off;    // word offset of the stage program
  int32_t n;           // qubits of the circuit
  int32_t init_zero;   // first stage of the circuit: start from |0..0>
  int32_t pad;
  int64_t g0, g1;      // group range of this item
};

// One item: a stage program over groups [g0, g1) of one circuit's state.
__device__ __forceinline__ void batch_item(float2* __restrict__ states, const BatchItem& item,
                                           const int* __restrict__ progs, const float2* __restrict__ mats,
                                           float2* lds, int tid) {
  const int* __restrict__ h = progs + item.prog_off;
  const int n_ops = h[H_NOPS], T = h[H_T], nnl = h[H_NNL];
  const int* __restrict__ tpos = h + H_POS;
  const int* __restrict__ rpos = tpos + T;
  const int* __restrict__ nlpos = rpos + R;
  const int* __restrict__ ops = h + h[H_OFF_OPS];
  const bool active = tid < (1 << T);
  uint32_t toff, roff[NS], soff, sroff[NS];   // load / store layout
  layout_offsets(tpos, h[H_M], T, (h[H_FLAGS] >> FL_LOAD_SHIFT) & FL_MASK, tid, toff, roff);
  layout_offsets(tpos, h[H_M], T, (h[H_FLAGS] >> FL_STORE_SHIFT) & FL_MASK, tid, soff, sroff);
  const uint64_t rest_mask = (1ull << nnl) - 1ull;
  float2* st = states + item.state_off;
  for (int64_t g = item.g0; g < item.g1; ++g) {
    const uint32_t goff = group_offset(nlpos, nnl, (uint64_t)g & rest_mask);
    float2 a[NS];
    if (item.init_zero) {
#pragma unroll
      for (int j = 0; j < NS; ++j) a[j] = make_float2((goff + toff + roff[j]) == 0 ? 1.f : 0.f, 0.f);
    } else if (active) {
      const float2* src = st + (uint64_t)goff + toff;
#pragma unroll
      for (int j = 0; j < NS; ++j) a[j] = src[roff[j]];
    }
    for (int o = 0; o < n_ops; ++o) {
      const int* __restrict__ w = ops + 8 * o;
      if ((w[0] & 15) == OP_REMAP) {
        const int* rec = h + w[7];
        remap(a, lds, rec, rec + T + R, T, tid, active);
        continue;
      }
      fwd_op<FEAT_MAT2 | FEAT_MATK>(a, w, mats + w[1], tid);
    }
    if (active) {
      float2* dst = st + (uint64_t)goff + soff;
      asm volatile("" : "+v"(dst));
#pragma unroll
      for (int j = 0; j < NS; ++j) dst[sroff[j]] = a[j];
    }
  }
}

__global__ __launch_bounds__(BATCH_THREADS) void k_batch_fwd(float2* __restrict__ states,
                                                             const BatchItem* __restrict__ items, int n_items,
                                                             const int* __restrict__ progs,
                                                             const float2* __restrict__ mats) {
  extern __shared__ float2 lds[];
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    batch_item(states, items[it], progs, mats, lds, threadIdx.x);
    __syncthreads();
  }
}

// ---- one persistent launch for every level of every circuit ---------------
// Items are listed level by level; item i of circuit c needs the `need`
// items of c's earlier levels to be done.  Workgroups take items in list
// order from a global counter (ctrl[0]), so whenever a workgroup waits, every
// item it waits for is already held by a running workgroup, and the earliest
// unfinished item never waits: the queue drains with any grid size.  A
// finished item is published with a device-scope release (its stores, then
// done[c] += 1); a waiting workgroup polls done[c] with acquire loads, at
// most spin_limit times: past that it sets ctrl[1] (the results are then
// wrong and the host reports it) and goes on, as does every later wait once
// ctrl[1] is set, so every wave reaches the queue's end.
struct QueueItem {
  BatchItem b;
  int32_t circ;   // circuit index (done[] slot)
  int32_t need;   // items of this circuit in earlier levels
};

__global__ void k_queue_reset(int* __restrict__ ctrl, int* __restrict__ done, int n_circ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) ctrl[i] = 0;
  if (i < n_circ) done[i] = 0;
}

// Control flow is uniform for the compiler throughout: per-wave choices test
// the wave index taken through readfirstlane, never a per-lane value.  (With
// "if (tid == 0)" around the claim and the release, the compiler treats the
// loop as divergent and may move lane 0's part of the iteration boundary out
// of the loop that the other lanes keep running -- past their barrier -- so
// the loop never ends.)  Only wave 0 claims, waits and publishes:
//   claim    one atomic add on the queue counter;
//   wait     relaxed polls of done[c], then one device-scope acquire fence
//            (it invalidates this CU's L1 and this XCD's L2, which the
//            workgroup's other waves, ordered after it by the barrier, share);
//   publish  every wave waits for its stores (workgroup-scope release), the
//            barrier, then wave 0's device-scope release (L2 write-back) and
//            a relaxed add to done[c].
__global__ __launch_bounds__(BATCH_THREADS) void k_batch_queue(float2* __restrict__ states,
                                                               const QueueItem* __restrict__ items, int n_items,
                                                               const int* __restrict__ progs,
                                                               const float2* __restrict__ mats,
                                                               int* __restrict__ ctrl, int* __restrict__ done,
                                                               int64_t spin_limit) {
  extern __shared__ float2 lds[];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const bool wave0 = (__builtin_amdgcn_readfirstlane(tid) >> 6) == 0;   // wave-uniform
  const int one = tid == 0 ? 1 : 0;
  for (;;) {
    if (wave0) {
      const int v = atomicAdd(&ctrl[0], one);
      s_item = __builtin_amdgcn_readfirstlane(v);
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) break;
    const QueueItem q = items[it];
    if (wave0 && q.need > 0) {
      int64_t spins = 0;
      for (;;) {
        const int d = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&done[q.circ], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (d >= q.need) break;
        if (++spins > spin_limit) {
          atomicExch(&ctrl[1], 1);
          break;
        }
        // once any wait has run out, the others stop waiting too
        if ((spins & 255) == 0 &&
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(&ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
          break;
        __builtin_amdgcn_s_sleep(8);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();                           // the wait is over, the previous item's LDS use too
    batch_item(states, q.b, progs, mats, lds, tid);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // this wave's stores are done
    __syncthreads();                           // ... in every wave; s_item has been read
    if (wave0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(&done[q.circ], one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// <Z-string> per circuit: sum_j |psi_j|^2 (-1)^{popcount(j & zmask)}; one
// partial per (circuit, chunk of 2^14 amplitudes), then a per-circuit sum.
struct ZCirc {
  int64_t state_off;
  int32_t n;
  uint32_t zmask;
  int64_t part_off;    // first partial slot of this circuit
};

constexpr int ZCHUNK_LG = 14;

__global__ __launch_bounds__(256) void k_batch_zpart(const float2* __restrict__ states,
                                                     const ZCirc* __restrict__ circ, const int64_t* __restrict__ chunk_circ,
                                                     int64_t n_chunks, float* __restrict__ part) {
  __shared__ float red[4];
  for (int64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const int c = (int)chunk_circ[ch];
    const ZCirc z = circ[c];
    const int64_t local = ch - z.part_off;
    const int64_t first = local << ZCHUNK_LG;
    const int64_t cnt = ((int64_t)1 << z.n) - first < ((int64_t)1 << ZCHUNK_LG) ? ((int64_t)1 << z.n) - first
                                                                               : ((int64_t)1 << ZCHUNK_LG);
    float acc = 0.f;
    for (int64_t i = threadIdx.x; i < cnt; i += blockDim.x) {
      const uint32_t j = (uint32_t)(first + i);
      const float2 v = states[z.state_off + first + i];
      const float p = v.x * v.x + v.y * v.y;
      acc += (__popc(j & z.zmask) & 1) ? -p : p;
    }
    acc = wave_sum(acc, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[ch] = red[0] + red[1] + red[2] + red[3];
  }
}

__global__ void k_batch_zsum(const ZCirc* __restrict__ circ, int n_circ, const float* __restrict__ part,
                             float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_circ) return;
  const ZCirc z = circ[c];
  const int64_t nch = (((int64_t)1 << z.n) + ((int64_t)1 << ZCHUNK_LG) - 1) >> ZCHUNK_LG;
  double s = 0.0;
  for (int64_t i = 0; i < nch; ++i) s += part[z.part_off + i];
  out[c] = (float)s;
}

// Hermitian-generator sums -> parameter gradients: for each parameter p of an
// op accumulated with F_MG, H_p = i D_p U^dag = a I + b X + c Y + d Z and
// grad[row, slot_p] += a S_I + b S_X + c S_Y + d S_Z.
struct MgItem {
  int32_t mg_slot;     // first of the 4 sums (S_I, S_X, S_Y, S_Z) in the grad row
  int32_t P;
  int32_t slot[4];     // parameter slots
  int64_t u_off;       // pool offset (complex) of U for row 0; D_p follows at u_off + 4 (p + 1)
  int64_t bstride;     // pool stride between rows (0 = shared)
};

__global__ void k_mg_to_grad(float* __restrict__ grad, int64_t gstride, int64_t bsz, const MgItem* __restrict__ items,
                             int n_items, const float2* __restrict__ mats) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= bsz * n_items) return;
  const int64_t row = t / n_items;
  const MgItem it = items[t % n_items];
  float* gr = grad + row * gstride;
  const float sI = gr[it.mg_slot], sX = gr[it.mg_slot + 1], sY = gr[it.mg_slot + 2], sZ = gr[it.mg_slot + 3];
  const float2* U = mats + it.u_off + row * it.bstride;
  for (int q = 0; q < it.P && q < 4; ++q) {
    const float2* D = U + 4 * (q + 1);
    // G = D U^dag: G[r][c] = sum_k D[r][k] conj(U[c][k]);  H = i G
    float2 G[4];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) G[2 * r + c] = cfmac(U[2 * c + 1], D[2 * r + 1], cmulc(U[2 * c], D[2 * r]));
    // H = i G: H[r][c] = (-G.y, G.x).  a = Re(H00 + H11)/2, d = Re(H00 - H11)/2,
    // b = Re(H01 + H10)/2, c = Im(H10 - H01)/2
    const float h00 = -G[0].y, h11 = -G[3].y;
    const float ca = 0.5f * (h00 + h11), cd = 0.5f * (h00 - h11);
    const float cb = 0.5f * (-G[1].y - G[2].y);
    const float cy = 0.5f * (G[2].x - G[1].x);
    gr[it.slot[q]] += ca * sI + cb * sX + cy * sY + cd * sZ;
  }
}

}  // namespace tqsv

using namespace tqsv;

TQSV_API int tqsv_mg_item_bytes(void) { return (int)sizeof(MgItem); }

// After tqsv_stage_adjoint: fold the matrix gradients of ops flagged F_MG into
// their parameter slots.  items_dev: MgItem array.
TQSV_API int tqsv_mg_to_grad(hipStream_t s, float* grad, int64_t grad_stride, int64_t bsz, const void* items_dev,
                             int n_items, const void* mats) {
  if (!grad || !items_dev || !mats || bsz < 1 || n_items < 0) { set_error("tqsv_mg_to_grad: bad arguments"); return E_ARG; }
  if (n_items == 0) return OK;
  const int64_t tot = bsz * n_items;
  k_mg_to_grad<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(grad, grad_stride, bsz, (const MgItem*)items_dev, n_items,
                                                             (const float2*)mats);
  return check_launch("tqsv_mg_to_grad");
}

// Launch shape of a stage program (also used by the run-time compiled stages).
int stage_launch_shape(const int* prog_host, int n, int64_t bsz, int grid_cap, int& T, int64_t& n_groups,
                              int& grid, int64_t& gpw, size_t& shm) {
  const int m = prog_host[H_M];
  T = prog_host[H_T];
  if (m < R || m > 13 || T != m - R || m > n || prog_host[H_NNL] != n - m) {
    set_error("stage header inconsistent: m=%d T=%d n=%d nnl=%d", m, T, n, prog_host[H_NNL]);
    return E_ARG;
  }
  if (prog_host[H_NGP] > MAX_GP) {
    set_error("stage has %d gradient parameters (max %d)", prog_host[H_NGP], MAX_GP);
    return E_LIMIT;
  }
  n_groups = bsz << (n - m);
  int64_t cap = grid_cap > 0 ? grid_cap : (int64_t)num_cu() * (T >= 8 ? 4 : 8);
  int64_t g = n_groups < cap ? n_groups : cap;
  gpw = (n_groups + g - 1) / g;
  grid = (int)((n_groups + gpw - 1) / gpw);
  shm = sizeof(float2) << m;
  return OK;
}

// prog_host: the same stage words as prog_dev, readable on the host (only the
// header is read, to size the launch).  mats: matrix pool.  state_in may
// equal state_out (in-place) or be null with init_zero = 1.
TQSV_API int tqsv_stage_forward(hipStream_t s, void* state_out, const void* state_in, int64_t bsz, int n,
                                const int* prog_host, const int* prog_dev, const void* mats, int init_zero,
                                int grid_cap) {
  if (!state_out || !prog_host || !prog_dev || !mats || (!state_in && !init_zero)) {
    set_error("tqsv_stage_forward: null pointer argument");
    return E_ARG;
  }
  if (bsz < 1 || n < R || n > 32) { set_error("tqsv_stage_forward: bad bsz/n (%lld, %d)", (long long)bsz, n); return E_ARG; }
  int T, grid;
  int64_t n_groups, gpw;
  size_t shm;
  int rc = stage_launch_shape(prog_host, n, bsz, grid_cap, T, n_groups, grid, gpw, shm);
  if (rc) return rc;
  StageArgs A{prog_dev, (const float2*)mats, n, init_zero, n_groups, gpw, nullptr, 0};
  const int feat = prog_host[H_FLAGS] & 3;
  if (feat == 0)
    k_stage_fwd<0><<<grid, 1 << T, shm, s>>>((float2*)state_out, (const float2*)state_in, prog_dev,
                                                      (const float2*)mats, A);
  else if (feat == FEAT_MAT2)
    k_stage_fwd<FEAT_MAT2><<<grid, 1 << T, shm, s>>>((float2*)state_out, (const float2*)state_in, prog_dev,
                                                      (const float2*)mats, A);
  else
    k_stage_fwd<FEAT_MAT2 | FEAT_MATK><<<grid, 1 << T, shm, s>>>((float2*)state_out, (const float2*)state_in, prog_dev,
                                                      (const float2*)mats, A);
  return check_launch("tqsv_stage_forward");
}

// psi / lam are updated in place; grad ([bsz, grad_stride] float) is
// accumulated into (the caller zeroes it once per backward pass).
// Rows a workgroup of gpw consecutive groups can touch (2^nnl groups per row).
int64_t stage_rows_per_wg(const int* prog_host, int64_t gpw) {
  const int64_t gpr = 1ll << prog_host[H_NNL];
  return (gpw - 1 + gpr - 1) / gpr + 1;
}

// Floats of partial buffer an adjoint launch of this stage needs.
int64_t stage_adjoint_work(const int* prog_host, int64_t gpw, int grid) {
  return (int64_t)grid * stage_rows_per_wg(prog_host, gpw) * prog_host[H_NGP];
}

// After an adjoint stage launch: add its partials into grad in a fixed order.
int stage_adjoint_reduce(hipStream_t s, const int* prog_host, const int* prog_dev, int64_t bsz, int64_t gpw,
                         int grid, const float* work, float* grad, int64_t grad_stride) {
  const int ngp = prog_host[H_NGP];
  if (ngp == 0) return OK;
  const int64_t total = bsz * ngp;
  const int threads = 256;
  k_stagepart_reduce<<<(unsigned)((total + threads - 1) / threads), threads, 0, s>>>(
      work, prog_dev, bsz, gpw, grid, stage_rows_per_wg(prog_host, gpw), grad, grad_stride);
  return OK;
}

TQSV_API int tqsv_stage_adjoint(hipStream_t s, void* psi, void* lam, int64_t bsz, int n, const int* prog_host,
                                const int* prog_dev, const void* mats, float* grad, int64_t grad_stride,
                                float* work, int64_t work_floats, int grid_cap) {
  if (!psi || !lam || !prog_host || !prog_dev || !mats) {
    set_error("tqsv_stage_adjoint: null pointer argument");
    return E_ARG;
  }
  if (bsz < 1 || n < R || n > 32) { set_error("tqsv_stage_adjoint: bad bsz/n"); return E_ARG; }
  if (prog_host[H_NGP] > 0 && (!grad || !work)) {
    set_error("tqsv_stage_adjoint: grad and partial buffers required");
    return E_ARG;
  }
  int T, grid;
  int64_t n_groups, gpw;
  size_t shm;
  int rc = stage_launch_shape(prog_host, n, bsz, grid_cap, T, n_groups, grid, gpw, shm);
  if (rc) return rc;
  const int64_t need = stage_adjoint_work(prog_host, gpw, grid);
  if (need > work_floats) {
    set_error("tqsv_stage_adjoint: partial buffer of %lld floats, the stage needs %lld "
              "(tqsv_stage_adjoint_work_floats)", (long long)work_floats, (long long)need);
    return E_ARG;
  }
  const int nwave = T >= 6 ? 1 << (T - 6) : 1;
  shm += sizeof(float) * (size_t)nwave * prog_host[H_NGP];
  StageArgs A{prog_dev, (const float2*)mats, n, 0, n_groups, gpw, work, stage_rows_per_wg(prog_host, gpw)};
  const int feat = prog_host[H_FLAGS] & 3;
  if (feat == 0)
    k_stage_adj<0><<<grid, 1 << T, shm, s>>>((float2*)psi, (float2*)lam, prog_dev, (const float2*)mats, A);
  else if (feat == FEAT_MAT2)
    k_stage_adj<FEAT_MAT2><<<grid, 1 << T, shm, s>>>((float2*)psi, (float2*)lam, prog_dev, (const float2*)mats, A);
  else
    k_stage_adj<FEAT_MAT2 | FEAT_MATK><<<grid, 1 << T, shm, s>>>((float2*)psi, (float2*)lam, prog_dev,
                                                                (const float2*)mats, A);
  if ((rc = check_launch("tqsv_stage_adjoint"))) return rc;
  stage_adjoint_reduce(s, prog_host, prog_dev, bsz, gpw, grid, work, grad, grad_stride);
  return check_launch("tqsv_stage_adjoint");
}

// Floats of the partial buffer tqsv_stage_adjoint / tqsv_jit_stage_adjoint
// need for this stage (same launch shape arguments).
TQSV_API int64_t tqsv_stage_adjoint_work_floats(const int* prog_host, int n, int64_t bsz, int grid_cap) {
  int T, grid;
  int64_t n_groups, gpw;
  size_t shm;
  if (!prog_host || stage_launch_shape(prog_host, n, bsz, grid_cap, T, n_groups, grid, gpw, shm)) return -1;
  return stage_adjoint_work(prog_host, gpw, grid);
}

// One launch over a list of (circuit, stage, group range) items; every item's
// stage program must have T <= 8 (m <= 12).  items_dev: BatchItem array;
// progs_dev: concatenated stage programs; mats: matrix pool.
TQSV_API int tqsv_batch_forward(hipStream_t s, void* states, const void* items_dev, int n_items,
                                const int* progs_dev, const void* mats, int max_m, int grid) {
  if (!states || !items_dev || !progs_dev || !mats || n_items < 0) {
    set_error("tqsv_batch_forward: null pointer argument");
    return E_ARG;
  }
  if (max_m < R || max_m > 12) { set_error("tqsv_batch_forward: max_m must be in [4, 12]"); return E_LIMIT; }
  if (n_items == 0) return OK;
  if (grid <= 0) grid = n_items < num_cu() * 4 ? n_items : num_cu() * 4;
  const size_t shm = sizeof(float2) << max_m;
  k_batch_fwd<<<grid, BATCH_THREADS, shm, s>>>((float2*)states, (const BatchItem*)items_dev, n_items,
                                               progs_dev, (const float2*)mats);
  return check_launch("tqsv_batch_forward");
}

// Every level of every circuit in one persistent launch (k_batch_queue).
// items_dev: QueueItem array, level-major; ctrl_dev: 2 ints (queue counter,
// error flag), done_dev: n_circ ints, both reset here by a small kernel (no
// memset, so the pair can be captured in a HIP graph).  After the stream
// reaches this point ctrl_dev[1] != 0 means a bounded wait ran out.
TQSV_API int tqsv_batch_queue(hipStream_t s, void* states, const void* items_dev, int n_items,
                              const int* progs_dev, const void* mats, int max_m, int* ctrl_dev, int* done_dev,
                              int n_circ, int64_t spin_limit, int grid) {
  if (!states || !items_dev || !progs_dev || !mats || !ctrl_dev || !done_dev || n_items < 0 || n_circ < 1) {
    set_error("tqsv_batch_queue: null pointer or bad count");
    return E_ARG;
  }
  if (max_m < R || max_m > 12) { set_error("tqsv_batch_queue: max_m must be in [4, 12]"); return E_LIMIT; }
  if (spin_limit < 1) { set_error("tqsv_batch_queue: spin_limit must be positive"); return E_ARG; }
  const int nr = n_circ > 2 ? n_circ : 2;
  k_queue_reset<<<(nr + 255) / 256, 256, 0, s>>>(ctrl_dev, done_dev, n_circ);
  int rc = check_launch("tqsv_batch_queue/reset");
  if (rc || n_items == 0) return rc;
  if (grid <= 0) grid = num_cu() * 4;
  if (grid > n_items) grid = n_items;
  const size_t shm = sizeof(float2) << max_m;
  k_batch_queue<<<grid, BATCH_THREADS, shm, s>>>((float2*)states, (const QueueItem*)items_dev, n_items, progs_dev,
                                                 (const float2*)mats, ctrl_dev, done_dev, spin_limit);
  return check_launch("tqsv_batch_queue");
}

TQSV_API int tqsv_batch_queue_item_bytes(void) { return (int)sizeof(QueueItem); }
TQSV_API int tqsv_batch_item_bytes(void) { return (int)sizeof(BatchItem); }
TQSV_API int tqsv_batch_zcirc_bytes(void) { return (int)sizeof(ZCirc); }

// Z-string expectation per circuit (circ_dev: ZCirc array, chunk_circ_dev:
// circuit index of every 2^14-amplitude chunk, part: n_chunks floats).
TQSV_API int tqsv_batch_zexpval(hipStream_t s, const void* states, const void* circ_dev, int n_circ,
                                const int64_t* chunk_circ_dev, int64_t n_chunks, float* part, float* out) {
  if (!states || !circ_dev || !chunk_circ_dev || !part || !out) { set_error("tqsv_batch_zexpval: null pointer"); return E_ARG; }
  if (n_circ <= 0 || n_chunks <= 0) return OK;
  const int64_t g = n_chunks < (int64_t)num_cu() * 8 ? n_chunks : (int64_t)num_cu() * 8;
  k_batch_zpart<<<(unsigned)g, 256, 0, s>>>((const float2*)states, (const ZCirc*)circ_dev, chunk_circ_dev, n_chunks,
                                            part);
  int rc = check_launch("tqsv_batch_zexpval/part");
  if (rc) return rc;
  k_batch_zsum<<<(n_circ + 127) / 128, 128, 0, s>>>((const ZCirc*)circ_dev, n_circ, part, out);
  return check_launch("tqsv_batch_zexpval/sum");
}
