// Kernels for the learned fusion-plan cost model (a small MLP): forward
// GEMM + bias + ReLU, the backward products, a column sum, the MSE loss
// gradient and an Adam step.  fp32 throughout; the network is tiny (batch of
// plan feature rows x 64 hidden units), so a 64x64 LDS-tiled VALU GEMM is
// the right tool -- these launches are latency-, not throughput-bound.
// (The reference has no learned component; this supports the planner.)
#include "tqsv_common.h"

namespace tqsv {

constexpr int GT = 32;   // tile edge

// C[M,N] = alpha * op(A) * op(B) + beta * C (+ bias[N]) (-> ReLU)
//   op(A) = A [M,K] (lda) or A^T with A stored [K,M]
//   op(B) = B [K,N] (ldb) or B^T with B stored [N,K]
__global__ __launch_bounds__(GT * 8) void k_gemm(int M, int N, int K, float alpha, const float* __restrict__ A,
                                                 int lda, int ta, const float* __restrict__ B, int ldb, int tb,
                                                 float beta, float* __restrict__ C, int ldc,
                                                 const float* __restrict__ bias, int relu) {
  __shared__ float As[GT][GT + 1];
  __shared__ float Bs[GT][GT + 1];
  const int tx = threadIdx.x % GT;      // column within tile
  const int ty = threadIdx.x / GT;      // 0..7, each thread computes 4 rows
  const int row0 = blockIdx.y * GT;
  const int col0 = blockIdx.x * GT;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += GT) {
    for (int e = threadIdx.x; e < GT * GT; e += blockDim.x) {
      const int r = e / GT, c = e % GT;
      const int gr = row0 + r, gk = k0 + c;
      float va = 0.f;
      if (gr < M && gk < K) va = ta ? A[(size_t)gk * lda + gr] : A[(size_t)gr * lda + gk];
      As[r][c] = va;
      const int gk2 = k0 + r, gc = col0 + c;
      float vb = 0.f;
      if (gk2 < K && gc < N) vb = tb ? B[(size_t)gc * ldb + gk2] : B[(size_t)gk2 * ldb + gc];
      Bs[r][c] = vb;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < GT; ++kk) {
      const float b = Bs[kk][tx];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = fmaf(As[ty + 8 * q][kk], b, acc[q]);
    }
    __syncthreads();
  }
  const int c = col0 + tx;
  if (c >= N) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = row0 + ty + 8 * q;
    if (r >= M) continue;
    float v = alpha * acc[q];
    if (beta != 0.f) v += beta * C[(size_t)r * ldc + c];
    if (bias) v += bias[c];
    if (relu) v = v > 0.f ? v : 0.f;
    C[(size_t)r * ldc + c] = v;
  }
}

// dZ = dY * (Y > 0)  (ReLU backward, in place on dY)
__global__ void k_relu_bwd(float* __restrict__ dY, const float* __restrict__ Y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && !(Y[i] > 0.f)) dY[i] = 0.f;
}

// out[n] = sum over rows of X[M, N]
__global__ void k_colsum(const float* __restrict__ X, int M, int N, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int r = 0; r < M; ++r) s += X[(size_t)r * N + c];
  out[c] = s;
}

// grad = 2 (pred - y) / M ; loss_out[0] = mean (pred - y)^2 (single block)
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ y, int M, float* __restrict__ grad,
                      float* __restrict__ loss_out) {
  __shared__ float red[4];
  float s = 0.f;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    const float d = pred[i] - y[i];
    grad[i] = 2.f * d / (float)M;
    s += d * d;
  }
  s = wave_sum(s, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    loss_out[0] = t / (float)M;
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace tqsv

using namespace tqsv;

// C = alpha op(A) op(B) + beta C (+ bias) (ReLU if relu); row-major.
TQSV_API int tqsv_gemm(hipStream_t s, int M, int N, int K, float alpha, const float* A, int lda, int ta,
                       const float* B, int ldb, int tb, float beta, float* C, int ldc, const float* bias, int relu) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) { set_error("tqsv_gemm: bad arguments"); return E_ARG; }
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  k_gemm<<<grid, GT * 8, 0, s>>>(M, N, K, alpha, A, lda, ta, B, ldb, tb, beta, C, ldc, bias, relu);
  return check_launch("tqsv_gemm");
}

TQSV_API int tqsv_relu_backward(hipStream_t s, float* dY, const float* Y, int64_t n) {
  if (!dY || !Y || n < 0) { set_error("tqsv_relu_backward: bad arguments"); return E_ARG; }
  if (n == 0) return OK;
  k_relu_bwd<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dY, Y, n);
  return check_launch("tqsv_relu_backward");
}

TQSV_API int tqsv_colsum(hipStream_t s, const float* X, int M, int N, float* out) {
  if (!X || !out || M <= 0 || N <= 0) { set_error("tqsv_colsum: bad arguments"); return E_ARG; }
  k_colsum<<<(N + 127) / 128, 128, 0, s>>>(X, M, N, out);
  return check_launch("tqsv_colsum");
}

TQSV_API int tqsv_mse_grad(hipStream_t s, const float* pred, const float* y, int M, float* grad, float* loss) {
  if (!pred || !y || !grad || !loss || M <= 0) { set_error("tqsv_mse_grad: bad arguments"); return E_ARG; }
  k_mse<<<1, 256, 0, s>>>(pred, y, M, grad, loss);
  return check_launch("tqsv_mse_grad");
}

// Adam with bias correction factors c1 = 1 - b1^t, c2 = 1 - b2^t (host computed).
TQSV_API int tqsv_adam_step(hipStream_t s, float* p, const float* g, float* m, float* v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  if (!p || !g || !m || !v || n < 0) { set_error("tqsv_adam_step: bad arguments"); return E_ARG; }
  if (n == 0) return OK;
  k_adam<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return check_launch("tqsv_adam_step");
}
