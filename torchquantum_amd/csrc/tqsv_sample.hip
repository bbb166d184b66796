// Computational-basis shot sampling: replaces the reference's measure()
// (torchquantum/measurement/measurements.py:37-74), which copies |psi| to the
// host and draws n_shots outcomes per batch row with random.choices over all
// 2^n bitstrings.  Here the state never leaves HBM:
//   1. k_block_probs: p_j = |a_j|^2 summed over blocks of SB = 4096
//      amplitudes (float64 partials), one workgroup per block;
//   2. k_block_scan: per row, an inclusive prefix of the block sums;
//   3. k_draw: one wavefront per shot draws u in [0, 1) from a counter-based
//      generator (splitmix64 of seed, row, shot), finds the block by binary
//      search over the prefix, then the amplitude inside the block with a
//      64-lane prefix over the block's probabilities.
// Only indices with p_j > 0 can be drawn.  The output is the flat in-row
// index (wire 0 = most significant bit, the reference's bitstring order).
#include "tqsv_common.h"

namespace tqsv {

constexpr int SB = 4096;            // amplitudes per block
constexpr int SB_THREADS = 256;     // block-sum workgroup (16 amplitudes per thread)

__device__ __forceinline__ float prob(const float2 a) { return a.x * a.x + a.y * a.y; }

// sums[row * nb + b] = sum_{j in block b} |a_j|^2 ; grid = (nb, bsz)
__global__ __launch_bounds__(SB_THREADS) void k_block_probs(const float2* __restrict__ st, int n, int64_t nb,
                                                            double* __restrict__ sums) {
  const uint64_t row = blockIdx.y;
  const uint64_t b = blockIdx.x;
  const uint64_t N = 1ull << n;
  const uint64_t base = (row << n) + b * SB;
  const uint64_t end = (b * SB + SB < N) ? SB : N - b * SB;
  double acc = 0.0;
  for (uint64_t j = threadIdx.x; j < end; j += SB_THREADS) acc += (double)prob(st[base + j]);
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  __shared__ double part[SB_THREADS / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < SB_THREADS / 64; ++w) s += part[w];
    sums[row * nb + b] = s;
  }
}

// inclusive prefix of sums over each row, in place; one workgroup per row
__global__ __launch_bounds__(256) void k_block_scan(double* __restrict__ sums, int64_t nb) {
  double* p = sums + (uint64_t)blockIdx.x * nb;
  const int64_t per = (nb + 255) / 256;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  double s = 0.0;
  for (int64_t i = lo; i < hi; ++i) s += p[i];
  __shared__ double tot[256];
  tot[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int t = 0; t < 256; ++t) {
      const double v = tot[t];
      tot[t] = run;
      run += v;
    }
  }
  __syncthreads();
  double run = tot[threadIdx.x];
  for (int64_t i = lo; i < hi; ++i) {
    run += p[i];
    p[i] = run;
  }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// one wavefront per (row, shot); blockDim = 256 (4 shots per workgroup)
__global__ __launch_bounds__(256) void k_draw(const float2* __restrict__ st, int n, int64_t nb,
                                              const double* __restrict__ prefix, int64_t n_shots, uint64_t seed,
                                              int64_t bsz, int64_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t wid = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid >= (uint64_t)bsz * (uint64_t)n_shots) return;  // whole wavefronts leave together
  const uint64_t row = wid / (uint64_t)n_shots;
  const double* P = prefix + row * nb;
  const double total = P[nb - 1];
  const uint64_t r = splitmix64(seed ^ splitmix64(wid));
  const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
  double t = u * total;
  // first block with prefix > t (binary search, every lane the same)
  int64_t lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (P[mid] > t) hi = mid; else lo = mid + 1;
  }
  const int64_t b = lo;
  t -= (b > 0 ? P[b - 1] : 0.0);
  const uint64_t N = 1ull << n;
  const uint64_t bstart = (uint64_t)b * SB;
  const uint64_t blen = (bstart + SB < N) ? SB : N - bstart;
  const uint64_t per = (blen + 63) / 64;
  const uint64_t lbeg = lane * per < blen ? lane * per : blen;
  const uint64_t lend = lbeg + per < blen ? lbeg + per : blen;
  const float2* a = st + (row << n) + bstart;
  double s = 0.0;
  for (uint64_t j = lbeg; j < lend; ++j) s += (double)prob(a[j]);
  // inclusive prefix over lanes
  double inc = s;
  for (int o = 1; o < 64; o <<= 1) {
    const double v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  const double exc = inc - s;
  const uint64_t hit = __ballot(inc > t && s > 0.0);
  const uint64_t nz = __ballot(s > 0.0);
  int sel;
  bool exact;
  if (hit) {
    sel = __builtin_ctzll(hit);
    exact = true;
  } else {  // rounding pushed t past the block's recomputed total: last lane with mass
    sel = 63 - __builtin_clzll(nz ? nz : 1ull);
    exact = false;
  }
  if (lane == sel) {
    // the element where the running sum passes t; without an exact hit, the
    // lane's last element with nonzero probability
    uint64_t pick = lbeg;
    double acc = exc;
    for (uint64_t j = lbeg; j < lend; ++j) {
      const double p = (double)prob(a[j]);
      if (p > 0.0) {
        pick = j;
        acc += p;
        if (exact && acc > t) break;
      }
    }
    out[wid] = (int64_t)(bstart + pick);
  }
}

}  // namespace tqsv

using namespace tqsv;

TQSV_API int64_t tqsv_sample_workspace_bytes(int64_t bsz, int n) {
  if (bsz < 1 || n < 0 || n > 32) return -1;
  const int64_t N = 1ll << n;
  const int64_t nb = (N + SB - 1) / SB;
  return bsz * nb * (int64_t)sizeof(double);
}

TQSV_API int tqsv_sample(hipStream_t s, const void* state, int64_t bsz, int n, int64_t n_shots, uint64_t seed,
                         int64_t* out, void* workspace, int64_t ws_bytes) {
  if (!state || !out || !workspace) { set_error("tqsv_sample: null pointer"); return E_ARG; }
  if (bsz < 1 || bsz > 65535 || n < 0 || n > 32) { set_error("tqsv_sample: bad bsz/n (bsz <= 65535, n <= 32)"); return E_ARG; }
  if (n_shots < 1) { set_error("tqsv_sample: n_shots must be >= 1"); return E_ARG; }
  const int64_t need = tqsv_sample_workspace_bytes(bsz, n);
  if (ws_bytes < need) { set_error("tqsv_sample: workspace %lld bytes < %lld", (long long)ws_bytes, (long long)need); return E_ARG; }
  const int64_t N = 1ll << n;
  const int64_t nb = (N + SB - 1) / SB;
  const uint64_t waves = (uint64_t)bsz * (uint64_t)n_shots;
  if (waves > (1ull << 26)) { set_error("tqsv_sample: bsz * n_shots exceeds 2^26"); return E_LIMIT; }
  double* sums = (double*)workspace;
  k_block_probs<<<dim3((unsigned)nb, (unsigned)bsz), SB_THREADS, 0, s>>>((const float2*)state, n, nb, sums);
  k_block_scan<<<(unsigned)bsz, 256, 0, s>>>(sums, nb);
  k_draw<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>((const float2*)state, n, nb, sums, n_shots, seed, bsz, out);
  return check_launch("tqsv_sample");
}
