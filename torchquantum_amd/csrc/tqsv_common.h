// Shared device helpers for the tqsv statevector kernels (gfx950 / CDNA4).
//
// State layout: a batch of rows, row b holds 2^n complex64 amplitudes
// (float2: x = real, y = imag) at [b << n, (b + 1) << n).  Wire w of the
// reference API is bit position (n - 1 - w) of the in-row index
// (torchquantum/device/devices.py:64-71); every kernel here speaks bit
// positions, the Python layer converts wires.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TQSV_API extern "C" __attribute__((visibility("default")))

namespace tqsv {

// ---- status codes (negative = error), mirrored in tqsv.h -------------------
enum Status : int {
  OK = 0,
  E_ARG = -1,       // invalid argument (shape, wire, size)
  E_LAUNCH = -2,    // HIP launch / runtime error
  E_LIMIT = -3,     // request exceeds a documented kernel limit
};

void set_error(const char* fmt, ...);
int check_launch(const char* what);

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
  return acc;
}
// acc + conj(a) * b
__device__ __forceinline__ float2 cfmac(float2 a, float2 b, float2 acc) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
  return acc;
}
// Re(conj(a) * b)
__device__ __forceinline__ float redot(float2 a, float2 b) {
  return fmaf(a.x, b.x, a.y * b.y);
}
__device__ __forceinline__ float2 csel(bool c, float2 a, float2 b) {
  return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}

// Launch shape of the streaming kernels: a full grid in which workgroup w owns
// the 256 * reps consecutive items [w * 256 * reps, (w + 1) * 256 * reps) and
// retires.  reps exceeds 1 only when the grid would pass 2^22 workgroups.  On
// an 8 GiB state an in-place pass launched this way moved 5.5-6.7 TB/s; the
// same pass as a grid-stride loop over 16 workgroups per CU moved 3.7-5.8 TB/s.
struct Tiles {
  unsigned grid;
  int reps;
};

static inline Tiles tiles_for(uint64_t items) {
  int reps = 1;
  uint64_t g = (items + 255) / 256;
  while (g > (1ull << 22)) {
    reps *= 2;
    g = (items + 256ull * reps - 1) / (256ull * reps);
  }
  return {(unsigned)(g ? g : 1), reps};
}

// This thread's first item; its r-th item is tile_first(reps) + 256 * r.
// Kernels using it are launched with 256 threads per workgroup.
__device__ __forceinline__ uint64_t tile_first(int reps) {
  return (uint64_t)blockIdx.x * 256u * (uint64_t)reps + threadIdx.x;
}

// 16-byte state accesses; NT marks them non-temporal (states larger than the
// 256 MB Infinity Cache, which would only evict each other).
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

// The state streams with non-temporal hints once it exceeds the Infinity Cache.
static inline bool stream_nt(int64_t bsz, int n) {
  return ((uint64_t)bsz << n) * 8ull > (256ull << 20);
}

// Insert zero bits at ascending positions pos[0..nb) of a dense item index,
// so that items enumerate exactly the amplitudes whose listed bits are 0.
struct InsertSpec {
  int nb;
  int pos[12];
};

__device__ __forceinline__ uint64_t insert_zeros(uint64_t i, const InsertSpec& s) {
#pragma unroll 1
  for (int q = 0; q < s.nb; ++q) {
    const int p = s.pos[q];
    const uint64_t lo = i & ((1ull << p) - 1ull);
    i = ((i >> p) << (p + 1)) | lo;
  }
  return i;
}

// Sum over the first `width` lanes of a wavefront (width: power of two <= 64;
// smaller than 64 only for workgroups narrower than one wavefront, whose
// upper lanes do not exist).  Lane 0 holds the total.
__device__ __forceinline__ float wave_sum(float v, int width = 64) {
  for (int o = width >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, width);
  return v;
}

// Reduce NV values per lane over a 64-lane wavefront with log2(NV) halving
// exchanges (each lane forwards the half it does not keep) followed by a
// plain tree on one value: 10 shuffles for NV = 8 instead of 48.  Returns,
// in every lane, the wave total of value index
//   idx = bits (5, 4, 3)[: log2 NV] of the lane id (msb first);
// lanes with (lane & (64 / NV - 1)) == 0 cover each index exactly once.
template <int NV>
__device__ __forceinline__ float wave_reduce_t(const float (&v)[NV], int lane) {
  static_assert(NV == 1 || NV == 2 || NV == 4 || NV == 8, "NV must be 1, 2, 4 or 8");
  float a[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) a[i] = v[i];
  int cnt = NV;
  int sh = 32;
#pragma unroll
  for (int step = 0; step < 3; ++step) {
    if (cnt == 1) break;
    const int half = cnt >> 1;
    const bool hi = lane & sh;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < half) {
        const float keep = hi ? a[i + half] : a[i];
        const float send = hi ? a[i] : a[i + half];
        a[i] = keep + __shfl_xor(send, sh, 64);
      }
    }
    cnt = half;
    sh >>= 1;
  }
  float c = a[0];
  for (int o = sh; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
  return c;
}

}  // namespace tqsv
