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

}  // namespace tqsv
