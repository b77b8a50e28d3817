// Shared device/host helpers for the gfx950 BSE kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>

// ---------------------------------------------------------------------------
// complex double with value semantics (interleaved re,im as in Fortran COMPLEX*16)
// ---------------------------------------------------------------------------
struct __attribute__((aligned(16))) zc {
  double re, im;
};

__host__ __device__ inline zc mkz(double r, double i) { zc z; z.re = r; z.im = i; return z; }
__host__ __device__ inline zc operator+(zc a, zc b) { return mkz(a.re + b.re, a.im + b.im); }
__host__ __device__ inline zc operator-(zc a, zc b) { return mkz(a.re - b.re, a.im - b.im); }
__host__ __device__ inline zc operator-(zc a) { return mkz(-a.re, -a.im); }
// (a.re*b.re - a.im*b.im) + i(a.re*b.im + a.im*b.re): the Fortran complex product
__host__ __device__ inline zc operator*(zc a, zc b) {
  return mkz(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ inline zc operator*(double s, zc a) { return mkz(s * a.re, s * a.im); }
__host__ __device__ inline zc operator*(zc a, double s) { return mkz(a.re * s, a.im * s); }
__host__ __device__ inline zc conj(zc a) { return mkz(a.re, -a.im); }
__host__ __device__ inline double norm2(zc a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline zc& operator+=(zc& a, zc b) { a.re += b.re; a.im += b.im; return a; }
__host__ __device__ inline zc& operator-=(zc& a, zc b) { a.re -= b.re; a.im -= b.im; return a; }
// complex float storage (COMPLEX*8) for the precision form; arithmetic stays in zc
struct __attribute__((aligned(8))) zf {
  float re, im;
};
__host__ __device__ inline zc ldz(const zc& v) { return v; }
__host__ __device__ inline zc ldz(const zf& v) { return mkz((double)v.re, (double)v.im); }
__host__ __device__ inline void stz(zc& d, zc v) { d = v; }
__host__ __device__ inline void stz(zf& d, zc v) {
  d.re = (float)v.re;
  d.im = (float)v.im;
}

// conj(a)*b accumulated with fused multiply-adds
__device__ inline zc fma_conj(zc a, zc b, zc acc) {
  acc.re = fma(a.re, b.re, fma(a.im, b.im, acc.re));
  acc.im = fma(a.re, b.im, fma(-a.im, b.re, acc.im));
  return acc;
}
__device__ inline zc fma_z(zc a, zc b, zc acc) {
  acc.re = fma(a.re, b.re, fma(-a.im, b.im, acc.re));
  acc.im = fma(a.re, b.im, fma(a.im, b.re, acc.im));
  return acc;
}

// sum of p[b * stride] over b = b0, b0 + step, ... < b1, added in that order, with the loads
// issued eight at a time (a plain loop waits for each load before the next).  The last
// group loads from clamped indices (b1 - 1 is in range whenever the group runs) and adds
// only the terms below b1, so short sums keep their loads in flight too.
template <class V>
__device__ inline V strided_sum(const V* __restrict__ p, size_t stride, int b0, int b1, int step, V acc) {
  int b = b0;
  for (; b + 7 * step < b1; b += 8 * step) {
    V v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(b + u * step) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  if (b < b1) {
    V v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int bu = b + u * step;
      v[u] = p[(size_t)(bu < b1 ? bu : b1 - 1) * stride];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b + u * step < b1) acc += v[u];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// MFMA lane layouts on gfx950 (measured with a lane probe):
//   v_mfma_f64_16x16x4_f64: lane l holds A[l%16][l/16] and B[l/16][l%16];
//     accumulator register r of lane l is D[(l/16) + 4*r][l%16].
//   v_mfma_f32_16x16x4_f32: same A and B; register r is D[4*(l/16) + r][l%16].
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16x16x4f32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// host error plumbing
// ---------------------------------------------------------------------------
namespace wtb {
void set_error(const std::string& msg);
}

#define WTB_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      wtb::set_error(std::string(#call) + ": " + hipGetErrorString(_e));     \
      return WTB_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define WTB_LAUNCH_CHECK()                                                   \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) {                                                  \
      wtb::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
      return WTB_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)
