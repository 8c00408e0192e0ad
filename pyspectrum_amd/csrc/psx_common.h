// Shared device/host helpers for the periodic-box estimator library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "psx.h"

namespace psx {

// ---------------------------------------------------------------- complex --
template <typename T> struct cplx { T re, im; };
typedef cplx<float> c32;
typedef cplx<double> c64;

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.re = a; r.im = b; return r; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) { return mk(a.re * s, a.im * s); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return mk(a.re, -a.im); }
// multiply by +i (s > 0) or -i (s < 0)
template <int S, typename T> __host__ __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) {
  return S > 0 ? mk(-a.im, a.re) : mk(a.im, -a.re);
}

// ------------------------------------------------------------ error state --
struct Error {
  int code = PSX_OK;
  std::string msg;
};

std::string& last_error_string();
int set_error(int code, const std::string& msg);

#define PSX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::psx::set_error(PSX_ERR_HIP, std::string(#call) + ": " +             \
                                               hipGetErrorString(_e));             \
  } while (0)

#define PSX_CHECK_LAUNCH()                                                         \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ::psx::set_error(PSX_ERR_HIP, std::string("kernel launch: ") +        \
                                               hipGetErrorString(_e));             \
  } while (0)

#define PSX_REQUIRE(cond, msg)                                                     \
  do {                                                                             \
    if (!(cond)) return ::psx::set_error(PSX_ERR_ARG, (msg));                      \
  } while (0)

#define PSX_TRY(call)                                                              \
  do {                                                                             \
    int _s = (call);                                                               \
    if (_s != PSX_OK) return _s;                                                   \
  } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// exact floor(sqrt(x)) for x < 2^52
__host__ __device__ __forceinline__ int64_t isqrt64(int64_t x) {
  int64_t r = (int64_t)sqrt((double)x);
  while (r * r > x) --r;
  while ((r + 1) * (r + 1) <= x) ++r;
  return r;
}

// Correctly rounded float32 division and square root, independent of how the
// compiler lowers '/' and sqrtf.  The quotient goes through IEEE float64
// division; rounding a 53-bit correctly rounded result to 24 bits equals
// direct rounding because 53 >= 2*24 + 2 (double rounding is innocuous).
// The square root starts from the float64 root and is then corrected with
// exact midpoint tests (midpoints have 25 bits, so their squares are exact in
// float64).
__device__ __forceinline__ float cr_divf(float a, float b) {
  return (float)((double)a / (double)b);
}

__device__ __forceinline__ float cr_sqrtf(float x) {
  if (!(x > 0.0f)) return sqrtf(x);  // 0, negative, NaN
  float r = (float)sqrt((double)x);
  const double xd = (double)x;
  const float up = nextafterf(r, INFINITY), dn = nextafterf(r, 0.0f);
  const double mhi = 0.5 * ((double)r + (double)up);
  const double mlo = 0.5 * ((double)r + (double)dn);
  if (mhi * mhi < xd || (mhi * mhi == xd && (__float_as_uint(up) & 1u) == 0)) r = up;
  else if (mlo * mlo > xd || (mlo * mlo == xd && (__float_as_uint(dn) & 1u) == 0)) r = dn;
  return r;
}

// Workgroup index b (of nb) renumbered so that each XCD runs a contiguous
// range: the dispatcher deals workgroups round-robin over the 8 XCDs, each
// with its own L2, so tiles t and t + 1 that share cache lines would
// otherwise be fetched by two L2s.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_local_block(int b, int nb) {
  constexpr int NX = 8;
  const int q = nb / NX, r = nb % NX, x = b % NX, i = b / NX;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Shell index of |k|^2 = m: int(sqrt(m)/step + 0.5) in exact integers
// (largest n with (2n-1)*step <= 2*sqrt(m)); step = 1 gives the P(k) bins.
__host__ __device__ __forceinline__ int shell_of_m(int64_t m, int step) {
  int64_t q = isqrt64((4 * m) / ((int64_t)step * step));
  return (int)((q + 1) / 2);
}

// Largest m with shell_of_m(m, step) <= j: shell_of_m is monotone in m, so
// shell j is exactly shell_mmax(j - 1) < m <= shell_mmax(j) (j >= 1; shell 0
// is m <= shell_mmax(0)).
__host__ __device__ __forceinline__ int64_t shell_mmax(int j, int step) {
  if (j < 0) return -1;
  const int64_t a = (int64_t)(2 * j + 1) * step;
  return (a * a - 1) / 4;
}

}  // namespace psx
