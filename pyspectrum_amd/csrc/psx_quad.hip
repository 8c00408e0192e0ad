// Quadrupole machinery of estimator.f: the Q_ij / Q_ijkl particle weights of
// assign_quad (:291-299) and the elementwise half-field passes FiveDelta2g_1
// (:514-540), FiveDelta2g_2 (:542-572) and build_quad (:574-604).
//
// Each half-field pass runs over the modes ix = 0..N/2, iy, iz of a half field
// (x fastest, the Fortran array (N/2+1, N, N)), with signed frequencies
// n <= N/2 ? n : n - N and rk = sqrt(float(kx^2 + ky^2 + kz^2)); modes with
// rk = 0 are left as they are.  The float path replays the Fortran's
// single-precision operations in its order, without contraction, with
// correctly rounded division and square root.  In FiveDelta2g_1/2 the unit
// vector components kxh, kyh, kzh are undeclared, hence INTEGER by Fortran's
// implicit typing: float(k)/rk truncated toward zero, i.e. +-1 on the axes
// and 0 elsewhere.  unit = 0 reproduces that; unit = 1 uses the real unit
// vectors (Scoccimarro 2015 eq. 20 as the comments intend).  The double path
// computes the same formulas in float64.
#include "psx_ctx.h"

namespace psx {

constexpr int QD_THREADS = 256;

__device__ __forceinline__ int qfreq(int n, int N) { return n <= N / 2 ? n : n - N; }

// Unit-vector components of one mode.  Returns false at k = 0.
template <typename T>
__device__ __forceinline__ bool quad_units(int ikx, int iky, int ikz, int unit, T& kx, T& ky, T& kz) {
#pragma clang fp contract(off)
  const int m = ikx * ikx + iky * iky + ikz * ikz;
  if constexpr (sizeof(T) == 4) {
    const float rk = cr_sqrtf((float)m);
    if (!(rk > 0.0f)) return false;
    const float ux = cr_divf((float)ikx, rk), uy = cr_divf((float)iky, rk), uz = cr_divf((float)ikz, rk);
    if (unit) { kx = ux; ky = uy; kz = uz; }
    else { kx = (float)(int)ux; ky = (float)(int)uy; kz = (float)(int)uz; }
  } else {
    if (m == 0) return false;
    const double rk = sqrt((double)m);
    const double ux = (double)ikx / rk, uy = (double)iky / rk, uz = (double)ikz / rk;
    if (unit) { kx = ux; ky = uy; kz = uz; }
    else { kx = (double)(int)ux; ky = (double)(int)uy; kz = (double)(int)uz; }
  }
  return true;
}

// dcgxx = 7.5 (dcgxx kxh^2 + dcgyy kyh^2 + dcgzz kzh^2)
template <typename T>
__global__ __launch_bounds__(QD_THREADS) void k_five_delta2_1(cplx<T>* __restrict__ xx,
                                                              const cplx<T>* __restrict__ yy,
                                                              const cplx<T>* __restrict__ zz, int N,
                                                              int unit) {
#pragma clang fp contract(off)
  const int Nh = N / 2 + 1;
  const int64_t n = (int64_t)Nh * N * N;
  for (int64_t id = (int64_t)blockIdx.x * QD_THREADS + threadIdx.x; id < n;
       id += (int64_t)gridDim.x * QD_THREADS) {
    const int ix = (int)(id % Nh), iy = (int)((id / Nh) % N), iz = (int)(id / ((int64_t)Nh * N));
    T kx, ky, kz;
    if (!quad_units<T>(ix, qfreq(iy, N), qfreq(iz, N), unit, kx, ky, kz)) continue;
    const T kx2 = kx * kx, ky2 = ky * ky, kz2 = kz * kz;
    const cplx<T> a = xx[id], b = yy[id], c = zz[id];
    const T re = ((a.re * kx2 + b.re * ky2) + c.re * kz2);
    const T im = ((a.im * kx2 + b.im * ky2) + c.im * kz2);
    xx[id] = mk(T(7.5) * re, T(7.5) * im);
  }
}

// dcgxx = dcgxx + 7.5 (2 dcgxy kxh kyh + 2 dcgyz kyh kzh + 2 dcgzx kzh kxh) - 2.5 dcg
template <typename T>
__global__ __launch_bounds__(QD_THREADS) void k_five_delta2_2(const cplx<T>* __restrict__ g,
                                                              cplx<T>* __restrict__ xx,
                                                              const cplx<T>* __restrict__ xy,
                                                              const cplx<T>* __restrict__ yz,
                                                              const cplx<T>* __restrict__ zx, int N,
                                                              int unit) {
#pragma clang fp contract(off)
  const int Nh = N / 2 + 1;
  const int64_t n = (int64_t)Nh * N * N;
  for (int64_t id = (int64_t)blockIdx.x * QD_THREADS + threadIdx.x; id < n;
       id += (int64_t)gridDim.x * QD_THREADS) {
    const int ix = (int)(id % Nh), iy = (int)((id / Nh) % N), iz = (int)(id / ((int64_t)Nh * N));
    T kx, ky, kz;
    if (!quad_units<T>(ix, qfreq(iy, N), qfreq(iz, N), unit, kx, ky, kz)) continue;
    const cplx<T> a = xx[id], p = xy[id], q = yz[id], r = zx[id], d = g[id];
    // ((2 p) kx) ky, left to right as written
    const T sre = (((T(2) * p.re) * kx) * ky + ((T(2) * q.re) * ky) * kz) + ((T(2) * r.re) * kz) * kx;
    const T sim = (((T(2) * p.im) * kx) * ky + ((T(2) * q.im) * ky) * kz) + ((T(2) * r.im) * kz) * kx;
    xx[id] = mk((a.re + T(7.5) * sre) - T(2.5) * d.re, (a.im + T(7.5) * sim) - T(2.5) * d.im);
  }
}

// dclr2 = (7.5 amu^2 - 2.5) dclr1, amu = k_los / rk (irsd 1, 2, 3 = x, y, z)
template <typename T>
__global__ __launch_bounds__(QD_THREADS) void k_build_quad(const cplx<T>* __restrict__ d1,
                                                           cplx<T>* __restrict__ d2, int irsd,
                                                           int N) {
#pragma clang fp contract(off)
  const int Nh = N / 2 + 1;
  const int64_t n = (int64_t)Nh * N * N;
  for (int64_t id = (int64_t)blockIdx.x * QD_THREADS + threadIdx.x; id < n;
       id += (int64_t)gridDim.x * QD_THREADS) {
    const int ikx = (int)(id % Nh), iky = qfreq((int)((id / Nh) % N), N),
              ikz = qfreq((int)(id / ((int64_t)Nh * N)), N);
    const int kl = irsd == 3 ? ikz : (irsd == 2 ? iky : ikx);
    const int m = ikx * ikx + iky * iky + ikz * ikz;
    T f;
    if constexpr (sizeof(T) == 4) {
      const float rk = cr_sqrtf((float)m);
      if (!(rk > 0.0f)) continue;
      const float amu = cr_divf((float)kl, rk);
      f = 7.5f * (amu * amu) - 2.5f;
    } else {
      if (m == 0) continue;
      const double amu = (double)kl / sqrt((double)m);
      f = 7.5 * (amu * amu) - 2.5;
    }
    const cplx<T> v = d1[id];
    d2[id] = mk(f * v.re, f * v.im);
  }
}

// we = w (ia = ib = ic = id = 0); w r_ia r_ib / |r|^2 (ic = id = 0);
// w r_ia r_ib r_ic r_id / |r|^4 otherwise.  Components 1..3 = x, y, z.
__global__ __launch_bounds__(QD_THREADS) void k_quad_weights(const float* __restrict__ xyz,
                                                             const float* __restrict__ w, int64_t np,
                                                             int ia, int ib, int ic, int id,
                                                             float* __restrict__ out) {
#pragma clang fp contract(off)
  for (int64_t i = (int64_t)blockIdx.x * QD_THREADS + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * QD_THREADS) {
    const float r[3] = {xyz[i], xyz[np + i], xyz[2 * np + i]};
    const float wi = w ? w[i] : 1.0f;
    float we;
    if (ia == 0 && ib == 0 && ic == 0 && id == 0) {
      we = wi;
    } else {
      const float rnorm = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
      if (ic == 0 && id == 0) we = cr_divf((wi * r[ia - 1]) * r[ib - 1], rnorm);
      else we = cr_divf((((wi * r[ia - 1]) * r[ib - 1]) * r[ic - 1]) * r[id - 1], rnorm * rnorm);
    }
    out[i] = we;
  }
}

static unsigned qd_blocks(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, QD_THREADS), 8192));
}

int launch_five_delta2_1(psx_ctx* ctx, void* xx, const void* yy, const void* zz, int N, int unit,
                         int fp) {
  const int64_t n = (int64_t)(N / 2 + 1) * N * N;
  if (fp == PSX_PREC_F64)
    hipLaunchKernelGGL(k_five_delta2_1<double>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (cplx<double>*)xx, (const cplx<double>*)yy, (const cplx<double>*)zz, N, unit);
  else
    hipLaunchKernelGGL(k_five_delta2_1<float>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (cplx<float>*)xx, (const cplx<float>*)yy, (const cplx<float>*)zz, N, unit);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_five_delta2_2(psx_ctx* ctx, const void* g, void* xx, const void* xy, const void* yz,
                         const void* zx, int N, int unit, int fp) {
  const int64_t n = (int64_t)(N / 2 + 1) * N * N;
  if (fp == PSX_PREC_F64)
    hipLaunchKernelGGL(k_five_delta2_2<double>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (const cplx<double>*)g, (cplx<double>*)xx, (const cplx<double>*)xy,
                       (const cplx<double>*)yz, (const cplx<double>*)zx, N, unit);
  else
    hipLaunchKernelGGL(k_five_delta2_2<float>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (const cplx<float>*)g, (cplx<float>*)xx, (const cplx<float>*)xy,
                       (const cplx<float>*)yz, (const cplx<float>*)zx, N, unit);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_build_quad(psx_ctx* ctx, const void* d1, void* d2, int irsd, int N, int fp) {
  const int64_t n = (int64_t)(N / 2 + 1) * N * N;
  if (fp == PSX_PREC_F64)
    hipLaunchKernelGGL(k_build_quad<double>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (const cplx<double>*)d1, (cplx<double>*)d2, irsd, N);
  else
    hipLaunchKernelGGL(k_build_quad<float>, dim3(qd_blocks(n)), dim3(QD_THREADS), 0, ctx->stream,
                       (const cplx<float>*)d1, (cplx<float>*)d2, irsd, N);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_quad_weights(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ia, int ib,
                        int ic, int id, float* out) {
  if (np == 0) return PSX_OK;
  hipLaunchKernelGGL(k_quad_weights, dim3(qd_blocks(np)), dim3(QD_THREADS), 0, ctx->stream, xyz, w,
                     np, ia, ib, ic, id, out);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
