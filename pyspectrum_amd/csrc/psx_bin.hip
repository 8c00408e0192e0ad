// k-space binning of |delta|^2: P(k) shells (Pk_periodic), the shell powers
// of the bispectrum (Parseval form of pyspectrum.py:404) and the P_l / P(k,mu)
// estimator of pk_pbox_rsd (estimator.f:155-264).
#include "psx_ctx.h"

namespace psx {

// ------------------------------------------------------------ bin power ---
// Bins the full field that reflect_delta (pyspectrum.py:1134-1157) builds from
// the half field: 1 <= x < N/2 count twice (the mirror has the same |delta|),
// x = 0 and x = N/2 once, the seven Nyquist corners with Re(delta)^2.
// bin = shell_of_m(|k|^2, step); bins 1..nbins are accumulated (k = kf*|k|).
// Output sums (k, p, n) per bin; the caller divides.
//
// Deterministic: every sum is formed in a fixed order, so the result is
// bit-identical run to run.  A wave takes whole (y, z) rows of the half field,
// 64 consecutive x at a time; |k|^2 = x^2 + c grows with x, so the lanes of
// one bin are one contiguous run (and a row or chunk that starts past the
// last bin is skipped with everything after it).  A segmented scan over the lanes (fixed shift order)
// sums each run, and the run's last lane adds it to the wave's own
// accumulators in LDS (one lane per bin: no two lanes share an address).
// The block then adds its waves in order and stores a partial per block;
// k_bin_reduce adds the blocks in a fixed order.
constexpr int BP_THREADS = 256, BP_WAVES = BP_THREADS / 64, BP_MAX_BLOCKS = 512;
constexpr int BP_MAX_BINS = 1024;  // LDS: BP_WAVES x 3 x nbins doubles

template <typename T>
__global__ __launch_bounds__(BP_THREADS) void k_bin_power(const cplx<T>* __restrict__ Hh, int N,
                                                          int step, int nbins, double kf,
                                                          double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* accw = reinterpret_cast<double*>(smem_raw);  // [BP_WAVES][3][nbins]
  for (int i = threadIdx.x; i < BP_WAVES * 3 * nbins; i += BP_THREADS) accw[i] = 0.0;
  __syncthreads();
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  double* acc = accw + wave * 3 * nbins;
  const int Nh = N / 2 + 1, hg = N / 2;
  const int64_t mlast = shell_mmax(nbins, step);  // largest |k|^2 of bin nbins
  // a wave takes whole (y, z) rows, 64 consecutive x at a time, in order
  for (int row = (int)blockIdx.x * BP_WAVES + wave; row < N * N; row += (int)gridDim.x * BP_WAVES) {
    const int y = row % N, z = row / N;
    const int ay = min(y, N - y), az = min(z, N - z);
    const int64_t c = (int64_t)ay * ay + (int64_t)az * az;
    if (c > mlast) continue;                       // the whole row lies past bin nbins
    for (int x0 = 0; x0 < Nh; x0 += 64) {
      if (c + (int64_t)x0 * x0 > mlast) break;     // this and later chunks lie past it
      const int x = x0 + lane;
      int key = -1;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
      if (x < Nh) {
        const int64_t m = c + (int64_t)x * x;
        const int b = shell_of_m(m, step);
        if (b >= 1 && b <= nbins) {
          const cplx<T> v = Hh[(int64_t)row * Nh + x];
          const bool xe = (x == 0 || x == hg);
          const bool corner = xe && (y == 0 || y == hg) && (z == 0 || z == hg);
          const double re = (double)v.re, im = corner ? 0.0 : (double)v.im;
          const double w = xe ? 1.0 : 2.0;
          key = b - 1;
          v0 = w * (kf * sqrt((double)m));
          v1 = w * (re * re + im * im);
          v2 = w;
        }
      }
      // keys -1 (outside 1..nbins or x >= Nh) only lead or trail the valid lanes
      if (__ballot(key >= 0) == 0) continue;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int ku = __shfl_up(key, off, 64);
        const double u0 = __shfl_up(v0, off, 64), u1 = __shfl_up(v1, off, 64),
                     u2 = __shfl_up(v2, off, 64);
        if (lane >= off && ku == key) { v0 += u0; v1 += u1; v2 += u2; }
      }
      const int kn = __shfl_down(key, 1, 64);
      if (key >= 0 && (lane == 63 || kn != key)) {
        acc[key] += v0;
        acc[nbins + key] += v1;
        acc[2 * nbins + key] += v2;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * nbins; i += BP_THREADS) {
    double s = accw[i];
#pragma unroll
    for (int w = 1; w < BP_WAVES; ++w) s += accw[w * 3 * nbins + i];
    part[(int64_t)blockIdx.x * 3 * nbins + i] = s;
  }
}

// out[e] = sum over blocks of part[block][e]: lane l adds blocks l, l+64, ...
// in order, then a fixed butterfly over the lanes.
__global__ __launch_bounds__(64) void k_bin_reduce(const double* __restrict__ part, int nblocks,
                                                   int nent, double* __restrict__ out) {
  const int e = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int b = lane; b < nblocks; b += 64) s += part[(int64_t)b * nent + e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[e] = s;
}

__global__ void k_fill_f64(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

int launch_bin_power(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                     int prec, double* out3) {
  PSX_REQUIRE(nbins >= 1 && nbins <= BP_MAX_BINS, "nbins must be in [1, 1024]");
  const size_t lds = (size_t)BP_WAVES * 3 * nbins * sizeof(double);
  const int blocks = (int)std::min<int64_t>(cdiv((int64_t)N * N, BP_WAVES), BP_MAX_BLOCKS);
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "bin_partials", (size_t)blocks * 3 * nbins * sizeof(double), &part));
  if (prec == PSX_PREC_F64) {
    PSX_TRY(set_lds(k_bin_power<double>, lds));
    hipLaunchKernelGGL(k_bin_power<double>, dim3(blocks), dim3(BP_THREADS), lds, ctx->stream,
                       (const cplx<double>*)half, N, step, nbins, kf, (double*)part);
  } else {
    PSX_TRY(set_lds(k_bin_power<float>, lds));
    hipLaunchKernelGGL(k_bin_power<float>, dim3(blocks), dim3(BP_THREADS), lds, ctx->stream,
                       (const cplx<float>*)half, N, step, nbins, kf, (double*)part);
  }
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bin_reduce, dim3(3 * nbins), dim3(64), 0, ctx->stream, (const double*)part,
                     blocks, 3 * nbins, out3);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

// p0k[j-1] = sum |delta|^2 / Nk[j] for j = max(j0,1)..nmax (NaN if Nk = 0),
// 0 elsewhere: the shell power of _Bk_periodic (pyspectrum.py:404) by Parseval.
__global__ void k_p0k(const double* __restrict__ bins3, int nmax, int j0, double* __restrict__ p0k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmax) return;
  const int j = i + 1;
  p0k[i] = (j >= j0) ? bins3[nmax + i] / bins3[2 * nmax + i] : 0.0;
}

int launch_p0k(psx_ctx* ctx, const double* bins3, int nmax, int j0, double* p0k) {
  hipLaunchKernelGGL(k_p0k, dim3((unsigned)cdiv(nmax, 64)), dim3(64), 0, ctx->stream, bins3, nmax,
                     j0, p0k);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

// --------------------------------------------------------- pk_pbox_rsd ----
// Restatement of the mode loop of estimator.f:196-244.  The Fortran visits
// the full grid, x > N/2 through the mirrored element; that mode is -k' of
// the half-field mode k' = (N - x, -y, -z), and in its float32 operations
// every sign flip is an exact negation, so it has the same rk, k bin, |mu|,
// mu bin, Legendre weights and |delta|^2.  The kernel therefore visits the
// half field once and counts 0 < x < N/2 twice (x = 0 and x = N/2 map to
// themselves): the same bins and counts, sums equal to round-off.  Except on
// the y or z Nyquist planes: there the mirror's component is +N/2 (index N/2
// maps to itself), not the negation, and the cos(pi/2) residue of the
// float32 line of sight makes its |mu| differ, so that mirror is replayed
// with the Fortran's own frequencies.  Every
// float32 operation that decides a bin is replayed in the reference's order
// without contraction (k bin: nint(Nbin*rk/(N/2)); mu bin:
// int((|mu| + dmu)/dmu) in float64 of float32 values).
struct RsdConst {
  float kf, mubin, cth, sth, cph, sph;
  int nbin, nmu;
};

__device__ __forceinline__ int fort_freq(int n, int N) { return n <= N / 2 ? n : n - N; }

template <typename T>
__global__ __launch_bounds__(256) void k_pk_rsd(const cplx<T>* __restrict__ Hh, int N, RsdConst c,
                                                double* __restrict__ gout, int use_lds) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // layout: [5][nbin] (k, p0, p2, p4, nk) then [4][nbin*nmu] (km, mk, pkm, nkm)
  const int nb = c.nbin, nm = c.nmu;
  const int ntot = 5 * nb + 4 * nb * nm;
  double* acc = use_lds ? reinterpret_cast<double*>(smem_raw) : gout;
  if (use_lds) {
    for (int i = threadIdx.x; i < ntot; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
  }
  const int Nh = N / 2 + 1;
  const int64_t n = (int64_t)Nh * N * N;
  const float fnb = (float)nb, fh = (float)(N / 2);
  const double md = (double)c.mubin;
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(id % Nh);
    const int iy = (int)((id / Nh) % N);
    const int iz = (int)(id / ((int64_t)Nh * N));
    const cplx<T> ct = Hh[id];
    double pk;
    if constexpr (sizeof(T) == 8) {
      pk = (double)ct.re * (double)ct.re + (double)ct.im * (double)ct.im;
    } else {  // the reference's float32 cabs(ct)**2
      const float a = hypotf(ct.re, ct.im);
      pk = (double)(a * a);
    }
    // one mode with frequencies (kx, ky, kz) and |delta|^2 = pk, counted wm times
    auto mode = [&](int kx, int ky, int kz, double wm) {
      float rk, mu32;
      int imk;
      {
#pragma clang fp contract(off)
        const float rkx = (float)kx, rky = (float)ky, rkz = (float)kz;
        rk = cr_sqrtf(rkx * rkx + rky * rky + rkz * rkz);
        imk = (int)roundf(cr_divf(fnb * rk, fh));  // NINT: half away from zero
        if (!(imk <= nb && imk != 0)) return;
        const float cot1 = cr_divf(rkz, rk);
        const float sit1 = cr_sqrtf(1.0f - cot1 * cot1);
        float cc = 0.0f;
        if (sit1 > 0.0f) {
          const float den = rk * sit1;
          const float cp = cr_divf(rkx, den), sp = cr_divf(rky, den);
          cc = c.sph * sp + c.cph * cp;
        }
        mu32 = c.cth * cot1 + (c.sth * sit1) * cc;
      }
      const double mu = (double)mu32;
      const int imu = (int)((fabs(mu) + md) / md);
      const double mu2 = mu * mu;
      const double Le2 = -0.5 + 1.5 * mu2;
      const double Le4 = 0.375 - 3.75 * mu2 + 4.375 * (mu2 * mu2);
      const double kk = (double)(c.kf * rk);
      const int b = imk - 1;
      atomicAdd(&acc[0 * nb + b], wm * kk);
      atomicAdd(&acc[1 * nb + b], wm * pk);
      atomicAdd(&acc[2 * nb + b], wm * (pk * 5.0 * Le2));
      atomicAdd(&acc[3 * nb + b], wm * (pk * 9.0 * Le4));
      atomicAdd(&acc[4 * nb + b], wm);
      if (imu <= nm && imu > 0) {
        const int bm = b * nm + (imu - 1);
        double* a2 = acc + 5 * nb;
        atomicAdd(&a2[0 * nb * nm + bm], wm * kk);
        atomicAdd(&a2[1 * nb * nm + bm], wm * fabs(mu));
        atomicAdd(&a2[2 * nb * nm + bm], wm * pk);
        atomicAdd(&a2[3 * nb * nm + bm], wm);
      }
    };
    const int ky = fort_freq(iy, N), kz = fort_freq(iz, N);
    if (ix == 0 || ix == N / 2) {
      mode(ix, ky, kz, 1.0);  // these planes map to themselves
    } else if (iy != N / 2 && iz != N / 2) {
      mode(ix, ky, kz, 2.0);  // the mirror is the exact negation
    } else {
      // on the y/z Nyquist planes the mirror's frequency there is +N/2 (index
      // N/2 maps to itself), not the negation: replay it separately
      mode(ix, ky, kz, 1.0);
      mode(-ix, fort_freq((N - iy) % N, N), fort_freq((N - iz) % N, N), 1.0);
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < ntot; i += blockDim.x)
      if (acc[i] != 0.0) atomicAdd(&gout[i], acc[i]);
  }
}

int launch_pk_rsd(psx_ctx* ctx, const void* half, int N, int irsd, int lbox, int nbin, int nmu,
                  int prec, double* gout) {
  PSX_REQUIRE(irsd >= 0 && irsd <= 2, "rsd must be 0, 1 or 2");
  PSX_REQUIRE(nbin >= 1 && nmu >= 1 && lbox > 0, "bad Nbin/Nmu/Lbox");
  RsdConst c;
  const float pi = 3.141592654f;
  const float tpi = 2.0f * pi;
  c.kf = tpi / (float)lbox;
  c.mubin = 1.0f / (float)nmu;
  const float th = irsd == 2 ? 0.0f : 0.5f * pi;
  const float ph = irsd == 1 ? 0.5f * pi : 0.0f;
  // sin/cos of these float32 constants, correctly rounded
  c.cth = (float)cos((double)th); c.sth = (float)sin((double)th);
  c.cph = (float)cos((double)ph); c.sph = (float)sin((double)ph);
  c.nbin = nbin; c.nmu = nmu;
  const int64_t ntot = 5 * (int64_t)nbin + 4 * (int64_t)nbin * nmu;
  hipLaunchKernelGGL(k_fill_f64, dim3(64), dim3(256), 0, ctx->stream, gout, ntot);
  PSX_CHECK_LAUNCH();
  const size_t lds = (size_t)ntot * sizeof(double);
  const int use_lds = lds <= 120 * 1024;
  const int64_t n = (int64_t)(N / 2 + 1) * N * N;
  const unsigned blocks = (unsigned)std::min<int64_t>(cdiv(n, 256), use_lds ? 1024 : 4096);
  const size_t shm = use_lds ? lds : 0;
  if (prec == PSX_PREC_F64) {
    if (shm > 65536) PSX_HIP(hipFuncSetAttribute((const void*)k_pk_rsd<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k_pk_rsd<double>, dim3(blocks), dim3(256), shm, ctx->stream,
                       (const cplx<double>*)half, N, c, gout, use_lds);
  } else {
    if (shm > 65536) PSX_HIP(hipFuncSetAttribute((const void*)k_pk_rsd<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k_pk_rsd<float>, dim3(blocks), dim3(256), shm, ctx->stream,
                       (const cplx<float>*)half, N, c, gout, use_lds);
  }
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
