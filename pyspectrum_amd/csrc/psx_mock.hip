// Input side of the serving campaign on the device: seeded lognormal mocks
// and the redshift-space shift of util.py:54-75, ahead of the assignment.
//
//   psx_mock_gaussian  white noise -> forward FFT -> P(k)^1/2 filter ->
//                      backward FFT -> field / sigma (float32, mesh^3)
//   psx_mock_sample    weights exp(g - 1/2) -> fixed-order prefix sums ->
//                      n inverse-CDF cell draws, uniform offsets in the cell,
//                      and optionally Gaussian line-of-sight velocities with
//                      the redshift-space shift applied in the same pass
//   psx_mock_velocities, psx_apply_rsd  the same pieces on their own
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed, with the
// counter (index low, index high, word, stream).  Every output depends only
// on (seed, index), so results are identical for any launch shape.
#include <algorithm>
#include <cmath>

#include "psx_ctx.h"

namespace psx {

// ------------------------------------------------------------- Philox4x32 --
struct Philox4 {
  uint32_t w[4];
};

// Ten rounds; the key advances by the Weyl constants between rounds.
__host__ __device__ __forceinline__ Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.w[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.w[2];
    Philox4 n;
    n.w[0] = (uint32_t)(p1 >> 32) ^ c.w[1] ^ k0;
    n.w[1] = (uint32_t)p1;
    n.w[2] = (uint32_t)(p0 >> 32) ^ c.w[3] ^ k1;
    n.w[3] = (uint32_t)p0;
    c = n;
  }
  return c;
}

// Streams of the counter's last word.
constexpr uint32_t RNG_NOISE = 0, RNG_SAMPLE = 1, RNG_VELOCITY = 2;

__device__ __forceinline__ Philox4 draw(uint64_t seed, int64_t idx, uint32_t word, uint32_t stream) {
  Philox4 c;
  c.w[0] = (uint32_t)(uint64_t)idx;
  c.w[1] = (uint32_t)((uint64_t)idx >> 32);
  c.w[2] = word;
  c.w[3] = stream;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Uniform in [0, 1) with 53 bits: 27 bits of a, 26 of b.
__device__ __forceinline__ double uniform53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Standard normal by the Box-Muller cosine branch (1 - u1 lies in (0, 1]).
__device__ __forceinline__ double box_muller(double u1, double u2) {
  return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

__device__ __forceinline__ double normal_of(const Philox4& r) {
  return box_muller(uniform53(r.w[0], r.w[1]), uniform53(r.w[2], r.w[3]));
}

// numpy's float32 remainder (the sign of the divisor, +0 for an exact multiple).
__device__ __forceinline__ float remainder_f32(float a, float b) {
  float m = fmodf(a, b);
  if (m != 0.0f) {
    if ((b < 0.0f) != (m < 0.0f)) m += b;
  } else {
    m = copysignf(0.0f, b);
  }
  return m;
}

// util.py:73-74 for float32 positions and velocities: the shift f v + Lbox
// in float64, added to the coordinate in float64 and rounded to float32 (the
// in-place add), then wrapped in float32.
__device__ __forceinline__ float rsd_move(float x, float v, double fac, double lbox, float lbox32) {
  const float moved = (float)((double)x + (fac * (double)v + lbox));
  return remainder_f32(moved, lbox32);
}

// ------------------------------------------------------------ Gaussian field --
__global__ void k_mock_noise(cplx<float>* __restrict__ g, int64_t n, uint64_t seed) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    g[c] = mk((float)normal_of(draw(seed, c, 0, RNG_NOISE)), 0.0f);
}

// g(k) *= k^(slope/2) exp(-(k/kcut)^2 / 2), k = kf |k|; zero at k = 0.  The
// amplitude depends on |k| only, so the Hermitian symmetry of the
// transformed real noise is kept and the backward transform is real.
__global__ void k_mock_filter(cplx<float>* __restrict__ g, int M, double kf, double half_slope,
                              double kcut) {
  const int64_t n = (int64_t)M * M * M;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i0 = (int)(c / ((int64_t)M * M)), i1 = (int)((c / M) % M), i2 = (int)(c % M);
    const int a = i0 <= M / 2 ? i0 : i0 - M, b = i1 <= M / 2 ? i1 : i1 - M,
              d = i2 <= M / 2 ? i2 : i2 - M;
    const int64_t m = (int64_t)a * a + (int64_t)b * b + (int64_t)d * d;
    float amp = 0.0f;
    if (m > 0) {
      const double k = kf * sqrt((double)m);
      const double q = k / kcut;
      amp = (float)(pow(k, half_slope) * exp(-0.5 * q * q));
    }
    g[c] = g[c] * amp;
  }
}

// Sums of g and g^2 (real parts) per block, in a fixed order.
constexpr int MOM_THREADS = 256, MOM_BLOCKS = 1024;
__global__ __launch_bounds__(MOM_THREADS) void k_mock_moments(const cplx<float>* __restrict__ g,
                                                              int64_t n, double* __restrict__ part) {
  __shared__ double s1[MOM_THREADS], s2[MOM_THREADS];
  double a = 0.0, b = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * MOM_THREADS + threadIdx.x; c < n;
       c += (int64_t)MOM_BLOCKS * MOM_THREADS) {
    const double v = (double)g[c].re;
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = MOM_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// sigma = sqrt((S2 - S1^2 / n) / (n - 1)) from the block sums (one block).
__global__ __launch_bounds__(MOM_THREADS) void k_mock_sigma(const double* __restrict__ part,
                                                            int64_t n, double* __restrict__ sigma) {
  __shared__ double s1[MOM_THREADS], s2[MOM_THREADS];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < MOM_BLOCKS; i += MOM_THREADS) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = MOM_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sigma[0] = sqrt((s2[0] - s1[0] * s1[0] / (double)n) / (double)(n - 1));
}

__global__ void k_mock_normalise(const cplx<float>* __restrict__ g, int64_t n,
                                 const double* __restrict__ sigma, float* __restrict__ out) {
  const double inv = 1.0 / sigma[0];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    out[c] = (float)((double)g[c].re * inv);
}

// ------------------------------------------------------------------ sampler --
// Inclusive prefix sums of w = exp(g - 1/2) in float64, in a fixed order:
// each block scans SCAN_CHUNK consecutive cells (per thread SCAN_PER
// sequential cells, then a Hillis-Steele scan of the thread sums); one block
// scans the chunk totals; a last pass adds each chunk's offset.
constexpr int SCAN_THREADS = 256, SCAN_PER = 16, SCAN_CHUNK = SCAN_THREADS * SCAN_PER;

__global__ __launch_bounds__(SCAN_THREADS) void k_mock_scan_chunks(const float* __restrict__ g,
                                                                   int64_t n, double* __restrict__ cdf,
                                                                   double* __restrict__ totals) {
  __shared__ double ts[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
  double run[SCAN_PER];
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < SCAN_PER; ++e) {
    const int64_t c = base + e;
    const double w = c < n ? exp((double)g[c] - 0.5) : 0.0;
    s += w;
    run[e] = s;
  }
  ts[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const double add = (int)threadIdx.x >= off ? ts[threadIdx.x - off] : 0.0;
    __syncthreads();
    ts[threadIdx.x] += add;
    __syncthreads();
  }
  const double before = threadIdx.x > 0 ? ts[threadIdx.x - 1] : 0.0;
#pragma unroll
  for (int e = 0; e < SCAN_PER; ++e) {
    const int64_t c = base + e;
    if (c < n) cdf[c] = before + run[e];
  }
  if (threadIdx.x == SCAN_THREADS - 1) totals[blockIdx.x] = ts[threadIdx.x];
}

// Exclusive prefix sums of the chunk totals in place (one block, each
// thread a contiguous run), and the grand total in totals[nchunks].
__global__ __launch_bounds__(SCAN_THREADS) void k_mock_scan_totals(double* __restrict__ totals,
                                                                   int64_t nchunks) {
  __shared__ double ts[SCAN_THREADS];
  const int64_t per = (nchunks + SCAN_THREADS - 1) / SCAN_THREADS;
  const int64_t lo = std::min<int64_t>((int64_t)threadIdx.x * per, nchunks);
  const int64_t hi = std::min<int64_t>(lo + per, nchunks);
  double s = 0.0;
  for (int64_t i = lo; i < hi; ++i) s += totals[i];
  ts[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const double add = (int)threadIdx.x >= off ? ts[threadIdx.x - off] : 0.0;
    __syncthreads();
    ts[threadIdx.x] += add;
    __syncthreads();
  }
  double run = threadIdx.x > 0 ? ts[threadIdx.x - 1] : 0.0;
  for (int64_t i = lo; i < hi; ++i) {
    const double t = totals[i];
    totals[i] = run;
    run += t;
  }
  if (threadIdx.x == SCAN_THREADS - 1) totals[nchunks] = ts[SCAN_THREADS - 1];
}

__global__ void k_mock_add_offsets(double* __restrict__ cdf, int64_t n,
                                   const double* __restrict__ totals) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    cdf[c] += totals[c / SCAN_CHUNK];
}

// Point i: target u * total, cell = first index with cdf >= target (clamped
// to the last cell), position ((cell index) + offset) * h along each axis
// (axis 0 = slowest index), wrapped into [0, Lbox).  With sigma_v > 0 the
// line-of-sight coordinate is then moved by rsd_move with the velocity that
// k_mock_velocities draws for that axis (same seed).
__global__ void k_mock_sample(const double* __restrict__ cdf, const double* __restrict__ totals,
                              int64_t ncell, int M, int64_t np, uint64_t seed, double lbox,
                              double sigma_v, double fac, int los, float* __restrict__ xyz) {
  const double total = totals[ncell / SCAN_CHUNK + (ncell % SCAN_CHUNK ? 1 : 0)];
  const double h = lbox / (double)M;
  const float lbox32 = (float)lbox;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * blockDim.x) {
    const Philox4 r0 = draw(seed, i, 0, RNG_SAMPLE), r1 = draw(seed, i, 1, RNG_SAMPLE);
    const double target = uniform53(r0.w[0], r0.w[1]) * total;
    int64_t lo = 0, hi = ncell - 1;  // first index with cdf >= target, or the last cell
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] >= target) hi = mid;
      else lo = mid + 1;
    }
    const int64_t cell = lo;
    const int64_t idx[3] = {cell / ((int64_t)M * M), (cell / M) % M, cell % M};
    const double off[3] = {uniform53(r0.w[2], r0.w[3]), uniform53(r1.w[0], r1.w[1]),
                           uniform53(r1.w[2], r1.w[3])};
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = (float)fmod(((double)idx[a] + off[a]) * h, lbox);
    if (sigma_v > 0.0) {
      const float v = (float)(sigma_v * normal_of(draw(seed, i, (uint32_t)los, RNG_VELOCITY)));
      p[los] = rsd_move(p[los], v, fac, lbox, lbox32);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) xyz[(int64_t)a * np + i] = p[a];
  }
}

// Velocities (3, n): sigma * N(0, 1) rounded to float32; axis a uses word a
// of the velocity stream.
__global__ void k_mock_velocities(int64_t np, uint64_t seed, double sigma, float* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int a = 0; a < 3; ++a) v[(int64_t)a * np + i] = (float)(sigma * normal_of(draw(seed, i, a, RNG_VELOCITY)));
}

__global__ void k_apply_rsd(const float* __restrict__ xyz, const float* __restrict__ v, int64_t np,
                            double fac, double lbox, int los, float* __restrict__ out) {
  const float lbox32 = (float)lbox;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np;
       i += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float x = xyz[(int64_t)a * np + i];
      out[(int64_t)a * np + i] = a == los ? rsd_move(x, v[(int64_t)a * np + i], fac, lbox, lbox32) : x;
    }
}

static unsigned grid_for(int64_t n, int threads) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, threads), 65536));
}

int launch_mock_gaussian(psx_ctx* ctx, int M, double lbox, uint64_t seed, double slope, double kcut,
                         float* out, double* sigma_host) {
  const int64_t n = (int64_t)M * M * M;
  void *gbuf = nullptr, *part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "mock_grid", (size_t)n * sizeof(cplx<float>), &gbuf));
  PSX_TRY(ctx_buffer(ctx, "mock_moments", (size_t)(2 * MOM_BLOCKS + 1) * sizeof(double), &part));
  cplx<float>* g = (cplx<float>*)gbuf;
  double* sig = (double*)part + 2 * MOM_BLOCKS;
  hipLaunchKernelGGL(k_mock_noise, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, n, seed);
  PSX_CHECK_LAUNCH();
  PSX_TRY(launch_fft3d(ctx, g, M, -1, PSX_PREC_F32));
  hipLaunchKernelGGL(k_mock_filter, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, M,
                     2.0 * M_PI / lbox, 0.5 * slope, kcut);
  PSX_CHECK_LAUNCH();
  PSX_TRY(launch_fft3d(ctx, g, M, 1, PSX_PREC_F32));
  hipLaunchKernelGGL(k_mock_moments, dim3(MOM_BLOCKS), dim3(MOM_THREADS), 0, ctx->stream, g, n,
                     (double*)part);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mock_sigma, dim3(1), dim3(MOM_THREADS), 0, ctx->stream, (const double*)part,
                     n, sig);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mock_normalise, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, n,
                     (const double*)sig, out);
  PSX_CHECK_LAUNCH();
  if (sigma_host) {
    PSX_HIP(hipMemcpyAsync(sigma_host, sig, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PSX_HIP(hipStreamSynchronize(ctx->stream));
  }
  return PSX_OK;
}

int launch_mock_sample(psx_ctx* ctx, const float* g, int M, double lbox, int64_t np, uint64_t seed,
                       double sigma_v, double fac, int los, float* xyz) {
  const int64_t n = (int64_t)M * M * M;
  const int64_t nchunks = cdiv(n, SCAN_CHUNK);
  void *cdf = nullptr, *tot = nullptr;
  PSX_TRY(ctx_buffer(ctx, "mock_cdf", (size_t)n * sizeof(double), &cdf));
  PSX_TRY(ctx_buffer(ctx, "mock_totals", (size_t)(nchunks + 1) * sizeof(double), &tot));
  hipLaunchKernelGGL(k_mock_scan_chunks, dim3((unsigned)nchunks), dim3(SCAN_THREADS), 0,
                     ctx->stream, g, n, (double*)cdf, (double*)tot);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mock_scan_totals, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream,
                     (double*)tot, nchunks);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mock_add_offsets, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,
                     (double*)cdf, n, (const double*)tot);
  PSX_CHECK_LAUNCH();
  if (np > 0) {
    hipLaunchKernelGGL(k_mock_sample, dim3(grid_for(np, 256)), dim3(256), 0, ctx->stream,
                       (const double*)cdf, (const double*)tot, n, M, np, seed, lbox, sigma_v, fac,
                       los, xyz);
    PSX_CHECK_LAUNCH();
  }
  return PSX_OK;
}

int launch_mock_velocities(psx_ctx* ctx, int64_t np, uint64_t seed, double sigma, float* v) {
  if (np == 0) return PSX_OK;
  hipLaunchKernelGGL(k_mock_velocities, dim3(grid_for(np, 256)), dim3(256), 0, ctx->stream, np,
                     seed, sigma, v);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_apply_rsd(psx_ctx* ctx, const float* xyz, const float* v, int64_t np, double fac,
                     double lbox, int los, float* out) {
  if (np == 0) return PSX_OK;
  hipLaunchKernelGGL(k_apply_rsd, dim3(grid_for(np, 256)), dim3(256), 0, ctx->stream, xyz, v, np,
                     fac, lbox, los, out);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
