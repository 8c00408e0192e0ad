// k-space binning of |delta|^2: P(k) shells (Pk_periodic), the shell powers
// of the bispectrum (Parseval form of pyspectrum.py:404) and the P_l / P(k,mu)
// estimator of pk_pbox_rsd (estimator.f:155-264).
#include <mutex>

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
  // box blockIdx.y of a batch: its half field and its block partials
  Hh += (int64_t)blockIdx.y * N * N * (N / 2 + 1);
  part += (int64_t)blockIdx.y * gridDim.x * 3 * nbins;
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
// in order, then a fixed butterfly over the lanes.  Box blockIdx.y of a
// batch: part [box][nblocks][nent], out [box][nent].
__global__ __launch_bounds__(64) void k_bin_reduce(const double* __restrict__ part, int nblocks,
                                                   int nent, double* __restrict__ out) {
  part += (int64_t)blockIdx.y * nblocks * nent;
  out += (int64_t)blockIdx.y * nent;
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

// Block partials of k_bin_power: one [3][nbins] record per block.
static int bin_power_blocks(int N) {
  return (int)std::min<int64_t>(cdiv((int64_t)N * N, BP_WAVES), BP_MAX_BLOCKS);
}

// k_bin_power + k_bin_reduce with the block partials in `part` (at least
// bin_power_blocks(N) * 3 * nbins doubles; every entry is written before it
// is read, so nothing needs clearing).
// nbox half fields back to back give out3 [nbox][3][nbins].
static int bin_power_into(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                          int prec, double* out3, void* part, int nbox = 1) {
  const size_t lds = (size_t)BP_WAVES * 3 * nbins * sizeof(double);
  const int blocks = bin_power_blocks(N);
  if (prec == PSX_PREC_F64) {
    PSX_TRY(set_lds(k_bin_power<double>, lds));
    hipLaunchKernelGGL(k_bin_power<double>, dim3(blocks, nbox), dim3(BP_THREADS), lds,
                       ctx->stream, (const cplx<double>*)half, N, step, nbins, kf, (double*)part);
  } else {
    PSX_TRY(set_lds(k_bin_power<float>, lds));
    hipLaunchKernelGGL(k_bin_power<float>, dim3(blocks, nbox), dim3(BP_THREADS), lds,
                       ctx->stream, (const cplx<float>*)half, N, step, nbins, kf, (double*)part);
  }
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bin_reduce, dim3(3 * nbins, nbox), dim3(64), 0, ctx->stream,
                     (const double*)part, blocks, 3 * nbins, out3);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_bin_power(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                     int prec, double* out3, int nbox) {
  PSX_REQUIRE(nbins >= 1 && nbins <= BP_MAX_BINS, "nbins must be in [1, 1024]");
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "bin_partials",
                     (size_t)nbox * bin_power_blocks(N) * 3 * nbins * sizeof(double), &part));
  return bin_power_into(ctx, half, N, step, nbins, kf, prec, out3, part, nbox);
}

// ------------------------------------------- bin power without allocation ---
// The block partials of psx_bin_power_noalloc live in this device array
// (NOALLOC_DOUBLES doubles = 6 MiB: 512 blocks x 3 sums x 512 bins), which
// each device holds until the process ends.  Host code never passes the
// array's name to a kernel: its device address is looked up once per device
// with hipGetSymbolAddress and kept in noalloc_ptr[device].  Every call on a
// device shares that device's array, so calls that overlap on one device
// race; each call writes every partial it reads, so nothing needs clearing.
constexpr int NOALLOC_MAX_BINS = 512;
constexpr size_t NOALLOC_DOUBLES = (size_t)BP_MAX_BLOCKS * 3 * NOALLOC_MAX_BINS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mu;

// The device address of `scratch` on the context's device (looked up on the
// first call with that device current).
static int noalloc_scratch(psx_ctx* ctx, void** out) {
  const int dev = ctx->device;
  PSX_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    int cur = 0;
    PSX_HIP(hipGetDevice(&cur));
    if (cur != dev) PSX_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(PSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return PSX_OK;
}

int launch_bin_power_noalloc(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                             int prec, double* out3) {
  PSX_REQUIRE(nbins >= 1 && nbins <= NOALLOC_MAX_BINS,
              "psx_bin_power_noalloc: nbins must be in [1, 512] (fixed scratch of 6 MiB)");
  PSX_REQUIRE((size_t)bin_power_blocks(N) * 3 * nbins <= NOALLOC_DOUBLES,
              "psx_bin_power_noalloc: partials exceed the fixed scratch");
  void* part = nullptr;
  PSX_TRY(noalloc_scratch(ctx, &part));
  return bin_power_into(ctx, half, N, step, nbins, kf, prec, out3, part);
}

// p0k[j-1] = sum |delta|^2 / Nk[j] for j = max(j0,1)..nmax (NaN if Nk = 0),
// 0 elsewhere: the shell power of _Bk_periodic (pyspectrum.py:404) by Parseval.
// Box blockIdx.y of a batch: bins3 [box][3][nmax], p0k [box][nmax].
__global__ void k_p0k(const double* __restrict__ bins3, int nmax, int j0, double* __restrict__ p0k) {
  bins3 += (int64_t)blockIdx.y * 3 * nmax;
  p0k += (int64_t)blockIdx.y * nmax;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmax) return;
  const int j = i + 1;
  p0k[i] = (j >= j0) ? bins3[nmax + i] / bins3[2 * nmax + i] : 0.0;
}

int launch_p0k(psx_ctx* ctx, const double* bins3, int nmax, int j0, double* p0k, int nbox) {
  hipLaunchKernelGGL(k_p0k, dim3((unsigned)cdiv(nmax, 64), nbox), dim3(64), 0, ctx->stream, bins3,
                     nmax, j0, p0k);
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
// with the Fortran's own frequencies as a second pass over those rows.
// Every float32 operation that decides a bin is replayed in the reference's
// order without contraction (k bin: nint(Nbin*rk/(N/2)); mu bin:
// int((|mu| + dmu)/dmu) in float64 of float32 values).
//
// Deterministic: a wave takes 64 consecutive x of one (y, z) row.  The k bin
// depends only on |k|^2 = x^2 + c, so it never decreases along the lanes;
// a segmented scan (fixed shift order) sums each run of equal keys and the
// run's last lane adds the run to the block's accumulators.  The (k, mu)
// key is scanned the same way; should a key come back after another key
// within the 64 lanes (two run ends with one key, found by a tag written and
// read back in LDS), those run ends are added one lane at a time in lane
// order.  The four waves of a block add their runs in turn, so the order of
// every addition is fixed; k_bin_reduce then adds the block partials in a
// fixed order.
struct RsdConst {
  float kf, mubin, cth, sth, cph, sph;
  int nbin, nmu;
  int64_t mlim;  // no |k|^2 above this has a k bin <= nbin
};

__device__ __forceinline__ int fort_freq(int n, int N) { return n <= N / 2 ? n : n - N; }

// One mode (kx, ky, kz) with |delta|^2 = pk counted wm times: k bin (0-based,
// -1 if outside 1..nbin), (k, mu) bin (-1 if outside), and the nine sums.
struct RsdMode {
  int kb, mb;
  double v[9];  // k, p0, p2, p4, n | km, mk, pkm, nkm
};

__device__ __forceinline__ RsdMode rsd_mode(int kx, int ky, int kz, double pk, double wm,
                                            const RsdConst& c, float fnb, float fh, double md) {
  RsdMode r;
  r.kb = -1;
  r.mb = -1;
#pragma unroll
  for (int e = 0; e < 9; ++e) r.v[e] = 0.0;
  float rk, mu32;
  int imk;
  {
#pragma clang fp contract(off)
    const float rkx = (float)kx, rky = (float)ky, rkz = (float)kz;
    rk = cr_sqrtf(rkx * rkx + rky * rky + rkz * rkz);
    imk = (int)roundf(cr_divf(fnb * rk, fh));  // NINT: half away from zero
    if (!(imk <= c.nbin && imk != 0)) return r;
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
  r.kb = imk - 1;
  r.v[0] = wm * kk;
  r.v[1] = wm * pk;
  r.v[2] = wm * (pk * 5.0 * Le2);
  r.v[3] = wm * (pk * 9.0 * Le4);
  r.v[4] = wm;
  if (imu <= c.nmu && imu > 0) {
    r.mb = r.kb * c.nmu + (imu - 1);
    r.v[5] = wm * kk;
    r.v[6] = wm * fabs(mu);
    r.v[7] = wm * pk;
    r.v[8] = wm;
  }
  return r;
}

// Segmented inclusive scan of the NV values v[] over runs of equal key (lane
// order, fixed shift order); returns whether this lane ends its run.
template <int NV>
__device__ __forceinline__ bool run_scan(int key, double* v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int ku = __shfl_up(key, off, 64);
    double u[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) u[e] = __shfl_up(v[e], off, 64);
    if (lane >= off && ku == key) {
#pragma unroll
      for (int e = 0; e < NV; ++e) v[e] += u[e];
    }
  }
  const int kn = __shfl_down(key, 1, 64);
  return key >= 0 && (lane == 63 || kn != key);
}

// Adds each run end's NV sums to acc[e * stride + key]; run ends that share a
// key (a key that came back) are added one lane at a time in lane order.
// The lanes of a wave exchange data through LDS here (the tag written and
// read back, the one-lane-at-a-time adds), so those accesses are volatile:
// issued in program order, never merged or forwarded.
template <int NV>
__device__ __forceinline__ void run_add(bool end, int key, const double* v, double* acc,
                                        int stride, int* tag, int lane) {
  volatile int* vt = tag;
  if (end) vt[key] = lane;
  const bool clash = end && vt[key] != lane;
  if (__ballot(clash) == 0) {
    if (end) {
#pragma unroll
      for (int e = 0; e < NV; ++e) acc[e * stride + key] += v[e];
    }
  } else {
    volatile double* va = acc;
    for (int l = 0; l < 64; ++l) {
      if (lane == l && end) {
#pragma unroll
        for (int e = 0; e < NV; ++e) va[e * stride + key] = va[e * stride + key] + v[e];
      }
    }
  }
}

constexpr int RSD_THREADS = 256, RSD_WAVES = RSD_THREADS / 64, RSD_MAX_BLOCKS = 512;

template <typename T>
__global__ __launch_bounds__(RSD_THREADS) void k_pk_rsd(const cplx<T>* __restrict__ Hh, int N,
                                                         RsdConst c, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // layout: [5][nbin] (k, p0, p2, p4, nk) then [4][nbin*nmu] (km, mk, pkm, nkm)
  const int nb = c.nbin, nm = c.nmu, nbm = nb * nm;
  const int ntot = 5 * nb + 4 * nbm;
  double* acc = reinterpret_cast<double*>(smem_raw);
  int* tagk = reinterpret_cast<int*>(acc + ntot);
  int* tagm = tagk + nb;
  for (int i = threadIdx.x; i < ntot; i += RSD_THREADS) acc[i] = 0.0;
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int Nh = N / 2 + 1, hN = N / 2;
  const float fnb = (float)nb, fh = (float)hN;
  const double md = (double)c.mubin;
  // work items of this wave: rows row0, row0 + rstride, ...; each row has one
  // pass (two on the y/z Nyquist planes: the original modes, then the
  // replayed mirrors) of 64-lane chunks x0 = 0, 64, ...
  const int rstride = (int)gridDim.x * RSD_WAVES;
  int row = (int)blockIdx.x * RSD_WAVES + wave, pass = 0, x0 = 0;
  auto row_c = [&](int r) {
    const int y = r % N, z = r / N;
    const int ky = fort_freq(y, N), kz = fort_freq(z, N);
    return (int64_t)ky * ky + (int64_t)kz * kz;
  };
  // advance to the next chunk that holds a mode below mlim
  auto settle = [&]() {
    while (row < N * N) {
      const int y = row % N, z = row / N;
      const int npass = (y == hN || z == hN) ? 2 : 1;
      const int64_t cc = row_c(row);
      if (pass < npass && x0 < Nh && cc + (int64_t)x0 * x0 <= c.mlim) return;
      if (pass + 1 < npass && x0 > 0) { ++pass; x0 = 0; continue; }
      row += rstride; pass = 0; x0 = 0;
    }
  };
  settle();
  __syncthreads();
  for (;;) {
    const bool has = row < N * N;
    if (!__syncthreads_or(has)) break;
    RsdMode m;
    m.kb = -1; m.mb = -1;
#pragma unroll
    for (int e = 0; e < 9; ++e) m.v[e] = 0.0;
    if (has) {
      const int y = row % N, z = row / N;
      const int x = x0 + lane;
      if (x < Nh) {
        const cplx<T> ct = Hh[(int64_t)row * Nh + x];
        double pk;
        if constexpr (sizeof(T) == 8) {
          pk = (double)ct.re * (double)ct.re + (double)ct.im * (double)ct.im;
        } else {  // the reference's float32 cabs(ct)**2
          const float a = hypotf(ct.re, ct.im);
          pk = (double)(a * a);
        }
        const int ky = fort_freq(y, N), kz = fort_freq(z, N);
        const bool nyq = (y == hN || z == hN);
        if (pass == 0) {
          // x = 0 and x = N/2 map to themselves; elsewhere the mirror is the
          // exact negation (weight 2) except on the Nyquist planes
          const double wm = (x == 0 || x == hN || nyq) ? 1.0 : 2.0;
          m = rsd_mode(x, ky, kz, pk, wm, c, fnb, fh, md);
        } else if (x != 0 && x != hN) {
          // the mirror on a y/z Nyquist plane, with the Fortran's frequencies
          m = rsd_mode(-x, fort_freq((N - y) % N, N), fort_freq((N - z) % N, N), pk, 1.0, c, fnb,
                       fh, md);
        }
      }
    }
    const bool endk = run_scan<5>(m.kb, m.v, lane);
    const bool endm = run_scan<4>(m.mb, m.v + 5, lane);
    for (int w = 0; w < RSD_WAVES; ++w) {
      if (w == wave && has) {
        run_add<5>(endk, m.kb, m.v, acc, nb, tagk, lane);
        run_add<4>(endm, m.mb, m.v + 5, acc + 5 * nb, nbm, tagm, lane);
      }
      __syncthreads();
    }
    if (has) {
      x0 += 64;
      settle();
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ntot; i += RSD_THREADS) part[(int64_t)blockIdx.x * ntot + i] = acc[i];
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
  // imk = nint(nbin*rk/(N/2)) <= nbin needs rk < (nbin + 1/2)(N/2)/nbin; the
  // bound below adds a margin far above float32 rounding
  const double rmax = (nbin + 0.5) * (N / 2) / (double)nbin * (1.0 + 1e-5) + 1e-3;
  c.mlim = (int64_t)std::ceil(rmax * rmax) + 1;
  const int64_t ntot = 5 * (int64_t)nbin + 4 * (int64_t)nbin * nmu;
  const size_t lds = (size_t)ntot * sizeof(double) + (size_t)(nbin + nbin * nmu) * sizeof(int);
  PSX_REQUIRE(lds <= 160 * 1024, "Nbin*Nmu too large for the binning kernel (LDS)");
  const int blocks = (int)std::min<int64_t>(cdiv((int64_t)N * N, RSD_WAVES), RSD_MAX_BLOCKS);
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "rsd_partials", (size_t)blocks * ntot * sizeof(double), &part));
  if (prec == PSX_PREC_F64) {
    PSX_TRY(set_lds(k_pk_rsd<double>, lds));
    hipLaunchKernelGGL(k_pk_rsd<double>, dim3(blocks), dim3(RSD_THREADS), lds, ctx->stream,
                       (const cplx<double>*)half, N, c, (double*)part);
  } else {
    PSX_TRY(set_lds(k_pk_rsd<float>, lds));
    hipLaunchKernelGGL(k_pk_rsd<float>, dim3(blocks), dim3(RSD_THREADS), lds, ctx->stream,
                       (const cplx<float>*)half, N, c, (double*)part);
  }
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bin_reduce, dim3((unsigned)ntot), dim3(64), 0, ctx->stream,
                     (const double*)part, blocks, (int)ntot, gout);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
