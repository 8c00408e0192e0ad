// Mesh front end of the periodic-box estimator: PCS assignment with
// interlacing, the interlacing/deconvolution combination, and the half <->
// full field conversions.
#include <cstring>

#include "psx_ctx.h"

namespace psx {

// ------------------------------------------------------------ assignment --
// Per-axis PCS stencil of estimator.f:302-378, replayed with the reference's
// float32 operation order (no contraction into FMA): rx = kf_ks*x + 1 + offset
// is the 1-based grid coordinate (offset = 0 for boxes, Ngrid/2 for survey
// coordinates centred on the observer, pyspectrum.py:819); grid A uses rx,
// grid B uses rx + 0.5.  Returns the 0-based base cell (wrapped into [0, N))
// and the four weights (x6) for cells base-1 .. base+2.
// (HIP's __fmul_rn/__fadd_rn are plain operators, so FMA contraction is
// switched off explicitly: the reference multiplies and adds separately.)
__device__ __forceinline__ float pcs_coord(float r, const PcsMap& pm, bool shift, int N, int* c) {
#pragma clang fp contract(off)
  float rx = pm.kf_ks * r + 1.0f;
  rx = rx + pm.offset;
  if (shift) rx = rx + 0.5f;
  const int i1 = (int)rx;  // rx >= 0 for positions inside the clip range: truncation = floor
  const int c0 = i1 - 1;   // in [-1, N]
  *c = c0 < 0 ? c0 + N : (c0 >= N ? c0 - N : c0);
  return rx - (float)i1;  // exact
}

// float32 weights, the reference's operation order
__device__ __forceinline__ void pcs_weights(float h, float* wts) {
#pragma clang fp contract(off)
  const float h2 = h * h;
  const float t = 1.0f - h;
  const float wm2 = (t * t) * t;
  const float wm1 = 4.0f + (3.0f * h - 6.0f) * h2;
  const float wp2 = h2 * h;
  const float wp1 = ((6.0f - wm2) - wm1) - wp2;
  wts[0] = wm2; wts[1] = wm1; wts[2] = wp1; wts[3] = wp2;
}

// float64 weights of the same float32 cell offset (FP64 pipeline)
__device__ __forceinline__ void pcs_weights(float hf, double* wts) {
  const double h = (double)hf;
  const double h2 = h * h;
  const double t = 1.0 - h;
  const double wm2 = t * t * t;
  const double wm1 = 4.0 + (3.0 * h - 6.0) * h2;
  const double wp2 = h2 * h;
  wts[0] = wm2; wts[1] = wm1; wts[2] = 6.0 - wm2 - wm1 - wp2; wts[3] = wp2;
}

template <typename W> __device__ __forceinline__ W wmul(W a, W b);
template <> __device__ __forceinline__ float wmul<float>(float a, float b) { return a * b; }
template <> __device__ __forceinline__ double wmul<double>(double a, double b) { return a * b; }

__device__ __forceinline__ float clip_pos(float v, const PcsMap& pm) {
  // np.clip to [lo, hi]; NaN is sent to lo so no cell index leaves the grid
  v = (v >= pm.lo) ? v : pm.lo;
  return (v <= pm.hi) ? v : pm.hi;
}

// The grid is assigned in tiles of whole x lines: a tile is a ty x tz block
// of (y, z) lines, ty*tz no more than the FFT engine's lines per tile.
// Particles are sorted by the tile of their grid-A base cell; the workgroup
// of a tile visits the particles of every tile whose stencils can reach it
// and adds the contributions to its own cells as int64 fixed point (scale
// 2^e) with LDS atomics -- integer sums, so the grid does not depend on the
// order of the adds -- then converts to the working precision and stores
// each line once (in FFT_periodic after the x pass of the FFT).
struct AssignTiling {
  int ty, tz, nty, ntz;  // lines per tile along y and z; tiles along y and z
};

// At most 8 lines per tile keeps the LDS of a tile (16 B per cell plus the
// complex lines) small enough for several workgroups per CU.
#ifndef PSX_ASSIGN_LINES
#define PSX_ASSIGN_LINES 8
#endif
#ifndef PSX_ASSIGN_TY
#define PSX_ASSIGN_TY 4
#endif
constexpr int ASSIGN_LINES = PSX_ASSIGN_LINES;
template <typename T, int NC>
constexpr int assign_lines_max() {
  return Tile<T, NC>::MAXLINES < ASSIGN_LINES ? Tile<T, NC>::MAXLINES : ASSIGN_LINES;
}

template <typename T>
static AssignTiling assign_tiling(int N) {
  const int tl = std::min(tile_lines_host<T>(N), ASSIGN_LINES);
  AssignTiling a;
  a.ty = tl >= PSX_ASSIGN_TY ? PSX_ASSIGN_TY : tl;
  a.tz = tl / a.ty;
  a.nty = (N + a.ty - 1) / a.ty;
  a.ntz = (N + a.tz - 1) / a.tz;
  return a;
}

// Particles are ordered by assignment tile with a counting sort: count per
// tile (the atomic's return value is the particle's rank in its tile), an
// exclusive scan of the counts, and a scatter of (x, y, z, w) records to
// offset + rank.  The order inside a tile depends on the atomics, which does
// not matter: the tile's workgroup adds integer contributions.
__device__ __forceinline__ unsigned tile_key(float y, float z, int N, const PcsMap& pm, AssignTiling at) {
  int cy, cz;
  pcs_coord(y, pm, false, N, &cy);
  pcs_coord(z, pm, false, N, &cz);
  return (unsigned)(cz / at.tz) * (unsigned)at.nty + (unsigned)(cy / at.ty);
}

__global__ void k_fill_u32(unsigned* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// The boxes of a batched assignment (kernel argument): box b's particles,
// its position map and fixed-point scale, and the offset of its particles
// in the sort buffers.  Tile counts and offsets are per box back to back.
struct AssignBoxes {
  const float* xyz[PSX_MAX_BATCH];
  const float* w[PSX_MAX_BATCH];
  int64_t np[PSX_MAX_BATCH];
  int64_t poff[PSX_MAX_BATCH];
  PcsMap pm[PSX_MAX_BATCH];
  double scale[PSX_MAX_BATCH];
};

__global__ void k_tile_count(const AssignBoxes bs, int N, AssignTiling at,
                             unsigned* __restrict__ count, unsigned* __restrict__ key,
                             unsigned* __restrict__ rank) {
  const int b = blockIdx.y;
  const int64_t np = bs.np[b];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float* xyz = bs.xyz[b];
  const PcsMap pm = bs.pm[b];
  count += (int64_t)b * at.nty * at.ntz;
  const unsigned k = tile_key(clip_pos(xyz[np + i], pm), clip_pos(xyz[2 * np + i], pm), N, pm, at);
  key[bs.poff[b] + i] = k;
  rank[bs.poff[b] + i] = atomicAdd(count + k, 1u);
}

// Exclusive scan of n counts into off[0..n] (off[n] = total), one workgroup:
// thread t owns the contiguous segment [t*seg, (t+1)*seg) and keeps up to
// SCAN_REG of its counts in registers (one round of loads, no re-read); the
// segment sums are scanned across each wave with shuffles and the 16 wave
// totals once in LDS.  Box blockIdx.y: count [box][n], off [box][n + 1].
constexpr int SCAN_THREADS = 1024, SCAN_WAVES = SCAN_THREADS / 64, SCAN_REG = 32;
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_scan(const unsigned* __restrict__ count,
                                                            int n, unsigned* __restrict__ off) {
  count += (int64_t)blockIdx.y * n;
  off += (int64_t)blockIdx.y * (n + 1);
  __shared__ unsigned wsum[SCAN_WAVES];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int seg = (n + SCAN_THREADS - 1) / SCAN_THREADS;
  const int b = min(n, t * seg), e = min(n, b + seg);
  unsigned c[SCAN_REG];
  unsigned s = 0;
  if (seg <= SCAN_REG) {
#pragma unroll
    for (int k = 0; k < SCAN_REG; ++k) c[k] = b + k < e ? count[b + k] : 0u;
#pragma unroll
    for (int k = 0; k < SCAN_REG; ++k) s += c[k];
  } else {
    for (int i = b; i < e; ++i) s += count[i];
  }
  // inclusive scan of the segment sums across the wave, then across waves
  unsigned inc = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned u = __shfl_up(inc, d, 64);
    if (lane >= d) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  unsigned before = 0;  // totals of the earlier waves
  for (int w = 0; w < wave; ++w) before += wsum[w];
  unsigned run = before + inc - s;  // exclusive prefix of this segment
  if (seg <= SCAN_REG) {
#pragma unroll
    for (int k = 0; k < SCAN_REG; ++k)
      if (b + k < e) { off[b + k] = run; run += c[k]; }
  } else {
    for (int i = b; i < e; ++i) { off[i] = run; run += count[i]; }
  }
  if (t == SCAN_THREADS - 1) off[n] = before + inc;
}

__global__ void k_tile_scatter(const AssignBoxes bs, const unsigned* __restrict__ key,
                               const unsigned* __restrict__ rank, const unsigned* __restrict__ off,
                               int ntiles, float4* __restrict__ rec) {
  const int b = blockIdx.y;
  const int64_t np = bs.np[b];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float* xyz = bs.xyz[b];
  const float* w = bs.w[b];
  const PcsMap pm = bs.pm[b];
  const int64_t p = bs.poff[b] + i;
  off += (int64_t)b * (ntiles + 1);
  rec[bs.poff[b] + (int64_t)off[key[p]] + rank[p]] =
      make_float4(clip_pos(xyz[i], pm), clip_pos(xyz[np + i], pm), clip_pos(xyz[2 * np + i], pm),
                  w ? w[i] : 1.0f);
}

// Tiles (width tw) along one axis holding the base cells c in [t0-3, t1]
// (cyclic) whose stencils reach the cells [t0, t1): grid A covers c-1..c+2
// and grid B (base c or c+1) c-1..c+3.  At most t1-t0+4 <= 8 tiles.
__device__ int source_tiles(int t0, int t1, int N, int tw, int* out) {
  int n = 0;
  const int span = min(t1 - t0 + 4, N);
  for (int k = 0; k < span; ++k) {
    const int c = ((t0 - 3 + k) % N + N) % N;
    const int t = c / tw;
    bool seen = false;
    for (int u = 0; u < n; ++u) seen = seen || out[u] == t;
    if (!seen) out[n++] = t;
  }
  return n;
}

__device__ __forceinline__ int wrap1(int c, int N) {
  // c in [-1, 2N): one correction
  return c < 0 ? c + N : (c >= N ? c - N : c);
}

// LDS: int64 accumulators acc[line][x][grid] (16 B per cell); with FFTX the
// same memory then holds the complex lines (row stride N+1) and the twiddles.
// Box blockIdx.y of a batch: its sorted records (from bs.poff), tile
// offsets [box][ntiles + 1] and grid out [box][N^3].
template <typename T, int NC, bool FFTX>
__global__ __launch_bounds__(FFT_THREADS) void k_assign_lines(
    const float4* __restrict__ rec, const unsigned* __restrict__ tile_off, const AssignBoxes bs,
    FftShape sh, AssignTiling at, const cplx<T>* __restrict__ tw, cplx<T>* __restrict__ out) {
  const int box = blockIdx.y;
  const int64_t np = bs.np[box];
  const PcsMap pm = bs.pm[box];
  const double scale = bs.scale[box];
  rec += bs.poff[box];
  tile_off += (int64_t)box * (at.nty * at.ntz + 1);
  out += (int64_t)box * n_of<NC>(sh) * n_of<NC>(sh) * n_of<NC>(sh);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem_raw);
  __shared__ int src[64];
  __shared__ int64_t sbeg[64];
  __shared__ int64_t pre[65];
  __shared__ int nsrc_s;
  const int N = n_of<NC>(sh);
  const int y0 = (int)(blockIdx.x % at.nty) * at.ty, z0 = (int)(blockIdx.x / at.nty) * at.tz;
  const int y1 = min(y0 + at.ty, N), z1 = min(z0 + at.tz, N);
  const int nl = at.ty * at.tz;
  const int ncell = nl * N;
  for (int i = threadIdx.x; i < 2 * ncell; i += FFT_THREADS) acc[i] = 0ull;
  if (threadIdx.x == 0) {
    int sy[8], sz[8];
    const int ny = source_tiles(y0, y1, N, at.ty, sy);
    const int nz = source_tiles(z0, z1, N, at.tz, sz);
    int n = 0;
    for (int b = 0; b < nz; ++b)
      for (int a = 0; a < ny; ++a) src[n++] = sz[b] * at.nty + sy[a];
    nsrc_s = n;
  }
  __syncthreads();
  const int nsrc = nsrc_s;
  if ((int)threadIdx.x < nsrc) {
    const unsigned key = (unsigned)src[threadIdx.x];
    // records of tile key: [tile_off[key], tile_off[key + 1])
    sbeg[threadIdx.x] = np > 0 ? (int64_t)tile_off[key] : 0;
    pre[threadIdx.x + 1] = np > 0 ? (int64_t)tile_off[key + 1] : 0;  // range end for now
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    pre[0] = 0;
    for (int u = 0; u < nsrc; ++u) {
      s += pre[u + 1] - sbeg[u];
      pre[u + 1] = s;
    }
  }
  __syncthreads();
  const int64_t total = pre[nsrc];
  int s = 0;
  constexpr int U = 4;  // particle records loaded ahead of their use
  for (int64_t q0 = threadIdx.x; q0 < total; q0 += FFT_THREADS * U) {
    float4 rq[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = q0 + (int64_t)u * FFT_THREADS;
      if (q < total) {
        while (q >= pre[s + 1]) ++s;
        rq[u] = rec[sbeg[s] + (q - pre[s])];
      }
    }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (q0 + (int64_t)u * FFT_THREADS >= total) break;
    const float4 r = rq[u];
    const float p[3] = {r.x, r.y, r.z};
    const T we = (T)r.w;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      int cy, cz;
      const float hy = pcs_coord(p[1], pm, g == 1, N, &cy);
      const float hz = pcs_coord(p[2], pm, g == 1, N, &cz);
      int ly[4], lz[4];
      unsigned my = 0, mz = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int yy = wrap1(cy + a - 1, N), zz = wrap1(cz + a - 1, N);
        ly[a] = yy - y0;
        lz[a] = zz - z0;
        if (yy >= y0 && yy < y1) my |= 1u << a;
        if (zz >= z0 && zz < z1) mz |= 1u << a;
      }
      if (my == 0 || mz == 0) continue;
      int cx;
      T wx[4], wy[4], wz[4];
      pcs_weights(pcs_coord(p[0], pm, g == 1, N, &cx), wx);
      pcs_weights(hy, wy);
      pcs_weights(hz, wz);
      int ix[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) ix[a] = wrap1(cx + a - 1, N);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (!((mz >> c) & 1u)) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (!((my >> b) & 1u)) continue;
          const int row = (lz[c] * at.ty + ly[b]) * N;
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            // estimator.f:380-508: hx*hy*hz*we, left to right
            const T v = wmul(wmul(wmul(wx[a], wy[b]), wz[c]), we);
            const long long qv = __double2ll_rn((double)v * scale);
            atomicAdd(acc + 2 * (row + ix[a]) + g, (unsigned long long)qv);
          }
        }
      }
    }
  }
  }
  __syncthreads();
  const double inv = 1.0 / scale;
  if constexpr (!FFTX) {
    for (int e = threadIdx.x; e < ncell; e += FFT_THREADS) {
      const int line = e / N, x = e - line * N;
      const int y = y0 + line % at.ty, z = z0 + line / at.ty;
      if (y < y1 && z < z1)
        out[((int64_t)z * N + y) * N + x] = mk((T)((double)(long long)acc[2 * e] * inv),
                                               (T)((double)(long long)acc[2 * e + 1] * inv));
    }
  } else {
    // Convert in place to complex lines of stride N+1.  Cell e moves from
    // byte 16e to byte sizeof(cplx<T>)*(e + line): for 8-byte complex the
    // destination lies in cells already read when rounds go up, for 16-byte
    // complex in cells already read when rounds go down.
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
    const int lstride = N + 1;
    const int nround = (ncell + FFT_THREADS - 1) / FFT_THREADS;
    for (int r = 0; r < nround; ++r) {
      const int rr = sizeof(cplx<T>) < 16 ? r : nround - 1 - r;
      const int e = rr * FFT_THREADS + threadIdx.x;
      cplx<T> v = mk(T(0), T(0));
      if (e < ncell)
        v = mk((T)((double)(long long)acc[2 * e] * inv), (T)((double)(long long)acc[2 * e + 1] * inv));
      __syncthreads();
      if (e < ncell) {
        const int line = e / N;
        lds[e + line] = v;
      }
      __syncthreads();
    }
    cplx<T>* twl = lds + nl * lstride;
    stage_twiddles(twl, tw, N);
    __syncthreads();
    fft_tile<NC, 1, T, assign_lines_max<T, NC>(), TileCfg<T>::ELEMS>(lds, sh, nl, lstride, twl);
    for (int e = threadIdx.x; e < ncell; e += FFT_THREADS) {
      const int line = e / N, x = e - line * N;
      const int y = y0 + line % at.ty, z = z0 + line / at.ty;
      if (y < y1 && z < z1) out[((int64_t)z * N + y) * N + x] = lds[line * lstride + x];
    }
  }
}

// Fixed-point exponent: every cell is bounded by 64 * sum|w| (per-axis PCS
// weight <= 4), keep that below 2^62.
int fixed_point_exponent(double wsum_abs) {
  double bound = 64.0 * (wsum_abs > 1.0 ? wsum_abs : 1.0);
  int e = 62 - (int)ceil(log2(bound));
  if (e > 40) e = 40;
  if (e < 0) e = 0;
  return e;
}

template <typename T, int NC>
static int assign_lines_n(psx_ctx* ctx, const float4* rec, const unsigned* off,
                          const AssignBoxes& bs, int nbox, int N, AssignTiling at, cplx<T>* grid,
                          bool fft_x) {
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  const size_t nl = (size_t)at.ty * at.tz;
  const size_t acc_bytes = 16 * nl * N;
  const dim3 blocks((unsigned)(at.nty * at.ntz), (unsigned)nbox);
  if (fft_x) {
    const size_t fft_bytes = sizeof(cplx<T>) * (nl * (N + 1) + N);
    const size_t lds = acc_bytes > fft_bytes ? acc_bytes : fft_bytes;
    auto kern = k_assign_lines<T, NC, true>;
    PSX_TRY(set_lds(kern, lds));
    launch_timed(ctx, "k_assign_lines", kern, blocks, dim3(FFT_THREADS), lds, rec, off, bs,
                 p->shape, at, plan_tw<T>(p), grid);
  } else {
    auto kern = k_assign_lines<T, 0, false>;
    PSX_TRY(set_lds(kern, acc_bytes));
    launch_timed(ctx, "k_assign_lines", kern, blocks, dim3(FFT_THREADS), acc_bytes, rec, off, bs,
                 p->shape, at, plan_tw<T>(p), grid);
  }
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

#define PSX_CASE_N(n)                                                                         \
  case n:                                                                                     \
    return assign_lines_n<T, n>(ctx, rec, o, bs, nbox, N, at, grid, fft_x);

// nbox boxes (particles xyz[b], weights w[b] or null, counts np[b], maps
// pm[b], fixed-point scales scale[b]) into nbox grids back to back.
template <typename T>
static int assign_grid_t(psx_ctx* ctx, int nbox, const float* const* xyz, const float* const* w,
                         const int64_t* np, const PcsMap* pm, const double* scale, int N,
                         cplx<T>* grid, bool fft_x) {
  PSX_REQUIRE(nbox >= 1 && nbox <= PSX_MAX_BATCH, "batch of 1 to 16 boxes");
  const AssignTiling at = assign_tiling<T>(N);
  AssignBoxes bs{};
  int64_t total = 0, maxnp = 0;
  for (int b = 0; b < nbox; ++b) {
    PSX_REQUIRE(np[b] >= 0 && np[b] < (int64_t)1 << 32, "more than 2^32 - 1 particles");
    bs.xyz[b] = xyz[b];
    bs.w[b] = w ? w[b] : nullptr;
    bs.np[b] = np[b];
    bs.poff[b] = total;
    bs.pm[b] = pm[b];
    bs.scale[b] = scale[b];
    total += np[b];
    maxnp = std::max(maxnp, np[b]);
  }
  const int64_t ntiles = (int64_t)at.nty * at.ntz;
  const float4* rec = nullptr;
  unsigned* o = nullptr;
  PSX_TRY(ctx_buffer(ctx, "tile_off", (size_t)nbox * (ntiles + 1) * 4, (void**)&o));
  if (total > 0) {
    // counting sort by assignment tile into (x, y, z, w) records, per box
    unsigned *count, *key, *rank;
    float4* r;
    PSX_TRY(ctx_buffer(ctx, "tile_count", (size_t)nbox * ntiles * 4, (void**)&count));
    PSX_TRY(ctx_buffer(ctx, "particle_tile", (size_t)total * 4, (void**)&key));
    PSX_TRY(ctx_buffer(ctx, "particle_rank", (size_t)total * 4, (void**)&rank));
    PSX_TRY(ctx_buffer(ctx, "sorted_particles", (size_t)total * sizeof(float4), (void**)&r));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)cdiv(nbox * ntiles, 256)), dim3(256), 0,
                       ctx->stream, count, nbox * ntiles);
    PSX_CHECK_LAUNCH();
    const dim3 pgrid((unsigned)cdiv(maxnp, 256), (unsigned)nbox);
    hipLaunchKernelGGL(k_tile_count, pgrid, dim3(256), 0, ctx->stream, bs, N, at, count, key, rank);
    PSX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_tile_scan, dim3(1, nbox), dim3(SCAN_THREADS), 0, ctx->stream,
                       (const unsigned*)count, (int)ntiles, o);
    PSX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_tile_scatter, pgrid, dim3(256), 0, ctx->stream, bs, (const unsigned*)key,
                       (const unsigned*)rank, (const unsigned*)o, (int)ntiles, r);
    PSX_CHECK_LAUNCH();
    rec = r;
  }
  // boxes without particles read no records (their np is 0)
  if (!fft_x) return assign_lines_n<T, 0>(ctx, rec, o, bs, nbox, N, at, grid, false);
  switch (N) {
    PSX_FFT_SIZES(PSX_CASE_N)
    default: return assign_lines_n<T, 0>(ctx, rec, o, bs, nbox, N, at, grid, true);
  }
}
#undef PSX_CASE_N

int launch_assign_grid_batch(psx_ctx* ctx, int nbox, const float* const* xyz,
                             const float* const* w, const int64_t* np, const PcsMap* pm,
                             const double* scale, int N, void* grid, int prec, bool fft_x) {
  if (prec == PSX_PREC_F64)
    return assign_grid_t<double>(ctx, nbox, xyz, w, np, pm, scale, N, (cplx<double>*)grid, fft_x);
  return assign_grid_t<float>(ctx, nbox, xyz, w, np, pm, scale, N, (cplx<float>*)grid, fft_x);
}

int launch_assign_grid(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int N,
                       PcsMap pm, double scale, void* grid, int prec, bool fft_x) {
  return launch_assign_grid_batch(ctx, 1, &xyz, &w, &np, &pm, &scale, N, grid, prec, fft_x);
}

// Unit field delta = 1 (the shell indicators of _counts_Bk123), complex double.
__global__ void k_fill_unit(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    p[2 * i] = 1.0;
    p[2 * i + 1] = 0.0;
  }
}

int launch_fill_unit(psx_ctx* ctx, double* half, int64_t n) {
  hipLaunchKernelGGL(k_fill_unit, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 8192)), dim3(256),
                     0, ctx->stream, half, n);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

// --------------------------------------------------------------- fcomb ----
// estimator.f:605-675.  One thread per loop iteration (ix, iy, iz) in
// [0, N/2]^3.  The eight elements {ix,-ix} x {iy,-iy} x {iz,-iz} are read,
// combined, and the eight stores are replayed in the reference's order on a
// register copy in which aliased indices (0 and N/2) share a slot, so the
// value that survives is the reference's.  Only the x <= N/2 elements are
// written (the half field).  rec[n] is the reference's twiddle recurrence
// xrec = recx^n (complex double); W is the per-axis PCS window.
template <typename T> __device__ __forceinline__ T window_val(const float* Wf, const double* Wd, int n);
template <> __device__ __forceinline__ float window_val<float>(const float* Wf, const double*, int n) { return Wf[n]; }
template <> __device__ __forceinline__ double window_val<double>(const float*, const double* Wd, int n) { return Wd[n]; }

template <typename T>
__device__ __forceinline__ cplx<T> slot_get(const cplx<T> (&m)[8], int s) {
  cplx<T> r = m[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) if (s == k) r = m[k];
  return r;
}
template <typename T>
__device__ __forceinline__ void slot_put(cplx<T> (&m)[8], int s, cplx<T> v) {
#pragma unroll
  for (int k = 0; k < 8; ++k) if (s == k) m[k] = v;
}

// Normalisations of the boxes of a batch (kernel argument).
template <typename T> struct FcombCoef {
  T cf[PSX_MAX_BATCH];
};

// Iteration (ix, iy, iz) of estimator.f:646-665: from the eight values v[s]
// at (xs[s>>2], ys[(s>>1)&1], zs[s&1]) (xs = {ix, -ix mod N}, ...) the
// eight stores in the reference's order; m[canon(xb, yb, zb)] is then the
// value the iteration leaves at (xs[xb], ys[yb], zs[zb]) -- where indices
// alias (0 and N/2) the last store wins.
struct FcombSlots {
  bool ax, ay, az;
  __device__ __forceinline__ int canon(int xb, int yb, int zb) const {
    return ((xb && !ax) ? 4 : 0) + ((yb && !ay) ? 2 : 0) + ((zb && !az) ? 1 : 0);
  }
};

template <typename T>
__device__ __forceinline__ void fcomb_iteration(cplx<T> (&m)[8], const FcombSlots& sl, int ix,
                                                int iy, int iz, T cf, const c64* __restrict__ rec,
                                                const float* __restrict__ Wf,
                                                const double* __restrict__ Wd) {
  // twiddles in double (estimator.f:650-653), rounded to T
  const c64 xr = rec[ix], yr = rec[iy], zr = rec[iz];
  const c64 ixr = mk(-xr.im, xr.re);
  const c64 cma = ixr * yr * zr;
  const c64 cmb = ixr * yr * conj(zr);
  const c64 cmc = ixr * conj(yr) * zr;
  const c64 cmd = ixr * conj(yr * zr);
  const cplx<T> one = mk(T(1), T(0));
  auto cv = [](c64 a) { return mk((T)a.re, (T)a.im); };
  const cplx<T> A = cv(cma), B = cv(cmb), C = cv(cmc), D = cv(cmd);
  const T cfac = cf / (window_val<T>(Wf, Wd, ix) * window_val<T>(Wf, Wd, iy) * window_val<T>(Wf, Wd, iz));
  const cplx<T> c000 = m[0] * (one - A) + conj(m[7]) * (one + A);
  const cplx<T> c001 = m[1] * (one - B) + conj(m[6]) * (one + B);
  const cplx<T> c010 = m[2] * (one - C) + conj(m[5]) * (one + C);
  const cplx<T> c011 = m[3] * (one - D) + conj(m[4]) * (one + D);
  slot_put(m, sl.canon(0, 0, 0), c000 * cfac);
  slot_put(m, sl.canon(0, 0, 1), c001 * cfac);
  slot_put(m, sl.canon(0, 1, 0), c010 * cfac);
  slot_put(m, sl.canon(0, 1, 1), c011 * cfac);
  slot_put(m, sl.canon(1, 0, 0), conj(slot_get(m, sl.canon(0, 1, 1))));
  slot_put(m, sl.canon(1, 0, 1), conj(slot_get(m, sl.canon(0, 1, 0))));
  slot_put(m, sl.canon(1, 1, 0), conj(slot_get(m, sl.canon(0, 0, 1))));
  slot_put(m, sl.canon(1, 1, 1), conj(slot_get(m, sl.canon(0, 0, 0))));
}

// Box blockIdx.y of a batch: grid F [box][N^3], half field Hh [box][N^2 (N/2+1)].
template <typename T>
__global__ __launch_bounds__(256) void k_fcomb(const cplx<T>* __restrict__ F, cplx<T>* __restrict__ Hh, int N,
                                               FcombCoef<T> coef, const c64* __restrict__ rec,
                                               const float* __restrict__ Wf,
                                               const double* __restrict__ Wd) {
  const int Nh = N / 2 + 1;
  F += (int64_t)blockIdx.y * N * N * N;
  Hh += (int64_t)blockIdx.y * N * N * Nh;
  const T cf = coef.cf[blockIdx.y];
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)Nh * Nh * Nh) return;
  const int ix = (int)(id % Nh);
  const int iy = (int)((id / Nh) % Nh);
  const int iz = (int)(id / ((int64_t)Nh * Nh));
  const int xs[2] = {ix, (N - ix) % N};
  const int ys[2] = {iy, (N - iy) % N};
  const int zs[2] = {iz, (N - iz) % N};
  const FcombSlots sl{xs[0] == xs[1], ys[0] == ys[1], zs[0] == zs[1]};
  cplx<T> m[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int xb = s >> 2, yb = (s >> 1) & 1, zb = s & 1;
    m[s] = F[((int64_t)zs[zb] * N + ys[yb]) * N + xs[xb]];
  }
  fcomb_iteration(m, sl, ix, iy, iz, cf, rec, Wf, Wd);
  // half-field elements owned by this iteration: x = ix, y in {iy,-iy}, z in {iz,-iz}
#pragma unroll
  for (int yb = 0; yb < 2; ++yb) {
#pragma unroll
    for (int zb = 0; zb < 2; ++zb) {
      if ((yb && sl.ay) || (zb && sl.az)) continue;
      Hh[((int64_t)zs[zb] * N + ys[yb]) * Nh + ix] = slot_get(m, sl.canon(0, yb, zb));
    }
  }
}

// The z pass of the backward grid FFT with fcomb as its epilogue.  The
// half-field line (x, y), x <= N/2, is the fcomb of the lines (x, y) and
// (-x, -y) after their z transforms: every value the store order of
// iteration (x, |y|, |z|) reads or leaves at (x, y, z) lies on one of them
// (where x, y or z alias, the other lines of the iteration coincide with
// these two).  A tile holds tp such line pairs, consecutive x at one y; the
// line transforms are those of k_fft_lines, the combination that of k_fcomb.
template <typename T, int NC>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_zcomb(const cplx<T>* __restrict__ F,
                                                            cplx<T>* __restrict__ Hh, FftShape sh,
                                                            const cplx<T>* __restrict__ tw,
                                                            FcombCoef<T> coef,
                                                            const c64* __restrict__ rec,
                                                            const float* __restrict__ Wf,
                                                            const double* __restrict__ Wd) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int N = n_of<NC>(sh), Nh = N / 2 + 1, lstride = N + 1;
  const int64_t NN = (int64_t)N * N;
  F += (int64_t)blockIdx.y * NN * N;
  Hh += (int64_t)blockIdx.y * NN * Nh;
  const T cf = coef.cf[blockIdx.y];
  constexpr int ML = Tile<T, NC>::MAXLINES;
  const int tl = NC > 0 ? ML : tile_lines_for(N, TileCfg<T>::ELEMS, TileCfg<T>::LINES);
  const int tp = tl / 2;  // line pairs per tile
  const int ntx = (Nh + tp - 1) / tp;
  // neighbouring x tiles read the two halves of the same 128-byte lines
  const int bx = xcd_local_block((int)blockIdx.x, (int)gridDim.x);
  const int y = bx / ntx;
  const int x0 = (bx - y * ntx) * tp;
  const int np = min(tp, Nh - x0);
  const int yb_line = (N - y) % N;
  // line 2c: (x0 + c, y); line 2c + 1: (-(x0 + c), -y); element n along z
  batched_fill<8, cplx<T>>(
      2 * tp * N,
      [&](int idx) {
        const int n = idx / (2 * tp), l = idx - n * (2 * tp);
        const int c = l >> 1;
        if (c >= np) return mk(T(0), T(0));
        const int x = (l & 1) ? (N - (x0 + c)) % N : x0 + c;
        const int yy = (l & 1) ? yb_line : y;
        return F[((int64_t)n * N + yy) * N + x];
      },
      [&](int idx, cplx<T> v) {
        const int n = idx / (2 * tp), l = idx - n * (2 * tp);
        if ((l >> 1) < np) lds[l * lstride + n] = v;
      });
  cplx<T>* twl = lds + tl * lstride;
  stage_twiddles(twl, tw, N);
  __syncthreads();
  fft_tile<NC, 1, T, ML, TileCfg<T>::ELEMS>(lds, sh, 2 * np, lstride, twl);
  const int yb = y > N / 2 ? 1 : 0, iy = yb ? N - y : y;
  for (int idx = threadIdx.x; idx < np * N; idx += FFT_THREADS) {
    const int z = idx / np, c = idx - z * np;
    const int ix = x0 + c;
    const int zb = z > N / 2 ? 1 : 0, iz = zb ? N - z : z;
    const int xs[2] = {ix, (N - ix) % N};
    const int ys[2] = {iy, (N - iy) % N};
    const int zs[2] = {iz, (N - iz) % N};
    const FcombSlots sl{xs[0] == xs[1], ys[0] == ys[1], zs[0] == zs[1]};
    const cplx<T>* la = lds + (2 * c) * lstride;      // (ix, y)
    const cplx<T>* lb = lds + (2 * c + 1) * lstride;  // (-ix, -y)
    cplx<T> out;
    if (!(sl.ax || sl.ay || sl.az)) {
      // no index aliases: the element keeps its own combination c_{0,yb,zb}
      // = v (1 - t) + conj(v') (1 + t) times cfac, v = F(k), v' = F(-k), with
      // the twiddle t of k_fcomb's cma/cmb/cmc/cmd (the same operations)
      const c64 xr = rec[ix], yr = rec[iy], zr = rec[iz];
      const c64 ixr = mk(-xr.im, xr.re);
      const c64 yv = yb ? conj(yr) : yr;
      c64 t64;
      if (yb && zb) t64 = ixr * conj(yr * zr);
      else t64 = ixr * yv * (zb ? conj(zr) : zr);
      const cplx<T> t = mk((T)t64.re, (T)t64.im);
      const cplx<T> one = mk(T(1), T(0));
      const T cfac = cf / (window_val<T>(Wf, Wd, ix) * window_val<T>(Wf, Wd, iy) *
                           window_val<T>(Wf, Wd, iz));
      out = (la[z] * (one - t) + conj(lb[(N - z) % N]) * (one + t)) * cfac;
    } else {
      cplx<T> m[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int X = xs[s >> 2], Y = ys[(s >> 1) & 1], Z = zs[s & 1];
        // every (X, Y) the iteration needs here is one of the two lines
        m[s] = (X == ix && Y == y) ? la[Z] : lb[Z];
      }
      fcomb_iteration(m, sl, ix, iy, iz, cf, rec, Wf, Wd);
      out = slot_get(m, sl.canon(0, yb, zb));
    }
    Hh[((int64_t)z * N + y) * Nh + ix] = out;
  }
}

template <typename T, int NC>
static int zcomb_n(psx_ctx* ctx, const FftPlan* p, const cplx<T>* F, cplx<T>* Hh, int nbox,
                   const FcombCoef<T>& c) {
  const int N = p->shape.N, Nh = N / 2 + 1;
  const int tl = tile_lines_host<T>(N), tp = tl / 2;
  const size_t lds = ((size_t)tl * (N + 1) + N) * sizeof(cplx<T>);
  auto kern = k_fft_zcomb<T, NC>;
  PSX_TRY(set_lds(kern, lds));
  const dim3 grid((unsigned)((int64_t)N * ((Nh + tp - 1) / tp)), (unsigned)nbox);
  launch_timed(ctx, "k_fft_zcomb", kern, grid, dim3(FFT_THREADS), lds, F, Hh, p->shape,
               plan_tw<T>(p), c, (const c64*)p->rec, p->Wf, p->Wd);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

#define PSX_CASE_N(n) \
  case n: return zcomb_n<T, n>(ctx, p, F, Hh, nbox, c);

template <typename T>
static int zcomb_t(psx_ctx* ctx, const FftPlan* p, const cplx<T>* F, cplx<T>* Hh, int nbox,
                   const FcombCoef<T>& c) {
  switch (p->shape.N) {
    PSX_FFT_SIZES(PSX_CASE_N)
    default: return zcomb_n<T, 0>(ctx, p, F, Hh, nbox, c);
  }
}
#undef PSX_CASE_N

int launch_fft_zcomb_batch(psx_ctx* ctx, const void* F, void* Hh, int N, int nbox,
                           const double* sumw, int prec) {
  PSX_REQUIRE(nbox >= 1 && nbox <= PSX_MAX_BATCH, "batch of 1 to 16 boxes");
  PSX_REQUIRE(tile_lines_host<double>(N) >= 2, "grid too large for the fused z pass");
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  if (prec == PSX_PREC_F64) {
    FcombCoef<double> c{};
    // cf = 1/(6^3 * 4 * N) with N = sum w as float32 (estimator.f:615)
    for (int b = 0; b < nbox; ++b) c.cf[b] = 1.0 / (216.0 * 4.0 * (double)(float)sumw[b]);
    return zcomb_t<double>(ctx, p, (const cplx<double>*)F, (cplx<double>*)Hh, nbox, c);
  }
  FcombCoef<float> c{};
  for (int b = 0; b < nbox; ++b) c.cf[b] = 1.0f / (864.0f * (float)sumw[b]);
  return zcomb_t<float>(ctx, p, (const cplx<float>*)F, (cplx<float>*)Hh, nbox, c);
}

// nbox grids back to back -> nbox half fields back to back.
int launch_fcomb_batch(psx_ctx* ctx, const void* F, void* Hh, int N, int nbox, const double* sumw,
                       int prec, const void* rec, const float* Wf, const double* Wd) {
  PSX_REQUIRE(nbox >= 1 && nbox <= PSX_MAX_BATCH, "batch of 1 to 16 boxes");
  const int64_t Nh = N / 2 + 1;
  const int64_t n = Nh * Nh * Nh;
  const dim3 grid((unsigned)cdiv(n, 256), nbox);
  if (prec == PSX_PREC_F64) {
    FcombCoef<double> c{};
    // cf = 1/(6^3 * 4 * N) with N = sum w as float32 (estimator.f:615)
    for (int b = 0; b < nbox; ++b) c.cf[b] = 1.0 / (216.0 * 4.0 * (double)(float)sumw[b]);
    hipLaunchKernelGGL(k_fcomb<double>, grid, dim3(256), 0, ctx->stream, (const cplx<double>*)F,
                       (cplx<double>*)Hh, N, c, (const c64*)rec, Wf, Wd);
  } else {
    FcombCoef<float> c{};
    for (int b = 0; b < nbox; ++b) c.cf[b] = 1.0f / (864.0f * (float)sumw[b]);
    hipLaunchKernelGGL(k_fcomb<float>, grid, dim3(256), 0, ctx->stream, (const cplx<float>*)F,
                       (cplx<float>*)Hh, N, c, (const c64*)rec, Wf, Wd);
  }
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_fcomb(psx_ctx* ctx, const void* F, void* Hh, int N, double sumw, int prec,
                 const void* rec, const float* Wf, const double* Wd) {
  return launch_fcomb_batch(ctx, F, Hh, N, 1, &sumw, prec, rec, Wf, Wd);
}

// ------------------------------------------------- half <-> full fields ----
// Hermitian part of an arbitrary full field (any strides, element type TI),
// kept on x <= N/2: H(k) = (F(k) + conj F(-k)) / 2.  Re FFT of a field equals
// the FFT of its Hermitian part, so the shell fields of H equal the
// reference's Re(fftn(F * mask)) exactly in exact arithmetic.
template <typename TI, typename T>
__global__ void k_herm_half(const cplx<TI>* __restrict__ F, int64_t sx, int64_t sy, int64_t sz,
                            cplx<T>* __restrict__ Hh, int N) {
  const int Nh = N / 2 + 1;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)Nh * N * N) return;
  const int x = (int)(id % Nh);
  const int y = (int)((id / Nh) % N);
  const int z = (int)(id / ((int64_t)Nh * N));
  const int mx = (N - x) % N, my = (N - y) % N, mz = (N - z) % N;
  const cplx<TI> a = F[x * sx + y * sy + z * sz];
  const cplx<TI> b = F[mx * sx + my * sy + mz * sz];
  Hh[id] = mk((T)(((double)a.re + (double)b.re) * 0.5), (T)(((double)a.im - (double)b.im) * 0.5));
}

// reflect_delta (pyspectrum.py:1134-1157) to a C-order [ix][iy][iz] full field.
template <typename T>
__global__ void k_reflect(const cplx<T>* __restrict__ Hh, cplx<T>* __restrict__ F, int N) {
  const int Nh = N / 2 + 1, hg = N / 2;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)N * N * N) return;
  const int z = (int)(id % N);
  const int y = (int)((id / N) % N);
  const int x = (int)(id / ((int64_t)N * N));
  cplx<T> v;
  if (x < Nh) {
    v = Hh[((int64_t)z * N + y) * Nh + x];
    const bool cx = (x == 0 || x == hg), cy = (y == 0 || y == hg), cz = (z == 0 || z == hg);
    if (cx && cy && cz && !(x == 0 && y == 0 && z == 0)) v.im = T(0);
  } else {
    const int mx = N - x, my = (N - y) % N, mz = (N - z) % N;
    v = conj(Hh[((int64_t)mz * N + my) * Nh + mx]);
  }
  F[id] = v;
}

int launch_herm_half(psx_ctx* ctx, const void* F, int in_prec, int64_t sx, int64_t sy, int64_t sz,
                     void* Hh, int N, int prec) {
  const int64_t n = (int64_t)(N / 2 + 1) * N * N;
  const dim3 grid((unsigned)cdiv(n, 256));
  if (in_prec == PSX_PREC_F64) {
    if (prec == PSX_PREC_F64)
      hipLaunchKernelGGL((k_herm_half<double, double>), grid, dim3(256), 0, ctx->stream,
                         (const cplx<double>*)F, sx, sy, sz, (cplx<double>*)Hh, N);
    else
      hipLaunchKernelGGL((k_herm_half<double, float>), grid, dim3(256), 0, ctx->stream,
                         (const cplx<double>*)F, sx, sy, sz, (cplx<float>*)Hh, N);
  } else {
    if (prec == PSX_PREC_F64)
      hipLaunchKernelGGL((k_herm_half<float, double>), grid, dim3(256), 0, ctx->stream,
                         (const cplx<float>*)F, sx, sy, sz, (cplx<double>*)Hh, N);
    else
      hipLaunchKernelGGL((k_herm_half<float, float>), grid, dim3(256), 0, ctx->stream,
                         (const cplx<float>*)F, sx, sy, sz, (cplx<float>*)Hh, N);
  }
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_reflect(psx_ctx* ctx, const void* Hh, void* F, int N, int prec) {
  const int64_t n = (int64_t)N * N * N;
  if (prec == PSX_PREC_F64)
    hipLaunchKernelGGL(k_reflect<double>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream,
                       (const cplx<double>*)Hh, (cplx<double>*)F, N);
  else
    hipLaunchKernelGGL(k_reflect<float>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream,
                       (const cplx<float>*)Hh, (cplx<float>*)F, N);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
