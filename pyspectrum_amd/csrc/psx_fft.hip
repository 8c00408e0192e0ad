// 3D FFTs of the periodic-box estimator on gfx950.
//
//  * Data transform (pyspectrum.py:1060-1080): the packed interlaced grid
//    A + iB, backward (e^{+i}), unnormalised, as three batched line passes
//    (x, y, z).  In FFT_periodic the x pass runs inside the assignment kernel
//    (psx_mesh.hip), which holds whole x lines in LDS; y and z follow here.
//  * Shell fields (pyspectrum.py:387-400): d_j(x) = Re FFT_forward[delta 1_j]
//    for every shell j, as a pruned Hermitian (c2r) transform of the half
//    field: a z pass over the (kx, ky) lines that meet the shell, a y pass
//    over kx <= kmax_j columns, and an x pass that builds two Hermitian lines
//    per complex FFT (their real and imaginary parts are the two real outputs).
//
// Every kernel is instantiated for the compile-time sizes of PSX_FFT_SIZES
// (N > 0) and once for a runtime N (N = 0).
//
// Grid layout everywhere: index (z*N + y)*N + x, x fastest; the half field is
// (z*N + y)*(N/2+1) + x.
#include "psx_ctx.h"

namespace psx {

// ------------------------------------------------------------------ passes --
// Batched line FFT.  Line (o, c) of box blockIdx.y starts at
// blockIdx.y*Bs + o*Os + c*Cs, element n at +n*Es.  A tile holds tl
// consecutive inner lines c of one outer index o.
template <typename T, int S, int NC>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_lines(
    const cplx<T>* in, cplx<T>* out, FftShape sh, const cplx<T>* __restrict__ tw, int64_t Os,
    int64_t Cs, int64_t Es, int n_inner, int64_t Bs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int N = n_of<NC>(sh);
  const int lstride = N + 1;
  const int tl = NC > 0 ? Tile<T, NC>::MAXLINES : tile_lines_for(N, TileCfg<T>::ELEMS, TileCfg<T>::LINES);
  const int ntile = (n_inner + tl - 1) / tl;
  const int64_t o = blockIdx.x / ntile;
  const int c0 = (int)(blockIdx.x - o * ntile) * tl;
  const int nl = min(tl, n_inner - c0);
  const int64_t base = (int64_t)blockIdx.y * Bs + o * Os + (int64_t)c0 * Cs;
  // offsets within a tile fit in 32 bits for N <= 1024; tl is a constant for
  // compile-time N, so the index divisions below are multiplications
  const int cs = (int)Cs, es = (int)Es;
  if (es == 1) {
    // x lines: the tile is one contiguous run of nl*N elements
    batched_fill<8, cplx<T>>(
        nl * N, [&](int idx) { return in[base + idx]; },
        [&](int idx, cplx<T> v) {
          const int c = idx / N;
          lds[c * lstride + (idx - c * N)] = v;
        });
  } else {
    batched_fill<8, cplx<T>>(
        tl * N,
        [&](int idx) {
          const int n = idx / tl, c = idx - n * tl;
          return c < nl ? in[base + (c * cs + n * es)] : mk(T(0), T(0));
        },
        [&](int idx, cplx<T> v) {
          const int n = idx / tl, c = idx - n * tl;
          if (c < nl) lds[c * lstride + n] = v;
        });
  }
  cplx<T>* twl = lds + tl * lstride;
  stage_twiddles(twl, tw, N);
  __syncthreads();
  fft_tile<NC, S, T, Tile<T, NC>::MAXLINES, TileCfg<T>::ELEMS>(lds, sh, nl, lstride, twl);
  for (int idx = threadIdx.x; idx < tl * N; idx += FFT_THREADS) {
    int c, n;
    if (es == 1) { c = idx / N; n = idx - c * N; }
    else { n = idx / tl; c = idx - n * tl; }
    if (c >= nl) continue;
    out[base + (c * cs + n * es)] = lds[c * lstride + n];
  }
}

template <typename T, int S, int NC>
static int run_pass_n(psx_ctx* ctx, const FftPlan* p, const cplx<T>* in, cplx<T>* out, int axis,
                      int nbox) {
  const int N = p->shape.N;
  const int64_t NN = (int64_t)N * N;
  int64_t Os, Cs, Es;
  if (axis == 0) { Os = NN; Cs = N; Es = 1; }        // x lines: inner y, outer z
  else if (axis == 1) { Os = NN; Cs = 1; Es = N; }   // y lines: inner x, outer z
  else { Os = N; Cs = 1; Es = NN; }                  // z lines: inner x, outer y
  const int tl = tile_lines_host<T>(N);
  const int64_t nblocks = (int64_t)N * cdiv(N, tl);
  const size_t lds = ((size_t)tl * (N + 1) + N) * sizeof(cplx<T>);
  auto kern = k_fft_lines<T, S, NC>;
  PSX_TRY(set_lds(kern, lds));
  launch_timed(ctx, "k_fft_lines", kern, dim3((unsigned)nblocks, (unsigned)nbox),
               dim3(FFT_THREADS), lds, in, out, p->shape, plan_tw<T>(p), Os, Cs, Es, N, NN * N);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

#define PSX_CASE_N(n)                                                                        \
  case n:                                                                                    \
    return run_pass_n<T, S, n>(ctx, p, in, out, axis, nbox);

template <typename T, int S>
static int run_pass(psx_ctx* ctx, const FftPlan* p, const cplx<T>* in, cplx<T>* out, int axis,
                    int nbox) {
  switch (p->shape.N) {
    PSX_FFT_SIZES(PSX_CASE_N)
    default: return run_pass_n<T, S, 0>(ctx, p, in, out, axis, nbox);
  }
}
#undef PSX_CASE_N

template <typename T, int S>
static int fft3d_t(psx_ctx* ctx, cplx<T>* g, int N, int first_axis, int nbox) {
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  for (int axis = first_axis; axis < 3; ++axis) PSX_TRY((run_pass<T, S>(ctx, p, g, g, axis, nbox)));
  return PSX_OK;
}

int launch_fft3d_axes_batch(psx_ctx* ctx, void* grid, int N, int sign, int prec, int first_axis,
                            int nbox) {
  if (prec == PSX_PREC_F64) {
    return sign > 0 ? fft3d_t<double, 1>(ctx, (cplx<double>*)grid, N, first_axis, nbox)
                    : fft3d_t<double, -1>(ctx, (cplx<double>*)grid, N, first_axis, nbox);
  }
  return sign > 0 ? fft3d_t<float, 1>(ctx, (cplx<float>*)grid, N, first_axis, nbox)
                  : fft3d_t<float, -1>(ctx, (cplx<float>*)grid, N, first_axis, nbox);
}

int launch_fft3d_axes(psx_ctx* ctx, void* grid, int N, int sign, int prec, int first_axis) {
  return launch_fft3d_axes_batch(ctx, grid, N, sign, prec, first_axis, 1);
}

int launch_fft_axis_batch(psx_ctx* ctx, void* grid, int N, int sign, int prec, int axis, int nbox) {
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  if (prec == PSX_PREC_F64) {
    cplx<double>* g = (cplx<double>*)grid;
    return sign > 0 ? run_pass<double, 1>(ctx, p, g, g, axis, nbox)
                    : run_pass<double, -1>(ctx, p, g, g, axis, nbox);
  }
  cplx<float>* g = (cplx<float>*)grid;
  return sign > 0 ? run_pass<float, 1>(ctx, p, g, g, axis, nbox)
                  : run_pass<float, -1>(ctx, p, g, g, axis, nbox);
}

int launch_fft3d(psx_ctx* ctx, void* grid, int N, int sign, int prec) {
  return launch_fft3d_axes(ctx, grid, N, sign, prec, 0);
}

// ------------------------------------------------------------ shell passes --
// meta per slot: {j, rmax, kxw, unused}
struct ShellDev {
  const int* meta;
  const int64_t* off;
  int step;
  int Nh;
};

__device__ __forceinline__ int signed_f(int n, int N) { return n <= N / 2 ? n : n - N; }

// Four consecutive values to global memory with vector stores (p is 4-element aligned).
__device__ __forceinline__ void store4(float* p, const float (&o)[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
}
__device__ __forceinline__ void store4(double* p, const double (&o)[4]) {
  reinterpret_cast<double2*>(p)[0] = make_double2(o[0], o[1]);
  reinterpret_cast<double2*>(p)[1] = make_double2(o[2], o[3]);
}

// z pass: tile (s, ky, kx0): NL kx columns of one ky row, masked to shell j.
template <typename T, int NC>
__global__ __launch_bounds__(FFT_THREADS) void k_shell_z(const cplx<T>* __restrict__ H,
                                                         cplx<T>* __restrict__ T1,
                                                         const Tile3* __restrict__ tiles,
                                                         ShellDev sd, FftShape sh,
                                                         const cplx<T>* __restrict__ tw,
                                                         int64_t hbs, int64_t tbs) {
  // box blockIdx.y of a batch: its half field and intermediate
  H += (int64_t)blockIdx.y * hbs;
  T1 += (int64_t)blockIdx.y * tbs;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int N = n_of<NC>(sh), Nh = sd.Nh, lstride = N + 1;
  const Tile3 t = tiles[blockIdx.x];
  const int s = t.s, ky = t.a, kx0 = t.b;
  const int j = sd.meta[4 * s], kxw = sd.meta[4 * s + 2];
  const int kys = signed_f(ky, N);
  // shell j is the |k|^2 range (mlo, mhi]
  const int64_t mlo = shell_mmax(j - 1, sd.step), mhi = shell_mmax(j, sd.step);
  constexpr int NL = TileCfg<T>::LINES;  // kx columns per tile (the plan's kt)
  // thread (row n0, column c) loads rows n0, n0 + RS, ... of kx column c, U
  // loads in flight; |k|^2 <= 3 (N/2)^2 fits in 32 bits, so the shell range
  // is tested on kx^2 + ky^2 + kz^2 in int with the bounds clamped to it
  constexpr int RS = FFT_THREADS / NL, U = 8;
  const int c = (int)threadIdx.x % NL, n0 = (int)threadIdx.x / NL, kx = kx0 + c;
  const int r2 = kx * kx + kys * kys;
  const int lo = (int)min(mlo, (int64_t)1 << 30), hi = (int)min(mhi, (int64_t)1 << 30);
  const bool colin = kx < Nh;
  const cplx<T>* src = H + (int64_t)ky * Nh + kx;
  const int64_t rstride = (int64_t)N * Nh;
  for (int nb = n0; nb < N; nb += RS * U) {
    cplx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = nb + u * RS;
      const int kzs = signed_f(n, N);
      const int m = r2 + kzs * kzs;
      v[u] = mk(T(0), T(0));
      if (colin && n < N && m > lo && m <= hi) v[u] = src[(int64_t)n * rstride];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = nb + u * RS;
      if (n < N) lds[c * lstride + n] = v[u];
    }
  }
  cplx<T>* twl = lds + NL * lstride;
  stage_twiddles(twl, tw, N);
  __syncthreads();
  fft_tile<NC, -1, T, NL, NL * 1024>(lds, sh, NL, lstride, twl);
  cplx<T>* dst = T1 + sd.off[s] + (int64_t)ky * kxw + kx0 + c;
  const int64_t dstride = (int64_t)N * kxw;
  for (int n = n0; n < N; n += RS) dst[(int64_t)n * dstride] = lds[c * lstride + n];
}

// y pass: tile (s, z, kx0): NL kx columns at one z; rows (ky, kx-tile) that
// the z pass skipped (|ky| > kyl, the tile's c) are zero and not read.
template <typename T, int NC>
__global__ __launch_bounds__(FFT_THREADS) void k_shell_y(const cplx<T>* __restrict__ T1,
                                                         cplx<T>* __restrict__ T2,
                                                         const Tile3* __restrict__ tiles,
                                                         ShellDev sd, FftShape sh,
                                                         const cplx<T>* __restrict__ tw,
                                                         int64_t tbs) {
  T1 += (int64_t)blockIdx.y * tbs;
  T2 += (int64_t)blockIdx.y * tbs;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int N = n_of<NC>(sh), lstride = N + 1;
  const Tile3 t = tiles[blockIdx.x];
  const int s = t.s, z = t.a, kx0 = t.b;
  const int kxw = sd.meta[4 * s + 2];
  const int64_t base = sd.off[s] + (int64_t)z * N * kxw + kx0;
  constexpr int NL = TileCfg<T>::LINES;
  // thread (row n0, column c) loads rows n0, n0 + RS, ... of column c, U loads
  // in flight; the rows |ky| <= kyl (kx0^2 + ky^2 <= mm) hold z-pass output
  constexpr int RS = FFT_THREADS / NL, U = 8;
  const int c = (int)threadIdx.x % NL, n0 = (int)threadIdx.x / NL;
  const int kyl = t.c;
  const cplx<T>* src = T1 + base + c;
  for (int nb = n0; nb < N; nb += RS * U) {
    cplx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = nb + u * RS;
      v[u] = mk(T(0), T(0));
      if (n < N && (n <= kyl || n >= N - kyl)) v[u] = src[(int64_t)n * kxw];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = nb + u * RS;
      if (n < N) lds[c * lstride + n] = v[u];
    }
  }
  cplx<T>* twl = lds + NL * lstride;
  stage_twiddles(twl, tw, N);
  __syncthreads();
  fft_tile<NC, -1, T, NL, NL * 1024>(lds, sh, NL, lstride, twl);
  cplx<T>* dst = T2 + base + c;
  for (int n = n0; n < N; n += RS) dst[(int64_t)n * kxw] = lds[c * lstride + n];
}

// x pass (Hermitian -> real): tile (s, z, y0) covers rows y0..y0+2NP-1 as NP
// pairs (NP = shell_x_pairs<T>(NC)); each pair is one complex FFT of
// X1 + i X2 whose real/imag parts are the two real rows.  kx = 0 and kx = N/2
// keep only their real part, which makes the result exactly Re FFT of the
// Hermitian extension.  Thread (p, lane) works on pair p throughout, so the
// gather, zero fill and store loops need no index division.
template <typename T, int NC>
__global__ __launch_bounds__(FFT_THREADS) void k_shell_x(const cplx<T>* __restrict__ T2,
                                                         T* __restrict__ D,
                                                         const Tile3* __restrict__ tiles,
                                                         ShellDev sd, FftShape sh,
                                                         const cplx<T>* __restrict__ tw,
                                                         bool vec4) {
  constexpr int NP = shell_x_pairs<T>(NC);
  constexpr int TPL = FFT_THREADS / NP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int N = n_of<NC>(sh), lstride = N + PSX_XPAD, hN = N / 2;
  const Tile3 t = tiles[blockIdx.x];
  const int s = t.s, z = t.a, y0 = t.b;
  const int kxw = sd.meta[4 * s + 2];
  const int kxl = min(kxw, hN + 1);
  const int npair = min(NP, (N - y0) / 2);
  const int p = (int)threadIdx.x / TPL, lane = (int)threadIdx.x - p * TPL;
  cplx<T>* L = lds + p * lstride;
  // zero the modes no gather writes: [kxl, N - kxl]
  if (p < npair)
    for (int n = kxl + lane; n <= N - kxl; n += TPL) L[n] = mk(T(0), T(0));
  cplx<T>* twl = lds + NP * lstride;
  stage_twiddles(twl, tw, N);
  if (p < npair) {
    const cplx<T>* r1 = T2 + sd.off[s] + ((int64_t)z * N + y0 + 2 * p) * kxw;
    const cplx<T>* r2 = r1 + kxw;
    constexpr int U = 4;
    for (int k0 = lane; k0 < kxl; k0 += TPL * U) {
      cplx<T> a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kx = k0 + u * TPL;
        if (kx < kxl) { a[u] = r1[kx]; b[u] = r2[kx]; }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kx = k0 + u * TPL;
        if (kx >= kxl) continue;
        if (kx == 0 || kx == hN) {
          L[kx] = mk(a[u].re, b[u].re);
        } else {
          L[kx] = mk(a[u].re - b[u].im, a[u].im + b[u].re);
          L[N - kx] = mk(a[u].re + b[u].im, b[u].re - a[u].im);
        }
      }
    }
  }
  __syncthreads();
  fft_tile<NC, -1, T, NP, NP * 1024>(lds, sh, npair, lstride, twl);
  // rows y0 .. y0 + 2 npair - 1 are contiguous in D: one coalesced stream
  const int64_t G = (int64_t)N * N * N;
  T* dst = D + (int64_t)s * G + ((int64_t)z * N + y0) * N;
  if (vec4) {
    // N % 4 == 0 and D 16-byte aligned: four consecutive values of one row
    // per thread, 16/32-byte stores
    for (int i4 = threadIdx.x; i4 < 2 * npair * N / 4; i4 += FFT_THREADS) {
      const int idx = 4 * i4, r = idx / N, n = idx - r * N;
      const cplx<T>* src = lds + (r >> 1) * lstride + n;
      T o[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = (r & 1) ? src[e].im : src[e].re;
      store4(dst + idx, o);
    }
  } else {
    for (int idx = threadIdx.x; idx < 2 * npair * N; idx += FFT_THREADS) {
      const int r = idx / N, n = idx - r * N;
      const cplx<T> v = lds[(r >> 1) * lstride + n];
      dst[idx] = (r & 1) ? v.im : v.re;
    }
  }
}

template <typename T, int NC>
static int shell_fields_n(psx_ctx* ctx, const FftPlan* p, const ShellPlan* sp, const cplx<T>* H,
                          int step, cplx<T>* t1, cplx<T>* t2, T* dout, bool do_x = true,
                          int nbox = 1) {
  const int64_t hbs = (int64_t)p->shape.N * p->shape.N * (p->shape.N / 2 + 1);
  const int N = p->shape.N;
  ShellDev sd{sp->d_meta, sp->d_off, step, N / 2 + 1};
  const size_t ldsz = ((size_t)TileCfg<T>::LINES * (N + 1) + N) * sizeof(cplx<T>);
  const size_t ldsx = ((size_t)shell_x_pairs<T>(NC) * (N + PSX_XPAD) + N) * sizeof(cplx<T>);
  const cplx<T>* tw = plan_tw<T>(p);
  if (sp->nz_tiles > 0) {
    PSX_TRY(set_lds(k_shell_z<T, NC>, ldsz));
    launch_timed(ctx, "k_shell_z", k_shell_z<T, NC>, dim3(sp->nz_tiles, nbox), dim3(FFT_THREADS),
                 ldsz, H, t1, sp->d_ztiles, sd, p->shape, tw, hbs, sp->t_elems);
    PSX_CHECK_LAUNCH();
  }
  if (sp->ny_tiles > 0) {
    PSX_TRY(set_lds(k_shell_y<T, NC>, ldsz));
    launch_timed(ctx, "k_shell_y", k_shell_y<T, NC>, dim3(sp->ny_tiles, nbox), dim3(FFT_THREADS),
                 ldsz, (const cplx<T>*)t1, t2, sp->d_ytiles, sd, p->shape, tw, sp->t_elems);
    PSX_CHECK_LAUNCH();
  }
  if (do_x && sp->nx_tiles > 0) {
    PSX_TRY(set_lds(k_shell_x<T, NC>, ldsx));
    launch_timed(ctx, "k_shell_x", k_shell_x<T, NC>, dim3(sp->nx_tiles), dim3(FFT_THREADS), ldsx,
                 (const cplx<T>*)t2, dout, sp->d_xtiles, sd, p->shape, tw,
                 N % 4 == 0 && ((uintptr_t)dout % 16) == 0);
    PSX_CHECK_LAUNCH();
  }
  return PSX_OK;
}

#define PSX_CASE_N(n) \
  case n: return shell_fields_n<T, n>(ctx, p, sp, H, step, (cplx<T>*)t1, (cplx<T>*)t2, (T*)dout);

template <typename T>
static int shell_fields_t(psx_ctx* ctx, const cplx<T>* H, int N, int step, int ncut, int nmax,
                          void* dout) {
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  const ShellPlan* sp;
  // the x tiles follow the pair count of the kernel instantiation used for N
  int nc = 0;
#define PSX_IF_N(n) if (N == n) nc = n;
  PSX_FFT_SIZES(PSX_IF_N)
#undef PSX_IF_N
  PSX_TRY(get_shell_plan(ctx, N, step, ncut, nmax, TileCfg<T>::LINES, shell_x_pairs<T>(nc), &sp));
  void *t1 = nullptr, *t2 = nullptr;
  PSX_TRY(ctx_buffer(ctx, sizeof(T) == 4 ? "shell_t1_f32" : "shell_t1_f64",
                     (size_t)sp->t_elems * sizeof(cplx<T>), &t1));
  PSX_TRY(ctx_buffer(ctx, sizeof(T) == 4 ? "shell_t2_f32" : "shell_t2_f64",
                     (size_t)sp->t_elems * sizeof(cplx<T>), &t2));
  switch (N) {
    PSX_FFT_SIZES(PSX_CASE_N)
    default: return shell_fields_n<T, 0>(ctx, p, sp, H, step, (cplx<T>*)t1, (cplx<T>*)t2, (T*)dout);
  }
}
#undef PSX_CASE_N

#define PSX_CASE_N(n) \
  case n: return shell_fields_n<float, n>(ctx, p, sp, H, step, t1, t2, nullptr, false, nbox);

int launch_shell_zy(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                    const ShellPlan** sp_out, const void** t2_out, int nbox) {
  const FftPlan* p;
  PSX_TRY(get_fft_plan(ctx, N, &p));
  int nc = 0;
#define PSX_IF_N(n) if (N == n) nc = n;
  PSX_FFT_SIZES(PSX_IF_N)
#undef PSX_IF_N
  const ShellPlan* sp;
  PSX_TRY(get_shell_plan(ctx, N, step, ncut, nmax, TileCfg<float>::LINES, shell_x_pairs<float>(nc), &sp));
  void *v1 = nullptr, *v2 = nullptr;
  PSX_TRY(ctx_buffer(ctx, "shell_t1_f32", (size_t)nbox * sp->t_elems * sizeof(cplx<float>), &v1));
  PSX_TRY(ctx_buffer(ctx, "shell_t2_f32", (size_t)nbox * sp->t_elems * sizeof(cplx<float>), &v2));
  *sp_out = sp;
  *t2_out = v2;
  const cplx<float>* H = (const cplx<float>*)half;
  cplx<float>* t1 = (cplx<float>*)v1;
  cplx<float>* t2 = (cplx<float>*)v2;
  switch (N) {
    PSX_FFT_SIZES(PSX_CASE_N)
    default: return shell_fields_n<float, 0>(ctx, p, sp, H, step, t1, t2, nullptr, false, nbox);
  }
}
#undef PSX_CASE_N

int launch_shell_fields(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                        int prec, void* dout) {
  if (prec == PSX_PREC_F64)
    return shell_fields_t<double>(ctx, (const cplx<double>*)half, N, step, ncut, nmax, dout);
  return shell_fields_t<float>(ctx, (const cplx<float>*)half, N, step, ncut, nmax, dout);
}

}  // namespace psx
