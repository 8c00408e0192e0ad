// LDS-tiled mixed-radix Stockham FFT engine for gfx950.
//
// A workgroup holds a tile of lines of length N (line-major in LDS, row
// stride N+1) and runs every radix stage in LDS: each stage gathers its
// butterflies into registers, synchronises, and scatters the autosorted
// result (Stockham, so there is no bit-reversal pass).  Radices 9, 8, 5, 4,
// 3, 2 cover every N = 2^a 3^b 5^c.
//
// Thread mapping: a line has TPL = 256/MAXLINES threads, thread (line, lane)
// owns butterflies j = lane + TPL*q of its line, so every LDS address is a
// per-thread base plus a compile-time offset; only k = j mod Ns (the twiddle
// row) is computed per butterfly.
//
// Stage order: odd radices first (their stride-R scatter in the first stage
// touches every LDS bank once per 16 lanes), then 4/2, then the 8s.
//
// Two instantiations of the same stage code:
//  * compile-time N (template parameter, the sizes the estimator uses): the
//    radix sequence, butterfly counts and all index divisions are constants,
//    and per-thread register arrays are sized exactly;
//  * runtime N (N = 0): any supported size, used for the others.
//
// Twiddles: one table of N entries per N, stage after stage, row k of a
// stage holding exp(-2 pi i r k / (Ns R)) for r = 1..R-1 (Ns = product of
// the earlier radices); computed in long double on the host and rounded
// once to the working precision.  The backward sign uses the conjugate.
#pragma once
#include "psx_common.h"

namespace psx {

constexpr int FFT_MAX_STAGES = 12;
constexpr int FFT_THREADS = 256;

struct FftShape {
  int N;
  int nstages;
  int radix[FFT_MAX_STAGES];
};

// Factor N = 2^a 3^b 5^c into stages: 9s, a 3, 5s, a 4 or 2, then 8s
// (order 2, the default: 9 x 5 x 8 at N = 360 measured fastest); order 0
// puts the 5s first, order 1 the 8s.
#ifndef PSX_FFT_ORDER
#define PSX_FFT_ORDER 2
#endif
__host__ __device__ constexpr FftShape fft_shape_of(int N) {
  FftShape sh{N, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int n = N, c2 = 0, c3 = 0, c5 = 0;
  while (n > 1 && n % 2 == 0) { n /= 2; ++c2; }
  while (n > 1 && n % 3 == 0) { n /= 3; ++c3; }
  while (n > 1 && n % 5 == 0) { n /= 5; ++c5; }
  if (n != 1) { sh.nstages = -1; return sh; }
  int r[FFT_MAX_STAGES + 8] = {};
  int ns = 0;
#if PSX_FFT_ORDER == 1
  for (int i = 0; i < c2 / 3; ++i) r[ns++] = 8;
  if (c2 % 3 == 1) r[ns++] = 2;
  if (c2 % 3 == 2) r[ns++] = 4;
  for (int i = 0; i < c3 / 2; ++i) r[ns++] = 9;
  if (c3 % 2) r[ns++] = 3;
  for (int i = 0; i < c5; ++i) r[ns++] = 5;
#elif PSX_FFT_ORDER == 2
  for (int i = 0; i < c3 / 2; ++i) r[ns++] = 9;
  if (c3 % 2) r[ns++] = 3;
  for (int i = 0; i < c5; ++i) r[ns++] = 5;
  if (c2 % 3 == 1) r[ns++] = 2;
  if (c2 % 3 == 2) r[ns++] = 4;
  for (int i = 0; i < c2 / 3; ++i) r[ns++] = 8;
#else
  for (int i = 0; i < c5; ++i) r[ns++] = 5;
  for (int i = 0; i < c3 / 2; ++i) r[ns++] = 9;
  if (c3 % 2) r[ns++] = 3;
  if (c2 % 3 == 1) r[ns++] = 2;
  if (c2 % 3 == 2) r[ns++] = 4;
  for (int i = 0; i < c2 / 3; ++i) r[ns++] = 8;
#endif
  if (ns > FFT_MAX_STAGES) { sh.nstages = -1; return sh; }
  for (int i = 0; i < ns; ++i) sh.radix[i] = r[i];
  sh.nstages = ns;
  return sh;
}

inline bool fft_factor(int N, FftShape* sh) {
  *sh = fft_shape_of(N);
  return N >= 2 && sh->nstages > 0;
}

// Lines per tile: up to 16 (FP32) / 8 (FP64) lines, at most ELEMS complex values.
template <typename T> struct TileCfg;
#ifndef PSX_TILE_LINES_F32
#define PSX_TILE_LINES_F32 16
#endif
template <> struct TileCfg<float> { static constexpr int ELEMS = 8192; static constexpr int LINES = PSX_TILE_LINES_F32; };
template <> struct TileCfg<double> { static constexpr int ELEMS = 4096; static constexpr int LINES = 8; };

__host__ __device__ constexpr int tile_lines_for(int N, int elems, int lines) {
  return elems / N < lines ? elems / N : lines;
}
template <typename T, int N> struct Tile {
  // compile-time N: exact; runtime N: the bound
  static constexpr int MAXLINES = N > 0 ? tile_lines_for(N, TileCfg<T>::ELEMS, TileCfg<T>::LINES)
                                        : TileCfg<T>::LINES;
};

template <typename T>
inline int tile_lines_host(int N) {
  return tile_lines_for(N, TileCfg<T>::ELEMS, TileCfg<T>::LINES);
}

template <int NC> __device__ __forceinline__ int n_of(const FftShape& sh) {
  if constexpr (NC > 0) return NC; else return sh.N;
}

// Fill LDS from global memory with U independent loads in flight per thread
// (a plain load/store loop waits for every load before its LDS store):
// load(idx) returns element idx < n, store(idx, v) puts it into LDS.
template <int U, typename V, typename Load, typename Store>
__device__ __forceinline__ void batched_fill(int n, Load load, Store store) {
  for (int i0 = threadIdx.x; i0 < n; i0 += FFT_THREADS * U) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = i0 + u * FFT_THREADS;
      if (idx < n) v[u] = load(idx);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = i0 + u * FFT_THREADS;
      if (idx < n) store(idx, v[u]);
    }
  }
}

// Copy the twiddle table into LDS (read by every radix stage).
template <typename T>
__device__ __forceinline__ void stage_twiddles(cplx<T>* twl, const cplx<T>* __restrict__ tw, int N) {
  for (int i = threadIdx.x; i < N; i += FFT_THREADS) twl[i] = tw[i];
}

// ------------------------------------------------------------ small DFTs --
// In-place DFT of length R on v[0..R), sign S of the exponent.
template <int S, typename T> __device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = a + b; v[1] = a - b;
}

template <int S, typename T> __device__ __forceinline__ void dft4(cplx<T>* v) {
  cplx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
  cplx<T> b0 = v[1] + v[3], b1 = mul_i<S>(v[1] - v[3]);
  v[0] = a0 + b0; v[2] = a0 - b0;
  v[1] = a1 + b1; v[3] = a1 - b1;
}

template <int S, typename T> __device__ __forceinline__ void dft3(cplx<T>* v) {
  const T c1 = T(-0.5);
  const T s1 = T(S) * T(0.86602540378443864676);  // sin(2 pi / 3)
  cplx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
  cplx<T> m = v[0] + t1 * c1;
  cplx<T> r = mk(-t2.im * s1, t2.re * s1);  // i * s1 * t2
  v[0] = v[0] + t1;
  v[1] = m + r;
  v[2] = m - r;
}

template <int S, typename T> __device__ __forceinline__ void dft5(cplx<T>* v) {
  const T c1 = T(0.30901699437494742410);   // cos(2 pi/5)
  const T c2 = T(-0.80901699437494742410);  // cos(4 pi/5)
  const T s1 = T(S) * T(0.95105651629515357212);  // sin(2 pi/5)
  const T s2 = T(S) * T(0.58778525229247312917);  // sin(4 pi/5)
  cplx<T> a1 = v[1] + v[4], b1 = v[1] - v[4];
  cplx<T> a2 = v[2] + v[3], b2 = v[2] - v[3];
  cplx<T> m1 = v[0] + a1 * c1 + a2 * c2;
  cplx<T> m2 = v[0] + a1 * c2 + a2 * c1;
  // i*(s1 b1 + s2 b2), i*(s2 b1 - s1 b2)
  cplx<T> n1 = b1 * s1 + b2 * s2;
  cplx<T> n2 = b1 * s2 - b2 * s1;
  cplx<T> r1 = mk(-n1.im, n1.re), r2 = mk(-n2.im, n2.re);
  v[0] = v[0] + a1 + a2;
  v[1] = m1 + r1; v[4] = m1 - r1;
  v[2] = m2 + r2; v[3] = m2 - r2;
}

template <int S, typename T> __device__ __forceinline__ void dft8(cplx<T>* v) {
  const T h = T(0.70710678118654752440);
  // radix-2 step between halves, twiddles w8^k, then two DFT4
  cplx<T> a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { a[k] = v[k] + v[k + 4]; b[k] = v[k] - v[k + 4]; }
  b[1] = mk((b[1].re - T(S) * b[1].im) * h, (b[1].im + T(S) * b[1].re) * h);
  b[2] = mul_i<S>(b[2]);
  b[3] = mk((-b[3].re - T(S) * b[3].im) * h, (-b[3].im + T(S) * b[3].re) * h);
  dft4<S>(a);
  dft4<S>(b);
#pragma unroll
  for (int k = 0; k < 4; ++k) { v[2 * k] = a[k]; v[2 * k + 1] = b[k]; }
}

// Length 9 as 3 x 3: n = 3 n1 + n2, k = k1 + 3 k2; DFT3 over n1, twiddle
// W9^(n2 k1), DFT3 over n2.
template <int S, typename T> __device__ __forceinline__ void dft9(cplx<T>* v) {
  const T c1 = T(0.76604444311897803520), s1 = T(S) * T(0.64278760968653932632);   // 2 pi/9
  const T c2 = T(0.17364817766693034885), s2 = T(S) * T(0.98480775301220805936);   // 4 pi/9
  const T c4 = T(-0.93969262078590838405), s4 = T(S) * T(0.34202014332566873304);  // 8 pi/9
  cplx<T> a[3][3];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    cplx<T> t[3] = {v[n2], v[3 + n2], v[6 + n2]};
    dft3<S>(t);
    a[n2][0] = t[0]; a[n2][1] = t[1]; a[n2][2] = t[2];
  }
  a[1][1] = a[1][1] * mk(c1, s1);
  a[1][2] = a[1][2] * mk(c2, s2);
  a[2][1] = a[2][1] * mk(c2, s2);
  a[2][2] = a[2][2] * mk(c4, s4);
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cplx<T> t[3] = {a[0][k1], a[1][k1], a[2][k1]};
    dft3<S>(t);
    v[k1] = t[0]; v[k1 + 3] = t[1]; v[k1 + 6] = t[2];
  }
}

template <int R, int S, typename T> __device__ __forceinline__ void dftR(cplx<T>* v) {
  if constexpr (R == 2) dft2<S>(v);
  else if constexpr (R == 3) dft3<S>(v);
  else if constexpr (R == 4) dft4<S>(v);
  else if constexpr (R == 5) dft5<S>(v);
  else if constexpr (R == 9) dft9<S>(v);
  else dft8<S>(v);
}

// ----------------------------------------------------------- one stage ---
template <int V> struct IntC {
  static constexpr int value = V;
  __device__ constexpr operator int() const { return V; }
};
template <typename I> struct IsOne { static constexpr bool value = false; };
template <> struct IsOne<IntC<1>> { static constexpr bool value = true; };

// lds: nlines lines, line l at lds + l*lstride.  Stage with radix R after the
// first Ns points have been combined (Ns = product of earlier radices); tw is
// this stage's twiddle rows [Ns][R-1].  Thread (line, lane), TPL lanes per
// line, owns butterflies j = lane + TPL*q, q < MAXQ.
template <int R, int S, typename T, int MAXQ, typename IntTPL, typename IntN, typename IntNs>
__device__ __forceinline__ void fft_stage(cplx<T>* lds, IntN N, IntNs Ns, IntTPL TPL, int nlines,
                                          int lstride, const cplx<T>* __restrict__ tw) {
  const int nb = N / R;
  const int line = (int)threadIdx.x / TPL;
  const int lane = (int)threadIdx.x - line * TPL;
  const bool act = line < nlines;
  cplx<T>* base = lds + line * lstride;
  cplx<T> v[MAXQ][R];
  int kk[MAXQ];
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int j = lane + q * TPL;
    kk[q] = 0;
    if (act && j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = base[j + r * nb];
      if constexpr (!IsOne<IntNs>::value) {
        const int k = j % Ns;
        kk[q] = k;
        const cplx<T>* w = tw + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> wr = w[r - 1];
          if (S > 0) wr.im = -wr.im;
          v[q][r] = v[q][r] * wr;
        }
      }
      dftR<R, S>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int j = lane + q * TPL;
    if (act && j < nb) {
      const int k = kk[q];
      cplx<T>* dst = base + (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// Compile-time N: stages unrolled by recursion over the radix sequence; TW
// is the offset of the stage's twiddle rows in the table.
template <int N, int STAGE, int NS, int TW, int S, typename T, int MAXLINES>
__device__ __forceinline__ void fft_stages_ct(cplx<T>* lds, int nlines, int lstride,
                                              const cplx<T>* __restrict__ tw) {
  constexpr FftShape sh = fft_shape_of(N);
  if constexpr (STAGE < sh.nstages) {
    constexpr int R = sh.radix[STAGE];
    constexpr int TPL = FFT_THREADS / MAXLINES;
    constexpr int MAXQ = (N / R + TPL - 1) / TPL;
    fft_stage<R, S, T, MAXQ>(lds, IntC<N>(), IntC<NS>(), IntC<TPL>(), nlines, lstride, tw + TW);
    fft_stages_ct<N, STAGE + 1, NS * R, TW + NS * (R - 1), S, T, MAXLINES>(lds, nlines, lstride, tw);
  }
}

// Butterflies per thread of the runtime-N engine: nb*nlines <= TILE_ELEMS/R
// and TPL = floor(256/nlines) >= (256 - nlines)/nlines, so nb/TPL <=
// (TILE_ELEMS/R)/(256 - MAXLINES).
template <int R, int TILE_ELEMS, int MAXLINES> constexpr int fft_maxq_rt() {
  return (TILE_ELEMS / R) / (FFT_THREADS - MAXLINES) + 1;
}

// Transform every line of the tile in place.  N > 0: compile-time length;
// N == 0: runtime length sh.N with at most MAXLINES*sh.N <= TILE_ELEMS.
template <int N, int S, typename T, int MAXLINES, int TILE_ELEMS>
__device__ __forceinline__ void fft_tile(cplx<T>* lds, const FftShape& sh, int nlines, int lstride,
                                         const cplx<T>* __restrict__ tw) {
  if constexpr (N > 0) {
    fft_stages_ct<N, 0, 1, 0, S, T, MAXLINES>(lds, nlines, lstride, tw);
  } else {
    const int TPL = FFT_THREADS / (nlines > 0 ? nlines : 1);
    int Ns = 1, off = 0;
    for (int s = 0; s < sh.nstages; ++s) {
      const int R = sh.radix[s];
      const cplx<T>* t = tw + off;
      switch (R) {
        case 9: fft_stage<9, S, T, fft_maxq_rt<9, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
        case 8: fft_stage<8, S, T, fft_maxq_rt<8, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
        case 5: fft_stage<5, S, T, fft_maxq_rt<5, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
        case 4: fft_stage<4, S, T, fft_maxq_rt<4, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
        case 3: fft_stage<3, S, T, fft_maxq_rt<3, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
        default: fft_stage<2, S, T, fft_maxq_rt<2, TILE_ELEMS, MAXLINES>()>(lds, sh.N, Ns, TPL, nlines, lstride, t); break;
      }
      off += Ns * (R - 1);
      Ns *= R;
    }
  }
}

// Host: the stage twiddle table of N entries (the last one unused).
template <typename C>
inline void fft_stage_twiddles(const FftShape& sh, const C* full /* exp(-2 pi i t/N), t < N */,
                               C* out, C one) {
  int Ns = 1, o = 0;
  for (int s = 0; s < sh.nstages; ++s) {
    const int R = sh.radix[s];
    const int stride = sh.N / (Ns * R);
    for (int k = 0; k < Ns; ++k)
      for (int r = 1; r < R; ++r) out[o++] = full[r * k * stride];
    Ns *= R;
  }
  while (o < sh.N) out[o++] = one;
}

// Sizes with a compile-time instantiation; others run the runtime-N engine.
#define PSX_FFT_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(360) X(384) X(512)

}  // namespace psx
