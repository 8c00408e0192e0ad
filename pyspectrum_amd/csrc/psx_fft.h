// LDS-tiled mixed-radix Stockham FFT engine for gfx950.
//
// A workgroup holds a tile of lines of length N (line-major in LDS, row
// stride N+1) and runs every radix stage in LDS: each stage gathers its
// butterflies into registers, synchronises, and scatters the autosorted
// result (Stockham, so there is no bit-reversal pass).  Radices 8, 5, 4, 3, 2
// cover every N = 2^a 3^b 5^c.
//
// Two instantiations of the same stage code:
//  * compile-time N (template parameter, the sizes the estimator uses): the
//    radix sequence, butterfly counts and all index divisions are constants,
//    and per-thread register arrays are sized exactly;
//  * runtime N (N = 0): any supported size, used for the others.
//
// Twiddles come from a per-N table tw[t] = exp(-2 pi i t / N) computed in
// long double on the host (rounded once to the working precision); the
// backward sign uses its conjugate.
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

// Factor N into the engine's radices (8 first, then 5, 4, 3, 2).
__host__ __device__ constexpr FftShape fft_shape_of(int N) {
  FftShape sh{N, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int n = N;
  const int order[5] = {8, 5, 4, 3, 2};
  for (int i = 0; i < 5; ++i) {
    const int r = order[i];
    while (n > 1 && n % r == 0 && sh.nstages < FFT_MAX_STAGES) {
      sh.radix[sh.nstages++] = r;
      n /= r;
    }
  }
  if (n != 1) sh.nstages = -1;
  return sh;
}

inline bool fft_factor(int N, FftShape* sh) {
  *sh = fft_shape_of(N);
  return N >= 2 && sh->nstages > 0;
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

template <int R, int S, typename T> __device__ __forceinline__ void dftR(cplx<T>* v) {
  if constexpr (R == 2) dft2<S>(v);
  else if constexpr (R == 3) dft3<S>(v);
  else if constexpr (R == 4) dft4<S>(v);
  else if constexpr (R == 5) dft5<S>(v);
  else dft8<S>(v);
}

// ----------------------------------------------------------- one stage ---
// lds: nlines lines, line l at lds + l*lstride.  Stage with radix R after the
// first Ns points have been combined (Ns = product of earlier radices).
// MAXQ = butterflies per thread for the largest tile.
template <int R, int S, typename T, int MAXQ, typename IntN, typename IntNs>
__device__ __forceinline__ void fft_stage(cplx<T>* lds, IntN N, IntNs Ns, int nlines, int lstride,
                                          const cplx<T>* __restrict__ tw) {
  const int nb = N / R;
  const int total = nb * nlines;
  const int tstride = N / (Ns * R);
  cplx<T> v[MAXQ][R];
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int b = threadIdx.x + q * FFT_THREADS;
    if (b < total) {
      const int line = b / nb;
      const int j = b - line * nb;
      const cplx<T>* src = lds + line * lstride;
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[j + r * nb];
      const int t = k * tstride;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = tw[r * t];
        if (S > 0) w.im = -w.im;
        v[q][r] = v[q][r] * w;
      }
      dftR<R, S>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int b = threadIdx.x + q * FFT_THREADS;
    if (b < total) {
      const int line = b / nb;
      const int j = b - line * nb;
      const int k = j % Ns;
      cplx<T>* dst = lds + line * lstride + (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

template <int V> struct IntC {
  static constexpr int value = V;
  __device__ constexpr operator int() const { return V; }
};

// Compile-time N: stages unrolled by recursion over the radix sequence.
template <int N, int STAGE, int NS, int S, typename T, int MAXLINES>
__device__ __forceinline__ void fft_stages_ct(cplx<T>* lds, int nlines, int lstride,
                                              const cplx<T>* __restrict__ tw) {
  constexpr FftShape sh = fft_shape_of(N);
  if constexpr (STAGE < sh.nstages) {
    constexpr int R = sh.radix[STAGE];
    constexpr int MAXQ = (MAXLINES * (N / R) + FFT_THREADS - 1) / FFT_THREADS;
    fft_stage<R, S, T, MAXQ>(lds, IntC<N>(), IntC<NS>(), nlines, lstride, tw);
    fft_stages_ct<N, STAGE + 1, NS * R, S, T, MAXLINES>(lds, nlines, lstride, tw);
  }
}

// Transform every line of the tile in place.  N > 0: compile-time length;
// N == 0: runtime length sh.N with at most MAXLINES*sh.N <= TILE_ELEMS.
template <int N, int S, typename T, int MAXLINES, int TILE_ELEMS>
__device__ __forceinline__ void fft_tile(cplx<T>* lds, const FftShape& sh, int nlines, int lstride,
                                         const cplx<T>* __restrict__ tw) {
  if constexpr (N > 0) {
    fft_stages_ct<N, 0, 1, S, T, MAXLINES>(lds, nlines, lstride, tw);
  } else {
    int Ns = 1;
    for (int s = 0; s < sh.nstages; ++s) {
      const int R = sh.radix[s];
      switch (R) {
        case 8: fft_stage<8, S, T, (TILE_ELEMS / 8 + FFT_THREADS - 1) / FFT_THREADS>(lds, sh.N, Ns, nlines, lstride, tw); break;
        case 5: fft_stage<5, S, T, (TILE_ELEMS / 5 + FFT_THREADS - 1) / FFT_THREADS>(lds, sh.N, Ns, nlines, lstride, tw); break;
        case 4: fft_stage<4, S, T, (TILE_ELEMS / 4 + FFT_THREADS - 1) / FFT_THREADS>(lds, sh.N, Ns, nlines, lstride, tw); break;
        case 3: fft_stage<3, S, T, (TILE_ELEMS / 3 + FFT_THREADS - 1) / FFT_THREADS>(lds, sh.N, Ns, nlines, lstride, tw); break;
        default: fft_stage<2, S, T, (TILE_ELEMS / 2 + FFT_THREADS - 1) / FFT_THREADS>(lds, sh.N, Ns, nlines, lstride, tw); break;
      }
      Ns *= R;
    }
  }
}

// Sizes with a compile-time instantiation; others run the runtime-N engine.
#define PSX_FFT_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(360) X(384) X(512)

}  // namespace psx
