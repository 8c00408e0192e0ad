// Shell-field x pass fused with the split-BF16 triangle contraction (gfx950).
//
// The shell fields d_j (pyspectrum.py:387-400) are formed one grid row at a
// time in LDS and contracted there (the einsum of pyspectrum.py:427-430), so
// the [shell][point] field array is never written to or read from HBM.  The z
// and y passes of the pruned inverse FFT (psx_fft.hip) leave T2: per shell,
// per (z, y) row, the kx < kxl spectrum of the row.  A workgroup owns a
// contiguous range of rows and, for each row:
//
//   1. builds one complex FFT line per shell pair (2p, 2p+1) from T2: the two
//      Hermitian rows packed as X_2p + i X_2p+1 (kx = 0 and N/2 keep their
//      real parts), only the kx < kxl entries and their mirrors are stored --
//      the first radix stage reads every other entry as zero;
//   2. runs the forward FFT of all lines in LDS (the Stockham engine of
//      psx_fft.h with 768 threads); the last stage writes the real parts to
//      shell row 2p and the imaginary parts to row 2p+1 (float32, the A
//      operand rows of the pair products) and both split into the bf16 hi/lo
//      planes of the B operand, in the MFMA K order;
//   3. contracts the row's points exactly as k_contract does (pair products
//      split on the fly, 16x16x32 bf16 MFMAs, float32 sums per row added into
//      float64 registers).
//
// The FFT lines and the float32 shell rows share one LDS region: line p
// occupies the floats of rows 2p and 2p+1, and the last stage reads all of
// its inputs before it writes.  The T2 values of the next row are loaded into
// registers while the current row is transformed and contracted.
//
// Row layout: a row of N points is contracted as KC = N rounded up to 32
// points; positions N..KC-1 are zero in every row and plane.  Row strides are
// 8 words modulo the 64 LDS banks (floats) and 8 words modulo 64 (bf16
// planes), the conflict-free layout of k_contract.
#include <algorithm>
#include <cstdio>
#include <map>

#include "psx_ctx.h"
#include "psx_mfma.h"

namespace psx {

// Waves per workgroup and pair tiles per wave: 12 x 2 (three waves per SIMD,
// at most 168 VGPRs each; the float32 and float64 accumulators of two tiles
// stay in registers through the FFT phase).  A SIMD overlaps little of one
// wave's VALU work with another's MFMAs, so its busy time is close to the sum
// of both; the third wave fills the LDS and barrier waits.
#ifndef PSX_XC_WAVES
#define PSX_XC_WAVES 12
#endif
#ifndef PSX_XC_MAXPT
#define PSX_XC_MAXPT 2
#endif
constexpr int XC_WAVES = PSX_XC_WAVES, XC_THREADS = XC_WAVES * 64, XC_MAXPT = PSX_XC_MAXPT;
constexpr int XC_NPF = (1536 + XC_THREADS - 1) / XC_THREADS;  // gathered (line, kx) items per thread and row
int xc_waves_per_group() { return XC_WAVES; }
int xc_maxpt() { return XC_MAXPT; }
#ifndef PSX_XC_FLUSH
#define PSX_XC_FLUSH 1     // rows per float32 accumulation chain (then float64)
#endif


template <int N> struct XcGeom {
  static constexpr int KC = (N + 31) / 32 * 32;
  static constexpr int L = KC + (72 - KC % 64) % 64;        // floats, 8 mod 64
  static constexpr int LDWH = KC + (144 - KC % 128) % 128;  // halves, 8 words mod 64
};

// LDS bytes of k_xcontract<N, ...> for ns shells.
template <int N>
static size_t xc_lds_bytes(int ns, int nlines_ct) {
  typedef XcGeom<N> Gm;
  size_t b = (size_t)2 * (ns / 2 + 1) * Gm::L * sizeof(float);        // lines / rows
  b += (size_t)2 * (ns + 1) * Gm::LDWH * sizeof(unsigned short);      // hi, lo planes
  b += (size_t)N * sizeof(cplx<float>);                               // twiddles
  b += (size_t)(nlines_ct + 3 * 64) * sizeof(int);                    // tables
  b += (size_t)64 * sizeof(int64_t);
  return b;
}

// One radix stage of the line FFTs (forward), out of place: the stage reads
// the lines at `in` and writes them at `out` (LAST: to emit(line, position,
// value)), so one barrier per stage suffices; the next stage swaps the two.
// Butterflies are numbered across all lines (b = line * N/R + j) and dealt to
// the threads in order, so the threads of a partly filled last pass sit in as
// few waves as possible and the other waves skip it.  FIRST: the input
// entries kxl <= n <= N - kxl of a line are zero (never stored); they are
// read from the zero entry zc (an address select instead of a branch).
template <int N, int NLINES, int STAGE, int NS, int TW, typename Emit>
__device__ __forceinline__ void xc_stages(const cplx<float>* in, cplx<float>* out, int nlines,
                                          const cplx<float>* __restrict__ tw,
                                          const int* __restrict__ kxl_line,
                                          const cplx<float>* zc, Emit& emit, int tid) {
  constexpr FftShape sh = fft_shape_of(N);
  if constexpr (STAGE < sh.nstages) {
    constexpr int R = sh.radix[STAGE];
    constexpr int NB = N / R;
    constexpr int MAXB = (NLINES * NB + XC_THREADS - 1) / XC_THREADS;
    constexpr bool FIRST = STAGE == 0, LAST = STAGE == sh.nstages - 1;
    constexpr int L = XcGeom<N>::L;
    const int nbf = nlines * NB;
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
      const int b = tid + q * XC_THREADS;
      if (b < nbf) {
        const int line = b / NB, j = b - line * NB;
        const cplx<float>* base = in + line * L;
        const int kxl = FIRST ? kxl_line[line] : 0;
        // the zero band kxl <= n <= N - kxl: n - kxl < zlen as unsigned
        const unsigned zlen = FIRST ? (unsigned)max(0, N - 2 * kxl + 1) : 0u;
        cplx<float> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int n = j + r * NB;
          if constexpr (FIRST)
            v[r] = *((unsigned)(n - kxl) < zlen ? zc : base + n);
          else
            v[r] = base[n];
        }
        int k = 0;
        if constexpr (NS > 1) {
          k = j % NS;
          const cplx<float>* w = tw + TW + k * (R - 1);
#pragma unroll
          for (int r = 1; r < R; ++r) v[r] = v[r] * w[r - 1];
        }
        dftR<R, -1>(v);
        cplx<float>* ob = out + line * L;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int pos = (j - k) * R + k + r * NS;
          if constexpr (LAST) emit(line, pos, v[r]);
          else ob[pos] = v[r];
        }
      }
    }
    __syncthreads();
    xc_stages<N, NLINES, STAGE + 1, NS * R, TW + NS * (R - 1)>(out, const_cast<cplx<float>*>(in),
                                                               nlines, tw, kxl_line, zc, emit, tid);
  }
}

// The MFMA blocks of one pair tile at one k step: for each column tile t in
// the mask m, acc_t += Ah Bh_t + Ah Bl_t + Al Bh_t (from zero when first),
// each accumulator updated in place under its own wave-uniform branch.
template <int NLT>
__device__ __forceinline__ void xc_mfma_blocks(f4v (&acc)[NLT], const bf8v& ah, const bf8v& al,
                                               const bf8v (&bh)[NLT], const bf8v (&bl)[NLT],
                                               bool first, int m) {
  const f4v zero = {0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < NLT; ++t)
    if ((m >> t) & 1) {
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], first ? zero : acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[t], 0, 0, 0);
    }
}

// T2: per shell s, rows (z*N + y) of kxw_s complex values at offset off[s];
// meta[4s + 1] = rmax_s, meta[4s + 2] = kxw_s.  items: (line, kx) pairs.
// partials: [box][gridDim.x][nslot][256] float64.
template <int N, int NLT, int NLINES>
__global__ __launch_bounds__(XC_THREADS) void k_xcontract(
    const cplx<float>* __restrict__ T2, const int64_t* __restrict__ off,
    const int* __restrict__ meta, int ns, const int* __restrict__ items, int nitems,
    const int* __restrict__ line_kxl, const cplx<float>* __restrict__ tw,
    const int* __restrict__ wave_pt, const int* __restrict__ wave_mask,
    const int* __restrict__ wave_slot, const int* __restrict__ pair_s, int rows_per_block,
    int nslot, double* __restrict__ partials, long long* __restrict__ phase_clocks,
    int64_t t2_bs) {
  // box blockIdx.y of a batch: its intermediate and its partials
  T2 += (int64_t)blockIdx.y * t2_bs;
  partials += (int64_t)blockIdx.y * gridDim.x * nslot * 256;
  typedef XcGeom<N> Gm;
  constexpr int KC = Gm::KC, L = Gm::L, LDWH = Gm::LDWH;
  constexpr int Nh = N / 2 + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nlines = (ns + 1) / 2;
  const int nrows = 2 * (ns / 2 + 1);
  float* rows = reinterpret_cast<float*>(smem_raw);
  cplx<float>* const lines_a = reinterpret_cast<cplx<float>*>(smem_raw);
  unsigned short* ldh = reinterpret_cast<unsigned short*>(rows + (size_t)nrows * L);
  unsigned short* ldl = ldh + (size_t)(ns + 1) * LDWH;
  cplx<float>* twl = reinterpret_cast<cplx<float>*>(ldl + (size_t)(ns + 1) * LDWH);
  int* kxl_l = reinterpret_cast<int*>(twl + N);
  int* s_kxw = kxl_l + NLINES;
  int* s_kxl = s_kxw + 64;
  int* s_pad = s_kxl + 64;  // [0], [1]: the zero complex entry of the first FFT stage
  int64_t* s_off = reinterpret_cast<int64_t*>(s_pad + 64);
  // The out-of-place x pass alternates between the lines over the float rows
  // (A) and lines over the bf16 planes (B), both free while a row is
  // transformed; the input goes where the last stage reads from B, whose
  // output (the float rows) lands in A.
  cplx<float>* const lines_b = reinterpret_cast<cplx<float>*>(ldh);
  constexpr bool IN_B = fft_shape_of(N).nstages % 2 == 1;
  cplx<float>* const lines = IN_B ? lines_b : lines_a;
  cplx<float>* const lines_o = IN_B ? lines_a : lines_b;

  // zero rows and planes once: the zero row (index ns) and the padding
  // columns N..KC-1 of the odd rows are never written again (the planes,
  // zero row and padding included, are rewritten from the rows every row)
  for (int i = threadIdx.x; i < nrows * L; i += XC_THREADS) rows[i] = 0.0f;
  for (int i = threadIdx.x; i < 2 * (ns + 1) * LDWH; i += XC_THREADS) ldh[i] = 0;
  for (int i = threadIdx.x; i < N; i += XC_THREADS) twl[i] = tw[i];
  for (int i = threadIdx.x; i < nlines; i += XC_THREADS) kxl_l[i] = line_kxl[i];
  if (threadIdx.x < 2) s_pad[threadIdx.x] = 0;  // the zero entry of the first FFT stage
  const cplx<float>* const zc = reinterpret_cast<const cplx<float>*>(s_pad);
  for (int s = threadIdx.x; s < ns; s += XC_THREADS) {
    s_kxw[s] = meta[4 * s + 2];
    s_kxl[s] = min(meta[4 * s + 1] + 1, Nh);
    s_off[s] = off[s];
  }

  const int wave = uniform((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, kq = lane >> 4;
  const int kbase = kq * 4, kbaseh = kq * 8;
  const unsigned neg_lo = in_register(0x0000BF80u), neg_hi = in_register(0xBF800000u);
  // column tiles of each pair tile, and of the wave
  int nq = 0, wmask = 0, msk[XC_MAXPT], oi[XC_MAXPT], oj[XC_MAXPT];
#pragma unroll
  for (int q = 0; q < XC_MAXPT; ++q) {
    const int p = uniform(wave_pt[wave * XC_MAXPT + q]);
    msk[q] = uniform(wave_mask[wave * XC_MAXPT + q]);
    wmask |= msk[q];
    nq += p >= 0 ? 1 : 0;
    const int si = p >= 0 ? pair_s[(p * 16 + r16) * 2] : -1;
    const int sj = p >= 0 ? pair_s[(p * 16 + r16) * 2 + 1] : -1;
    oi[q] = (si >= 0 ? si : ns) * L + kbase;
    oj[q] = (sj >= 0 ? sj : ns) * L + kbase;
  }
  int obh[NLT];
#pragma unroll
  for (int t = 0; t < NLT; ++t) {
    const int s = t * 16 + r16;
    obh[t] = ((s < ns) ? s : ns) * LDWH + kbaseh;
  }
  // MFMA accumulators restart every PSX_XC_FLUSH rows (at most 3 x 12 MFMAs
  // per row, so the low bias of the bf16 MFMA's internal rounding never
  // builds up over long chains) and are then added into float64 registers.
  typedef f4v Acc;
  const Acc zero = {0, 0, 0, 0};
  Acc acc[XC_MAXPT][NLT];
  d4v acc2[XC_MAXPT][NLT];
#pragma unroll
  for (int q = 0; q < XC_MAXPT; ++q)
#pragma unroll
    for (int t = 0; t < NLT; ++t) {
      acc[q][t] = zero;
#pragma unroll
      for (int e = 0; e < 4; ++e) acc2[q][t][e] = 0.0;
    }

  // this thread's (line, kx) items, the same for every row
  int it_p[XC_NPF], it_kx[XC_NPF];
#pragma unroll
  for (int u = 0; u < XC_NPF; ++u) {
    const int i = threadIdx.x + u * XC_THREADS;
    it_p[u] = i < nitems ? items[2 * i] : -1;
    it_kx[u] = i < nitems ? items[2 * i + 1] : 0;
  }
  __syncthreads();

  // per item, the T2 offsets of its two entries in row 0 and their row
  // strides (stride -1: the entry is outside its shell's kx range, zero)
  int64_t g_off[XC_NPF][2];
  int g_w[XC_NPF][2];
#pragma unroll
  for (int u = 0; u < XC_NPF; ++u) {
    const int p = it_p[u], kx = it_kx[u];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int sh = 2 * p + h, shc = max(0, min(sh, ns - 1));
      const bool in = p >= 0 && sh < ns && kx < s_kxl[shc];
      g_off[u][h] = in ? s_off[shc] + kx : 0;
      g_w[u][h] = in ? s_kxw[shc] : -1;
    }
  }
  cplx<float> pa[XC_NPF], pb[XC_NPF];
  auto fetch = [&](int row) {
#pragma unroll
    for (int u = 0; u < XC_NPF; ++u) {
      pa[u] = mk(0.0f, 0.0f);
      pb[u] = mk(0.0f, 0.0f);
      if (g_w[u][0] >= 0) pa[u] = T2[g_off[u][0] + (int64_t)row * g_w[u][0]];
      if (g_w[u][1] >= 0) pb[u] = T2[g_off[u][1] + (int64_t)row * g_w[u][1]];
    }
  };
  // the two Hermitian rows a (shell 2p) and b (shell 2p+1) as one line
  auto store_line = [&]() {
#pragma unroll
    for (int u = 0; u < XC_NPF; ++u) {
      const int p = it_p[u], kx = it_kx[u];
      if (p < 0) continue;
      cplx<float>* ln = lines + p * L;
      const cplx<float> a = pa[u], b = pb[u];
      if (kx == 0 || kx == N / 2) {
        ln[kx] = mk(a.re, b.re);
      } else {
        ln[kx] = mk(a.re - b.im, a.im + b.re);
        ln[N - kx] = mk(a.re + b.im, b.re - a.im);
      }
    }
  };
  // last FFT stage: real parts to row 2p, imaginary parts to row 2p+1
  struct Emit {
    float* rows;
    int ns;
    __device__ __forceinline__ void operator()(int line, int pos, cplx<float> v) {
      const int s0 = 2 * line, s1 = s0 + 1;
      rows[s0 * L + pos] = v.re;
      rows[s1 * L + pos] = s1 < ns ? v.im : 0.0f;
    }
  } emit{rows, ns};
  // The float rows 0 .. ns (row ns is zero) split into the bf16 hi/lo planes,
  // four consecutive points per thread; positions N..KC-1 of the even rows
  // lie inside the FFT lines and are zeroed here first.
  auto split_rows = [&](int tid) {
    if constexpr (KC > N) {
      for (int i = tid; i < nlines * (KC - N); i += XC_THREADS) {
        const int l = i / (KC - N);
        rows[2 * l * L + N + (i - l * (KC - N))] = 0.0f;
      }
      __syncthreads();
    }
    constexpr int Q = KC / 4;
    auto split4 = [&](const float* src, unsigned short* dh, unsigned short* dl) {
      const f4v x = *reinterpret_cast<const f4v*>(src);
      u2v h, l;
      const unsigned h0 = pack_bf16(x[0], x[1]), h1 = pack_bf16(x[2], x[3]);
      float r0, r1, r2, r3;
      bf16_residuals(h0, x[0], x[1], neg_lo, neg_hi, r0, r1);
      bf16_residuals(h1, x[2], x[3], neg_lo, neg_hi, r2, r3);
      h[0] = h0; h[1] = h1;
      l[0] = pack_bf16(r0, r1); l[1] = pack_bf16(r2, r3);
      *reinterpret_cast<u2v*>(dh) = h;
      *reinterpret_cast<u2v*>(dl) = l;
    };
    // point k = 32m + 16h + 4g is stored at 32m + 8g + 4h (the K order)
    auto korder = [](int k) { return (k & ~31) + ((k & 12) << 1) + ((k & 16) >> 2); };
    if constexpr (XC_THREADS % Q == 0) {
      // item i = tid + XC_THREADS * p is row s0 + RPP * p, quad c: the same
      // items as the flat loop below, with the divisions taken once
      constexpr int RPP = XC_THREADS / Q;
      const int s0 = tid / Q, k = (tid - s0 * Q) * 4, kp = korder(k);
      const float* src = rows + s0 * L + k;
      unsigned short* dh = ldh + s0 * LDWH + kp;
      unsigned short* dl = ldl + s0 * LDWH + kp;
      for (int s = s0; s < ns + 1; s += RPP) {
        split4(src, dh, dl);
        src += RPP * L; dh += RPP * LDWH; dl += RPP * LDWH;
      }
    } else {
      for (int i = tid; i < (ns + 1) * Q; i += XC_THREADS) {
        const int s = i / Q, k = (i - s * Q) * 4;
        split4(rows + s * L + k, ldh + s * LDWH + korder(k), ldl + s * LDWH + korder(k));
      }
    }
  };

  const int r_begin = (int)blockIdx.x * rows_per_block;
  const int r_end = min(N * N, r_begin + rows_per_block);
  if (r_begin < r_end) fetch(r_begin);
  int chain = 0;
#ifdef PSX_XC_PHASE_CLOCKS
  // diagnostic build: core-clock cycles per phase, per wave, summed over rows
  long long clk[6] = {0, 0, 0, 0, 0, 0}, c0 = clock64(), c1;
#define PSX_XC_TICK(i) (c1 = clock64(), clk[i] += c1 - c0, c0 = c1)
#else
#define PSX_XC_TICK(i) ((void)0)
#endif
  for (int row = r_begin; row < r_end; ++row) {
    store_line();
    __syncthreads();
    if (row + 1 < r_end) fetch(row + 1);
    PSX_XC_TICK(0);
    // the thread index passes through a register each row, so the compiler
    // recomputes the FFT and split indices per row instead of holding them
    // in registers through the contraction
    const int tid = (int)in_register(threadIdx.x);
    xc_stages<N, NLINES, 0, 1, 0>(lines, lines_o, nlines, twl, kxl_l, zc, emit, tid);
    PSX_XC_TICK(1);
    split_rows(tid);
    __syncthreads();
    PSX_XC_TICK(2);
    // the k steps of the row: the B fragments of the step, then per pair tile
    // (a slot with an empty mask is skipped) its products, split, and MFMAs
    if constexpr (PSX_XC_FLUSH > 1) {
      if (chain == 0) {
#pragma unroll
        for (int q = 0; q < XC_MAXPT; ++q)
#pragma unroll
          for (int t = 0; t < NLT; ++t) acc[q][t] = zero;
      }
    }
#pragma unroll
    for (int kb = 0; kb < KC / 32; ++kb) {
      const int k0 = kb * 32;
      bf8v bh[NLT], bl[NLT];
#pragma unroll
      for (int t = 0; t < NLT; ++t) {
        bh[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const us8v*>(ldh + obh[t] + k0));
        bl[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const us8v*>(ldl + obh[t] + k0));
      }
#pragma unroll
      for (int q = 0; q < XC_MAXPT; ++q) {
        if (msk[q] == 0) continue;
        const f4v* a4 = reinterpret_cast<const f4v*>(rows + oi[q] + k0);
        const f4v* b4 = reinterpret_cast<const f4v*>(rows + oj[q] + k0);
        bf8v ah, al;
        product_split_bf16(a4[0], a4[4], b4[0], b4[4], neg_lo, neg_hi, ah, al);
        xc_mfma_blocks<NLT>(acc[q], ah, al, bh, bl, PSX_XC_FLUSH == 1 && kb == 0, msk[q]);
      }
    }
    PSX_XC_TICK(3);
    if (++chain == PSX_XC_FLUSH || row + 1 == r_end) {
      chain = 0;
#pragma unroll
      for (int q = 0; q < XC_MAXPT; ++q) {
        if (q >= nq) break;
#pragma unroll
        for (int t = 0; t < NLT; ++t)
          if ((msk[q] >> t) & 1) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc2[q][t][e] += (double)acc[q][t][e];
          }
      }
    }
    PSX_XC_TICK(4);
    __syncthreads();
    PSX_XC_TICK(5);
  }
  double* out = partials + (int64_t)blockIdx.x * nslot * 256;
#pragma unroll
  for (int q = 0; q < XC_MAXPT; ++q) {
#pragma unroll
    for (int t = 0; t < NLT; ++t) {
      const int slot = uniform(wave_slot[(wave * XC_MAXPT + q) * 4 + t]);
      if (slot < 0) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) out[(int64_t)slot * 256 + lane * 4 + e] = acc2[q][t][e];
    }
  }
#ifdef PSX_XC_PHASE_CLOCKS
  if (lane == 0)
    for (int i = 0; i < 6; ++i) phase_clocks[((int64_t)blockIdx.x * XC_WAVES + wave) * 8 + i] = clk[i];
#else
  (void)phase_clocks;
#endif
#undef PSX_XC_TICK
}

// Compile-time sizes of the fused path and the line-count classes.
#ifndef PSX_XC_SIZES
#define PSX_XC_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(360) X(384)
#endif

static int xc_nlines_ct(int ns) {
  const int nl = (ns + 1) / 2;
  return nl <= 8 ? 8 : nl <= 16 ? 16 : nl <= 20 ? 20 : nl <= 24 ? 24 : -1;
}

static bool xc_size_ok(int N) {
  bool ok = false;
#define PSX_IS_N(n) ok = ok || N == n;
  PSX_XC_SIZES(PSX_IS_N)
#undef PSX_IS_N
  return ok;
}

static size_t xc_lds_for(int N, int ns, int nl) {
  size_t b = 0;
#define PSX_LDS_N(n) if (N == n) b = xc_lds_bytes<n>(ns, nl);
  PSX_XC_SIZES(PSX_LDS_N)
#undef PSX_LDS_N
  return b;
}

int xcontract_supported(psx_ctx* ctx, int N, int step, int ncut, int nmax, int prec) {
  if (prec != PSX_PREC_BF16X3 || !xc_size_ok(N)) return PSX_ERR_UNSUPPORTED;
  const int ns = nmax - ncut / step + 1;
  const int nl = xc_nlines_ct(ns);
  if (nl < 0 || ns > 48) return PSX_ERR_UNSUPPORTED;
  const ContractPlan* cp;
  PSX_TRY(get_contract_plan(ctx, PSX_PREC_BF16X3 | CT_LAYOUT_FUSED, step, ncut, nmax, &cp));
  if (cp->waves != XC_WAVES || cp->maxpt != XC_MAXPT || cp->nlt > 3) return PSX_ERR_UNSUPPORTED;
  // the line items of the shells (at most XC_NPF per thread)
  int items = 0;
  const int Nh = N / 2 + 1, hN = N / 2;
  const int64_t mgrid = 3 * (int64_t)hN * hN;
  for (int p = 0; p < (ns + 1) / 2; ++p) {
    int k = 0;
    for (int s = 2 * p; s < std::min(2 * p + 2, ns); ++s) {
      const int64_t mm = std::min(shell_mmax(ncut / step + s, step), mgrid);
      k = std::max(k, (int)std::min<int64_t>(isqrt64(mm), hN) + 1);
    }
    items += std::min(k, Nh);
  }
  if (items > XC_THREADS * XC_NPF) return PSX_ERR_UNSUPPORTED;
  if (xc_lds_for(N, ns, nl) > 160 * 1024) return PSX_ERR_UNSUPPORTED;
  return PSX_OK;
}

template <int N, int NLT, int NLINES>
static int xcontract_t(psx_ctx* ctx, const ContractPlan* cp, const ShellPlan* sp,
                       const cplx<float>* t2, const FftPlan* fp, double* braw, int nbox) {
  const int ns = cp->ns;
  PSX_REQUIRE(sp->nxlines <= NLINES && sp->nxitems <= XC_THREADS * XC_NPF,
              "fused contraction: shell plan outside the kernel's limits");
  auto kern = k_xcontract<N, NLT, NLINES>;
  const size_t lds = xc_lds_bytes<N>(ns, NLINES);
  PSX_TRY(set_lds(kern, lds));
  static thread_local std::map<std::tuple<int, const void*, size_t>, int> slots_cache;
  int& slots = slots_cache[std::make_tuple(ctx->device, (const void*)kern, lds)];
  if (slots == 0) {
    int per_cu = 0, cus = 0;
    PSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, XC_THREADS, lds));
    PSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    slots = std::max(1, per_cu) * cus;
  }
  const int nrows = N * N;
  const int rows_per_block = (int)cdiv(nrows, slots);
  const int nwg = (int)cdiv(nrows, rows_per_block);
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "contract_partials",
                     (size_t)nbox * nwg * cp->nslot * 256 * sizeof(double), &part));
  long long* clocks = nullptr;
#ifdef PSX_XC_PHASE_CLOCKS
  PSX_TRY(ctx_buffer(ctx, "xc_phase_clocks", (size_t)nwg * XC_WAVES * 8 * sizeof(long long),
                     (void**)&clocks));
#endif
  launch_timed(ctx, "k_xcontract", kern, dim3((unsigned)nwg, (unsigned)nbox), dim3(XC_THREADS), lds,
               t2, sp->d_off, sp->d_meta, ns, sp->d_xitems, sp->nxitems, sp->d_xline_kxl,
               plan_tw<float>(fp), cp->d_wave_pt, cp->d_wave_mask, cp->d_wave_slot, cp->d_pair_s,
               rows_per_block, cp->nslot, (double*)part, clocks, sp->t_elems);
  PSX_CHECK_LAUNCH();
#ifdef PSX_XC_PHASE_CLOCKS
  {
    // diagnostic build only: mean cycles per phase for each wave slot
    std::vector<long long> h((size_t)nwg * XC_WAVES * 8);
    PSX_HIP(hipStreamSynchronize(ctx->stream));
    PSX_HIP(hipMemcpy(h.data(), clocks, h.size() * sizeof(long long), hipMemcpyDeviceToHost));
    fprintf(stderr, "k_xcontract<%d> phase clocks per workgroup (gather fft split mfma flush wait):\n", N);
    for (int w = 0; w < XC_WAVES; ++w) {
      double m[6] = {0, 0, 0, 0, 0, 0};
      for (int b = 0; b < nwg; ++b)
        for (int i = 0; i < 6; ++i) m[i] += (double)h[((size_t)b * XC_WAVES + w) * 8 + i] / nwg;
      fprintf(stderr, "  wave %d: %.3g %.3g %.3g %.3g %.3g %.3g\n", w, m[0], m[1], m[2], m[3], m[4], m[5]);
    }
  }
#endif
  return launch_contract_reduce(ctx, cp, (const double*)part, nwg, 1, braw, nbox);
}

template <int N>
static int xcontract_n(psx_ctx* ctx, const ContractPlan* cp, const ShellPlan* sp,
                       const cplx<float>* t2, const FftPlan* fp, double* braw, int nbox) {
  const int nl = xc_nlines_ct(cp->ns);
  switch (cp->nlt) {
    case 1: return xcontract_t<N, 1, 8>(ctx, cp, sp, t2, fp, braw, nbox);
    case 2: return xcontract_t<N, 2, 16>(ctx, cp, sp, t2, fp, braw, nbox);
    case 3:
      return nl <= 20 ? xcontract_t<N, 3, 20>(ctx, cp, sp, t2, fp, braw, nbox)
                      : xcontract_t<N, 3, 24>(ctx, cp, sp, t2, fp, braw, nbox);
    default: return set_error(PSX_ERR_UNSUPPORTED, "fused contraction: at most 48 shells");
  }
}

int launch_xcontract(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                     double* braw, int nbox) {
  PSX_TRY(xcontract_supported(ctx, N, step, ncut, nmax, PSX_PREC_BF16X3));
  const ContractPlan* cp;
  PSX_TRY(get_contract_plan(ctx, PSX_PREC_BF16X3 | CT_LAYOUT_FUSED, step, ncut, nmax, &cp));
  const FftPlan* fp;
  PSX_TRY(get_fft_plan(ctx, N, &fp));
  const ShellPlan* sp;
  const void* t2;
  PSX_TRY(launch_shell_zy(ctx, half, N, step, ncut, nmax, &sp, &t2, nbox));
  switch (N) {
#define PSX_CASE_N(n) case n: return xcontract_n<n>(ctx, cp, sp, (const cplx<float>*)t2, fp, braw, nbox);
    PSX_XC_SIZES(PSX_CASE_N)
#undef PSX_CASE_N
    default: return set_error(PSX_ERR_UNSUPPORTED, "fused contraction: unsupported Ngrid");
  }
}

}  // namespace psx
