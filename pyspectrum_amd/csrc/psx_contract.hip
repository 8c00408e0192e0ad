// Triangle contraction B_ijl = sum_x d_i(x) d_j(x) d_l(x) on MFMA (gfx950).
//
// Recast as C[(i,j), l] = sum_x P[(i,j), x] D[l, x] with P = d_i * d_j formed
// on the fly: rows are the shell pairs (i >= j, i <= 2j) sorted by (j, i) in
// tiles of 16, columns the shells in tiles of 16, and only the 16x16 output
// blocks that hold a triangle (l in [max(i-j, j0), j]) are computed.  A
// workgroup streams a contiguous range of grid points through LDS in chunks;
// each wave owns up to MAXPT pair tiles and keeps their accumulators in
// registers for the whole range.  The launch has one workgroup per resident
// slot (occupancy x CUs), so the whole grid is one wave of workgroups.
// Per-workgroup partial blocks are reduced over workgroups in float64 in a
// fixed order (deterministic).
//
//   PSX_PREC_F64    v_mfma_f64_16x16x4f64, float64 shells and accumulators
//   PSX_PREC_F32    v_mfma_f32_16x16x4f32, float32 shells and accumulators
//   PSX_PREC_BF16X3 v_mfma_f32_16x16x32_bf16 on split operands
//                   x = hi + lo (hi = bf16(x), lo = bf16(x - hi)):
//                   P D ~ Ph Dh + Ph Dl + Pl Dh, float32 accumulators
//
// LDS per chunk: the shell values of KC points (one padded row per shell plus
// an all-zero row that padded pair rows and shell columns point at).  The
// split-BF16 mode also stores the B operand pre-split (hi and lo planes), so
// every wave reads its B fragments with one 16-byte LDS load per plane.
//
// K order.  The sum over grid points is order-free, so each mode maps the
// MFMA k index to points in the way that gives 16-byte LDS reads: BF16X3
// lane group kq holds points 4kq..4kq+3 and 16+4kq..16+4kq+3 of a 32-point
// step (two 16-byte reads of a float row; the hi/lo planes store each
// 32-point step in that order, so the B fragment is one 16-byte read);
// F32/F64 lane group kq owns the quarter [kq*KC/4, (kq+1)*KC/4) of the chunk
// and walks it one point per MFMA, reading 4 points (4 MFMAs) per vector load.
//
// LDS banks (measured on gfx950): a ds_read_b128 serves 8 lanes of two lane
// groups (16 lanes, 16 x 16 B = all 64 banks) per cycle, so it is
// conflict-free when, among those, the 16-byte units (row base + column)
// differ modulo 16.  Rows whose stride is 8 words modulo 64 banks and lane
// groups 4 words apart achieve that for 8 distinct rows (modulo 8); rows 4
// words apart with lane groups 8 words apart run at half
// the LDS bandwidth.
#include <algorithm>

#include "psx_ctx.h"
#include "psx_mfma.h"

namespace psx {

// Compile-time tuning: waves per workgroup (FP64/FP32, and split-BF16 at 4 l
// tiles), waves and pair tiles per wave of the split-BF16 mode up to 3 l
// tiles, points per chunk in the split-BF16 mode, and the minimum waves per
// SIMD that register allocation must allow (1 = no constraint).
#ifndef PSX_CT_WAVES
#define PSX_CT_WAVES 8
#endif
#ifndef PSX_CT_WAVES_BF
#define PSX_CT_WAVES_BF 12
#endif
#ifndef PSX_CT_WAVES_F32
#define PSX_CT_WAVES_F32 12
#endif
#ifndef PSX_CT_MAXPT_BF
#define PSX_CT_MAXPT_BF 2
#endif
#ifndef PSX_CT_MINW
#define PSX_CT_MINW 1
#endif
#ifndef PSX_CT_KC_BF
#define PSX_CT_KC_BF 256
#endif
#ifndef PSX_CT_KC_BF4
#define PSX_CT_KC_BF4 128
#endif
constexpr int CT_MAXLT = 4;   // l tiles in the plan layout (ns <= 64)

template <int MODE> struct CtTraits;
template <> struct CtTraits<PSX_PREC_F64> {
  typedef double TD; typedef d4v Acc; typedef double TP;
  static constexpr int KC = 128, MAXPT = 4, PAD = 2;
};
#ifndef PSX_CT_KC_F32
#define PSX_CT_KC_F32 128
#endif
template <> struct CtTraits<PSX_PREC_F32> {
  typedef float TD; typedef f4v Acc; typedef double TP;
  static constexpr int KC = PSX_CT_KC_F32, MAXPT = 2, PAD = 4;
};
#ifndef PSX_CT_PAD_BF
#define PSX_CT_PAD_BF 8     // float rows: 264 words = 8 modulo 64 banks
#endif
#ifndef PSX_CT_PADH_BF
#define PSX_CT_PADH_BF 16   // bf16 planes: 272 halves = 136 words = 8 modulo 64
#endif
template <> struct CtTraits<PSX_PREC_BF16X3> {
  typedef float TD; typedef f4v Acc; typedef double TP;
  static constexpr int KC = PSX_CT_KC_BF, MAXPT = 3, PAD = PSX_CT_PAD_BF;
};

// Waves per workgroup and pair tiles per wave.  Split-BF16 up to 3 l tiles:
// 12 waves x 2 tiles (168 VGPRs, so 3 waves per SIMD overlap each other's LDS,
// VALU and MFMA latencies; measured 4.5 % faster than 8 x 4 at Nmax 40).
// FP32: 12 waves x 2 tiles (at most 166 VGPRs): 24 pair tiles per group, so
// at 40 shells (21 tiles) one group reads D once (8 x 2 needed two groups,
// each reading all of D).  Otherwise 8 waves with the mode's tiles per wave
// (split-BF16 at 4 l tiles: 3, whose float32 and float64 accumulators fit
// 256 VGPRs at 2 waves per SIMD).
template <int MODE, int NLT> __host__ __device__ constexpr int ct_waves() {
  return (MODE == PSX_PREC_BF16X3 && NLT <= 3) ? PSX_CT_WAVES_BF
         : MODE == PSX_PREC_F32                ? PSX_CT_WAVES_F32
                                               : PSX_CT_WAVES;
}
template <int MODE, int NLT> __host__ __device__ constexpr int ct_maxpt() {
  return (MODE == PSX_PREC_BF16X3 && NLT <= 3) ? PSX_CT_MAXPT_BF : CtTraits<MODE>::MAXPT;
}

// Row strides (elements): padded so that consecutive rows start in
// different 16-byte bank groups and every row start stays 16-byte aligned.
// Points per chunk.  Split-BF16 at 4 l tiles: PSX_CT_KC_BF4 = 128.  Its 8
// waves x 3 tiles fill the 256 VGPRs of 2 waves per SIMD; with 256-point
// chunks the prefetch registers spilled to scratch (C4: 45.8 ms), with 128
// they fit (36.7 ms).
template <int MODE, int NLT> __host__ __device__ constexpr int ct_kc() {
  return (MODE == PSX_PREC_BF16X3 && NLT > 3) ? PSX_CT_KC_BF4 : CtTraits<MODE>::KC;
}
template <int MODE, int NLT> __host__ __device__ constexpr int ct_ldw() {
  return ct_kc<MODE, NLT>() + CtTraits<MODE>::PAD;
}
template <int MODE, int NLT> __host__ __device__ constexpr int ct_ldwh() {
  return ct_kc<MODE, NLT>() + PSX_CT_PADH_BF;
}

template <int MODE, int NLT>
static size_t ct_lds_bytes(int ns) {
  size_t b = (size_t)(ns + 1) * ct_ldw<MODE, NLT>() * sizeof(typename CtTraits<MODE>::TD);
  if (MODE == PSX_PREC_BF16X3) b += (size_t)2 * (ns + 1) * ct_ldwh<MODE, NLT>() * sizeof(unsigned short);
  return b;
}

// D: [ns][G] shell fields; partials: [gridDim.y][gridDim.x][nslot][256].
// NLT: l tiles read (ns <= 16*NLT).
template <int MODE, int NLT>
__global__ __launch_bounds__((ct_waves<MODE, NLT>() * 64), PSX_CT_MINW) void k_contract(
    const typename CtTraits<MODE>::TD* __restrict__ D, int64_t G, int ns,
    const int* __restrict__ wave_pt, const int* __restrict__ wave_mask,
    const int* __restrict__ wave_slot, const int* __restrict__ pair_s, int64_t pts_per_block,
    int nslot, typename CtTraits<MODE>::TP* __restrict__ partials) {
  typedef CtTraits<MODE> Tr;
  typedef typename Tr::TD TD;
  typedef typename Tr::Acc Acc;
  constexpr int KC = ct_kc<MODE, NLT>(), MAXPT = ct_maxpt<MODE, NLT>();
  constexpr int CT_WAVES = ct_waves<MODE, NLT>(), CT_THREADS = CT_WAVES * 64;
  constexpr int LDW = ct_ldw<MODE, NLT>();
  constexpr int LDWH = ct_ldwh<MODE, NLT>();
  constexpr bool BF = MODE == PSX_PREC_BF16X3;
  constexpr int KQ = KC / 4;  // F32/F64: points per lane group per chunk
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TD* lds = reinterpret_cast<TD*>(smem_raw);
  unsigned short* ldh = reinterpret_cast<unsigned short*>(smem_raw + (size_t)(ns + 1) * LDW * sizeof(TD));
  unsigned short* ldl = ldh + (size_t)(ns + 1) * LDWH;
  for (int k = threadIdx.x; k < LDW; k += CT_THREADS) lds[ns * LDW + k] = TD(0);
  if constexpr (BF) {
    for (int k = threadIdx.x; k < LDWH; k += CT_THREADS) {
      ldh[ns * LDWH + k] = 0;
      ldl[ns * LDWH + k] = 0;
    }
  }

  const int wave = uniform((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15;   // row (A) / column (B) within a 16-block
  const int kq = lane >> 4;    // k group
  const int gw = (int)blockIdx.y * CT_WAVES + wave;
  const int kbase = BF ? kq * 4 : kq * KQ;   // float rows
  const int kbaseh = kq * 8;                  // bf16 planes (split-BF16)
  // bf16 pairs (lo, hi) = (-1, 0) and (0, -1) for the split residuals
  const unsigned neg_lo = in_register(0x0000BF80u), neg_hi = in_register(0xBF800000u);

  // wave-uniform tile list (pair tiles fill q = 0 .. nq-1)
  int nq = 0, msk[MAXPT], oi[MAXPT], oj[MAXPT];
#pragma unroll
  for (int q = 0; q < MAXPT; ++q) {
    const int p = uniform(wave_pt[gw * MAXPT + q]);
    msk[q] = uniform(wave_mask[gw * MAXPT + q]);
    nq += p >= 0 ? 1 : 0;
    const int si = p >= 0 ? pair_s[(p * 16 + r16) * 2] : -1;
    const int sj = p >= 0 ? pair_s[(p * 16 + r16) * 2 + 1] : -1;
    // padded rows read the zero row
    oi[q] = (si >= 0 ? si : ns) * LDW + kbase;
    oj[q] = (sj >= 0 ? sj : ns) * LDW + kbase;
  }
  int ob[NLT], obh[NLT];  // B rows (shells); padded columns read the zero row
#pragma unroll
  for (int t = 0; t < NLT; ++t) {
    const int s = t * 16 + r16;
    ob[t] = ((s < ns) ? s : ns) * LDW + kbase;
    obh[t] = ((s < ns) ? s : ns) * LDWH + kbaseh;
  }
  // F32/BF16X3: the MFMA accumulators restart from zero every chunk and are
  // added into float64 accumulators (acc2) at the end of the chunk, so no
  // float32 accumulation chain is longer than one chunk.  (The bf16 MFMA's
  // internal sum of 32 products rounds low; over a long chain that bias is
  // systematic.)
  constexpr bool TWO = MODE != PSX_PREC_F64;
  constexpr int M2 = TWO ? MAXPT : 1, N2 = TWO ? NLT : 1;
  Acc acc[MAXPT][NLT];
  d4v acc2[M2][N2];
  const Acc zero = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < MAXPT; ++q)
#pragma unroll
    for (int t = 0; t < NLT; ++t) acc[q][t] = zero;
#pragma unroll
  for (int q = 0; q < M2; ++q)
#pragma unroll
    for (int t = 0; t < N2; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc2[q][t][e] = 0.0;

  const int64_t x_begin = (int64_t)blockIdx.x * pts_per_block;
  const int64_t x_end = min(G, x_begin + pts_per_block);
  // Chunks are prefetched into registers one chunk ahead, so the HBM loads of
  // chunk c+1 are in flight while chunk c is contracted.  Each thread owns NV
  // groups of 4 consecutive values (ns <= 16*NLT shells).
  constexpr int NG = KC / 4;                       // 4-value groups per shell row
  constexpr int NV = (NLT * 16 * NG + CT_THREADS - 1) / CT_THREADS;
  TD pre[NV][4];
  auto fetch = [&](int64_t xc) {
    const int kc = (int)min((int64_t)KC, x_end - xc);
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int idx = threadIdx.x + u * CT_THREADS;
      if (idx < ns * NG) {
        const int s = idx / NG, k = (idx - s * NG) * 4;
        const TD* src = D + (int64_t)s * G + xc + k;
        if (k + 4 <= kc) {
          if constexpr (sizeof(TD) == 4) {
            const f4v t4 = *reinterpret_cast<const f4v*>(src);
            pre[u][0] = t4[0]; pre[u][1] = t4[1]; pre[u][2] = t4[2]; pre[u][3] = t4[3];
          } else {
            const d2v t0 = *reinterpret_cast<const d2v*>(src);
            const d2v t1 = *reinterpret_cast<const d2v*>(src + 2);
            pre[u][0] = t0[0]; pre[u][1] = t0[1]; pre[u][2] = t1[0]; pre[u][3] = t1[1];
          }
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) pre[u][e] = (k + e < kc) ? src[e] : TD(0);
        }
      }
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      const int idx = threadIdx.x + u * CT_THREADS;
      if (idx < ns * NG) {
        const int s = idx / NG, k = (idx - s * NG) * 4;
        if constexpr (sizeof(TD) == 4) {
          f4v v;
          v[0] = pre[u][0]; v[1] = pre[u][1]; v[2] = pre[u][2]; v[3] = pre[u][3];
          *reinterpret_cast<f4v*>(lds + s * LDW + k) = v;
        } else {
          d2v v0, v1;
          v0[0] = pre[u][0]; v0[1] = pre[u][1]; v1[0] = pre[u][2]; v1[1] = pre[u][3];
          *reinterpret_cast<d2v*>(lds + s * LDW + k) = v0;
          *reinterpret_cast<d2v*>(lds + s * LDW + k + 2) = v1;
        }
        if constexpr (BF) {
          // x = hi + lo, hi = bf16(x), lo = bf16(x - hi) (x - hi is exact)
          u2v h, l;
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const float x0 = (float)pre[u][2 * p], x1 = (float)pre[u][2 * p + 1];
            const unsigned hb = pack_bf16(x0, x1);
            float r0, r1;
            bf16_residuals(hb, x0, x1, neg_lo, neg_hi, r0, r1);
            h[p] = hb;
            l[p] = pack_bf16(r0, r1);
          }
          // point k = 32m + 16h + 4g is stored at 32m + 8g + 4h (the K order)
          const int kp = (k & ~31) + ((k & 12) << 1) + ((k & 16) >> 2);
          *reinterpret_cast<u2v*>(ldh + s * LDWH + kp) = h;
          *reinterpret_cast<u2v*>(ldl + s * LDWH + kp) = l;
        }
      }
    }
  };
  if (x_begin < x_end) fetch(x_begin);
  for (int64_t x0 = x_begin; x0 < x_end; x0 += KC) {
    commit();
    __syncthreads();
    if (x0 + KC < x_end) fetch(x0 + KC);
    if constexpr (BF) {
#pragma unroll
      for (int k0 = 0; k0 < KC; k0 += 32) {
        bf8v bh[NLT], bl[NLT];
#pragma unroll
        for (int t = 0; t < NLT; ++t) {
          bh[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const us8v*>(ldh + obh[t] + k0));
          bl[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const us8v*>(ldl + obh[t] + k0));
        }
#pragma unroll
        for (int q = 0; q < MAXPT; ++q) {
          if (q >= nq) break;
          const f4v* pa = reinterpret_cast<const f4v*>(lds + oi[q] + k0);
          const f4v* pb = reinterpret_cast<const f4v*>(lds + oj[q] + k0);
          bf8v ah, al;
          product_split_bf16(pa[0], pa[4], pb[0], pb[4], neg_lo, neg_hi, ah, al);
#pragma unroll
          for (int t = 0; t < NLT; ++t) {
            if ((msk[q] >> t) & 1) {
              acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], k0 == 0 ? zero : acc[q][t], 0, 0, 0);
              acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[q][t], 0, 0, 0);
              acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[q][t], 0, 0, 0);
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int s0 = 0; s0 < KQ; s0 += 4) {
        TD b[NLT][4];
#pragma unroll
        for (int t = 0; t < NLT; ++t) {
          if constexpr (sizeof(TD) == 4) {
            const f4v v = *reinterpret_cast<const f4v*>(lds + ob[t] + s0);
            b[t][0] = v[0]; b[t][1] = v[1]; b[t][2] = v[2]; b[t][3] = v[3];
          } else {
            const d2v v0 = *reinterpret_cast<const d2v*>(lds + ob[t] + s0);
            const d2v v1 = *reinterpret_cast<const d2v*>(lds + ob[t] + s0 + 2);
            b[t][0] = v0[0]; b[t][1] = v0[1]; b[t][2] = v1[0]; b[t][3] = v1[1];
          }
        }
#pragma unroll
        for (int q = 0; q < MAXPT; ++q) {
          if (q >= nq) break;
          TD a[4];
          if constexpr (sizeof(TD) == 4) {
            const f4v ai = *reinterpret_cast<const f4v*>(lds + oi[q] + s0);
            const f4v aj = *reinterpret_cast<const f4v*>(lds + oj[q] + s0);
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = ai[e] * aj[e];
          } else {
            const d2v i0 = *reinterpret_cast<const d2v*>(lds + oi[q] + s0);
            const d2v i1 = *reinterpret_cast<const d2v*>(lds + oi[q] + s0 + 2);
            const d2v j0 = *reinterpret_cast<const d2v*>(lds + oj[q] + s0);
            const d2v j1 = *reinterpret_cast<const d2v*>(lds + oj[q] + s0 + 2);
            a[0] = i0[0] * j0[0]; a[1] = i0[1] * j0[1]; a[2] = i1[0] * j1[0]; a[3] = i1[1] * j1[1];
          }
#pragma unroll
          for (int t = 0; t < NLT; ++t) {
            if ((msk[q] >> t) & 1) {
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                if constexpr (MODE == PSX_PREC_F64)
                  acc[q][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[e], b[t][e], acc[q][t], 0, 0, 0);
                else
                  acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                      a[e], b[t][e], (s0 == 0 && e == 0) ? zero : acc[q][t], 0, 0, 0);
              }
            }
          }
        }
      }
    }
    if constexpr (TWO) {
#pragma unroll
      for (int q = 0; q < M2; ++q) {
        if (q >= nq) break;
#pragma unroll
        for (int t = 0; t < N2; ++t)
          if ((msk[q] >> t) & 1) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc2[q][t][e] += (double)acc[q][t][e];
          }
      }
    }
    __syncthreads();
  }
  // write the valid blocks of this workgroup
  typename Tr::TP* out = partials + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * nslot * 256;
#pragma unroll
  for (int q = 0; q < MAXPT; ++q) {
#pragma unroll
    for (int t = 0; t < NLT; ++t) {
      const int slot = uniform(wave_slot[(gw * MAXPT + q) * CT_MAXLT + t]);
      if (slot < 0) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if constexpr (TWO) out[(int64_t)slot * 256 + lane * 4 + e] = acc2[q][t][e];
        else out[(int64_t)slot * 256 + lane * 4 + e] = acc[q][t][e];
      }
    }
  }
}

// Cross-workgroup reduction in float64, in a fixed order (deterministic).
// Stage 1: block (slot, quarter, group) of 64 entries x 16 segments; thread
// (e, seg) sums its segment of the workgroups, then segment sums are added in
// order.  Stage 2 gathers each triangle's entry, tri_loc[t] = slot*256 +
// lane*4 + reg of its (row, col) in the MFMA output layout.
constexpr int RED_SEG = 16;

// Batches: blockIdx.z = box * ngroups + group, partials and red are per box
// back to back.
template <typename TP>
__global__ __launch_bounds__(64 * RED_SEG) void k_contract_reduce(const TP* __restrict__ partials,
                                                                  int64_t nwg, int nslot,
                                                                  double* __restrict__ red) {
  __shared__ double part[RED_SEG][64];
  const int slot = blockIdx.x, g = blockIdx.z;  // g runs over boxes x groups
  const int e = threadIdx.x & 63, seg = threadIdx.x >> 6;
  const int entry = (int)blockIdx.y * 64 + e;
  const int64_t per = (nwg + RED_SEG - 1) / RED_SEG;
  const int64_t w0 = seg * per;
  const int64_t w1 = min(nwg, w0 + per);
  const int64_t stride = (int64_t)nslot * 256;
  const TP* p = partials + ((int64_t)g * nwg * nslot + slot) * 256 + entry;
  double s = 0.0;
  for (int64_t w = w0; w < w1; ++w) s += (double)p[w * stride];
  part[seg][e] = s;
  __syncthreads();
  if (seg == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < RED_SEG; ++k) t += part[k][e];
    red[((int64_t)g * nslot + slot) * 256 + entry] = t;
  }
}

// Box blockIdx.y: its reduced blocks and its ntri outputs.
__global__ void k_contract_gather(const double* __restrict__ red, int nslot, int ngroups,
                                  const int* __restrict__ tri_loc, const int* __restrict__ tri_grp,
                                  int ntri, double* __restrict__ braw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  red += (int64_t)blockIdx.y * ngroups * nslot * 256;
  braw[(int64_t)blockIdx.y * ntri + t] = red[(int64_t)tri_grp[t] * nslot * 256 + tri_loc[t]];
}

int launch_contract_reduce(psx_ctx* ctx, const ContractPlan* cp, const double* part, int64_t nwg,
                           int ngroups, double* braw, int nbox) {
  const int nslot = cp->nslot;
  void* red = nullptr;
  PSX_TRY(ctx_buffer(ctx, "contract_red", (size_t)nbox * ngroups * nslot * 256 * sizeof(double),
                     &red));
  hipLaunchKernelGGL(k_contract_reduce<double>, dim3((unsigned)nslot, 4, nbox * ngroups),
                     dim3(64 * RED_SEG), 0, ctx->stream, part, nwg, nslot, (double*)red);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_contract_gather, dim3((unsigned)cdiv(cp->ntri, 128), nbox), dim3(128), 0,
                     ctx->stream, (const double*)red, nslot, ngroups, cp->d_tri_loc,
                     cp->d_tri_grp, cp->ntri, braw);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

// Workgroups resident at once on this device for the contraction kernel.
template <int MODE, int NLT>
static int contract_slots(psx_ctx* ctx, size_t lds, int* out) {
  static thread_local std::map<std::tuple<int, int, int, size_t>, int> cache;
  const auto key = std::make_tuple(ctx->device, MODE, NLT, lds);
  auto it = cache.find(key);
  if (it != cache.end()) { *out = it->second; return PSX_OK; }
  int per_cu = 0, cus = 0;
  PSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_contract<MODE, NLT>,
                                                       ct_waves<MODE, NLT>() * 64, lds));
  PSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if (per_cu < 1) per_cu = 1;
  *out = cache[key] = per_cu * cus;
  return PSX_OK;
}

template <int MODE, int NLT>
static int contract_t(psx_ctx* ctx, const ContractPlan* cp, const void* D, int N, double* braw) {
  typedef CtTraits<MODE> Tr;
  PSX_REQUIRE(cp->maxpt == (ct_maxpt<MODE, NLT>()), "contraction plan built for another precision");
  const int64_t G = (int64_t)N * N * N;
  constexpr int CT_WAVES = ct_waves<MODE, NLT>();
  PSX_REQUIRE(cp->waves % CT_WAVES == 0, "contraction plan built for another wave layout");
  const int ngroups = cp->waves / CT_WAVES;
  const size_t lds = ct_lds_bytes<MODE, NLT>(cp->ns);
  PSX_TRY(set_lds(k_contract<MODE, NLT>, lds));
  int slots = 0;
  PSX_TRY((contract_slots<MODE, NLT>(ctx, lds, &slots)));
  constexpr int KC = ct_kc<MODE, NLT>();
  const int64_t nchunks = cdiv(G, KC);
  int64_t nwg = std::max<int64_t>(1, slots / ngroups);
  if (nwg > nchunks) nwg = nchunks;
  const int64_t pts_per_block = cdiv(nchunks, nwg) * KC;
  nwg = cdiv(G, pts_per_block);
  const int nslot = cp->nslot;
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "contract_partials",
                     (size_t)ngroups * nwg * nslot * 256 * sizeof(typename Tr::TP), &part));
  launch_timed(ctx, "k_contract", k_contract<MODE, NLT>, dim3((unsigned)nwg, ngroups),
               dim3(CT_WAVES * 64), lds, (const typename Tr::TD*)D, G, cp->ns, cp->d_wave_pt,
               cp->d_wave_mask, cp->d_wave_slot, cp->d_pair_s, pts_per_block, nslot,
               (typename Tr::TP*)part);
  PSX_CHECK_LAUNCH();
  return launch_contract_reduce(ctx, cp, (const double*)part, nwg, ngroups, braw);
}

template <int MODE>
static int contract_mode(psx_ctx* ctx, const void* D, int N, int step, int ncut, int nmax,
                         double* braw) {
  const ContractPlan* cp;
  PSX_TRY(get_contract_plan(ctx, MODE, step, ncut, nmax, &cp));
  switch (cp->nlt) {
    case 1: return contract_t<MODE, 1>(ctx, cp, D, N, braw);
    case 2: return contract_t<MODE, 2>(ctx, cp, D, N, braw);
    case 3: return contract_t<MODE, 3>(ctx, cp, D, N, braw);
    case 4: return contract_t<MODE, 4>(ctx, cp, D, N, braw);
    default: return set_error(PSX_ERR_UNSUPPORTED, "at most 64 shells (Nmax - Ncut//step + 1 <= 64)");
  }
}

int contract_waves_per_group(int prec, int nlt) {
  if (prec == PSX_PREC_F64) return ct_waves<PSX_PREC_F64, 3>();
  if (prec == PSX_PREC_F32) return ct_waves<PSX_PREC_F32, 3>();
  return nlt > 3 ? ct_waves<PSX_PREC_BF16X3, 4>() : ct_waves<PSX_PREC_BF16X3, 3>();
}

int contract_maxpt(int prec, int nlt) {
  if (prec == PSX_PREC_F64) return CtTraits<PSX_PREC_F64>::MAXPT;
  if (prec == PSX_PREC_F32) return CtTraits<PSX_PREC_F32>::MAXPT;
  return nlt > 3 ? ct_maxpt<PSX_PREC_BF16X3, 4>() : ct_maxpt<PSX_PREC_BF16X3, 3>();
}

int launch_contract(psx_ctx* ctx, const void* D, int N, int step, int ncut, int nmax, int prec,
                    double* braw) {
  if (prec == PSX_PREC_F64) return contract_mode<PSX_PREC_F64>(ctx, D, N, step, ncut, nmax, braw);
  if (prec == PSX_PREC_F32) return contract_mode<PSX_PREC_F32>(ctx, D, N, step, ncut, nmax, braw);
  return contract_mode<PSX_PREC_BF16X3>(ctx, D, N, step, ncut, nmax, braw);
}

// ----------------------------------------------------------- finalize -----
// _Bk_periodic (pyspectrum.py:418-445) and Bk_periodic (:346-355) in float64.
// tri: [ntri][3] shell indices (i, j, l); p0k[j-1]; cnt[t] = the normalisation
// sum_x I_i I_j I_l.  Zero-count triangles produce zeros before the unit
// conversion, as in the reference.
// Box blockIdx.y of a batch: braw [box][ntri], p0k [box][p0k_stride],
// out [box][7][ntri], its own nbar and kf^3, kf^6.
struct BkConst {
  double G, a3, a6;
  double nbar[PSX_MAX_BATCH], b3[PSX_MAX_BATCH], b6[PSX_MAX_BATCH];
};

__global__ void k_bk_finalize(const double* __restrict__ braw, const double* __restrict__ p0k,
                              const double* __restrict__ cnt, const int* __restrict__ tri, int ntri,
                              int p0k_stride, BkConst bc, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  const int box = blockIdx.y;
  braw += (int64_t)box * ntri;
  p0k += (int64_t)box * p0k_stride;
  out += (int64_t)box * 7 * ntri;
  struct {
    double G, nbar, a3, b3, a6, b6;
  } c = {bc.G, bc.nbar[box], bc.a3, bc.b3[box], bc.a6, bc.b6[box]};
  const int i = tri[3 * t], j = tri[3 * t + 1], l = tri[3 * t + 2];
  double fac = 1.0;
  if (i == j && j == l) fac = 6.0;
  else if (i == j || j == l) fac = 2.0;
  const double n = cnt[t];
  double p1 = 0, p2 = 0, p3 = 0, b = 0, cc = 0;
  if (n > 0) {
    p1 = p0k[i - 1]; p2 = p0k[j - 1]; p3 = p0k[l - 1];
    b = braw[t] / n;
    cc = n / (fac * c.G);
  }
  // same operation order as the reference's numpy expressions
  const double inv = 1.0 / c.nbar;
  p1 = p1 * c.a3 / c.b3 - inv;
  p2 = p2 * c.a3 / c.b3 - inv;
  p3 = p3 * c.a3 / c.b3 - inv;
  const double sn = (p1 + p2 + p3) / c.nbar + 1.0 / (c.nbar * c.nbar);
  const double b123 = b * c.a6 / c.b6 - sn;
  out[0 * (int64_t)ntri + t] = p1;
  out[1 * (int64_t)ntri + t] = p2;
  out[2 * (int64_t)ntri + t] = p3;
  out[3 * (int64_t)ntri + t] = b123;
  out[4 * (int64_t)ntri + t] = b123 / (p1 * p2 + p1 * p3 + p2 * p3);
  out[5 * (int64_t)ntri + t] = cc;
  out[6 * (int64_t)ntri + t] = sn;
}

int launch_bk_finalize_batch(psx_ctx* ctx, int nbox, const double* braw, const double* p0k,
                             int p0k_stride, const double* cnt, const int* tri, int ntri, double G,
                             const double* nbar, double a3, const double* b3, double a6,
                             const double* b6, double* out) {
  PSX_REQUIRE(nbox >= 1 && nbox <= PSX_MAX_BATCH, "batch of 1 to 16 boxes");
  BkConst c{};
  c.G = G; c.a3 = a3; c.a6 = a6;
  for (int b = 0; b < nbox; ++b) { c.nbar[b] = nbar[b]; c.b3[b] = b3[b]; c.b6[b] = b6[b]; }
  hipLaunchKernelGGL(k_bk_finalize, dim3((unsigned)cdiv(ntri, 128), nbox), dim3(128), 0,
                     ctx->stream, braw, p0k, cnt, tri, ntri, p0k_stride, c, out);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int launch_bk_finalize(psx_ctx* ctx, const double* braw, const double* p0k, const double* cnt,
                       const int* tri, int ntri, double G, double nbar, double a3, double b3,
                       double a6, double b6, double* out) {
  return launch_bk_finalize_batch(ctx, 1, braw, p0k, 0, cnt, tri, ntri, G, &nbar, a3, &b3, a6, &b6,
                                  out);
}

}  // namespace psx
