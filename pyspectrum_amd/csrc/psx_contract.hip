// Triangle contraction B_ijl = sum_x d_i(x) d_j(x) d_l(x) on MFMA (gfx950).
//
// Recast as C[(i,j), l] = sum_x P[(i,j), x] D[l, x] with P = d_i * d_j formed
// on the fly: rows are the shell pairs (i >= j, i <= 2j) sorted by (j, i) in
// tiles of 16, columns the shells in tiles of 16, and only the 16x16 output
// blocks that hold a triangle (l in [max(i-j, j0), j]) are computed.  A
// workgroup streams a contiguous range of grid points through LDS in chunks;
// each wave owns a fixed set of pair tiles and keeps their accumulators in
// registers for the whole range.  Per-workgroup partial blocks are reduced
// over workgroups in float64 in a fixed order (deterministic).
//
//   PSX_PREC_F64    v_mfma_f64_16x16x4f64, float64 shells and accumulators
//   PSX_PREC_F32    v_mfma_f32_16x16x4f32, float32 shells and accumulators
//   PSX_PREC_BF16X3 v_mfma_f32_16x16x32_bf16 on split operands
//                   x = hi + lo (hi = bf16(x), lo = bf16(x - hi)):
//                   P D ~ Ph Dh + Ph Dl + Pl Dh, float32 accumulators
//
// LDS per chunk: the shell values of KC points, rows padded by 4 elements
// (16 B) so row starts stay 16-byte aligned and consecutive rows fall in
// different bank groups.  The split-BF16 mode also stores the B operand
// pre-split (hi and lo planes), so every wave reads its B fragments with one
// 16-byte LDS load per plane instead of re-splitting them.
#include "psx_ctx.h"

namespace psx {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8v __attribute__((ext_vector_type(8)));

constexpr int CT_WAVES = 8;
constexpr int CT_THREADS = CT_WAVES * 64;
constexpr int CT_MAXLT = 4;

template <int MODE> struct CtTraits;
template <> struct CtTraits<PSX_PREC_F64> {
  typedef double TD; typedef d4v Acc; typedef double TP;
  static constexpr int KSTEP = 4, KC = 128, MAXPT = 4;
};
template <> struct CtTraits<PSX_PREC_F32> {
  typedef float TD; typedef f4v Acc; typedef float TP;
  static constexpr int KSTEP = 4, KC = 256, MAXPT = 8;
};
template <> struct CtTraits<PSX_PREC_BF16X3> {
  typedef float TD; typedef f4v Acc; typedef float TP;
  static constexpr int KSTEP = 32, KC = 128, MAXPT = 8;
};

template <int MODE> __host__ __device__ constexpr int ct_ldw() { return CtTraits<MODE>::KC + 4; }
// bf16 plane row length (elements): KC + 8 keeps 16-byte alignment
template <int MODE> __host__ __device__ constexpr int ct_ldwh() { return CtTraits<MODE>::KC + 8; }

// ns shell rows plus one all-zero row (index ns) that padded pair rows and
// padded shell columns point at.
template <int MODE>
static size_t ct_lds_bytes(int ns) {
  size_t b = (size_t)(ns + 1) * ct_ldw<MODE>() * sizeof(typename CtTraits<MODE>::TD);
  if (MODE == PSX_PREC_BF16X3) b += (size_t)2 * (ns + 1) * ct_ldwh<MODE>() * sizeof(unsigned short);
  return b;
}

__device__ __forceinline__ unsigned short bf16_bits(__bf16 h) {
  return __builtin_bit_cast(unsigned short, h);
}

// x = hi + lo with hi = bf16(x) (round to nearest even), lo = bf16(x - hi)
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf8v& hi, bf8v& lo) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const __bf16 h = (__bf16)x[k];
    hi[k] = h;
    lo[k] = (__bf16)(x[k] - (float)h);
  }
}

// D: [ns][G] shell fields; partials: [gridDim.y][gridDim.x][nslot][256].
template <int MODE>
__global__ __launch_bounds__(CT_THREADS) void k_contract(
    const typename CtTraits<MODE>::TD* __restrict__ D, int64_t G, int ns, int nlt,
    const int* __restrict__ wave_pt, const int* __restrict__ wave_mask,
    const int* __restrict__ wave_slot, const int* __restrict__ pair_s, int64_t pts_per_block,
    int nslot, typename CtTraits<MODE>::TP* __restrict__ partials) {
  typedef CtTraits<MODE> Tr;
  typedef typename Tr::TD TD;
  typedef typename Tr::Acc Acc;
  constexpr int KC = Tr::KC, KS = Tr::KSTEP, MAXPT = Tr::MAXPT;
  constexpr int LDW = ct_ldw<MODE>();
  constexpr int LDWH = ct_ldwh<MODE>();
  constexpr bool BF = MODE == PSX_PREC_BF16X3;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TD* lds = reinterpret_cast<TD*>(smem_raw);
  unsigned short* ldh = reinterpret_cast<unsigned short*>(smem_raw + (size_t)(ns + 1) * LDW * sizeof(TD));
  unsigned short* ldl = ldh + (size_t)(ns + 1) * LDWH;
  for (int k = threadIdx.x; k < LDW; k += CT_THREADS) lds[ns * LDW + k] = TD(0);
  if constexpr (MODE == PSX_PREC_BF16X3) {
    for (int k = threadIdx.x; k < LDWH; k += CT_THREADS) {
      ldh[ns * LDWH + k] = 0;
      ldl[ns * LDWH + k] = 0;
    }
  }

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15;   // row (A) / column (B) within a 16-block
  const int kq = lane >> 4;    // k group
  const int gw = blockIdx.y * CT_WAVES + wave;
  constexpr int KL = BF ? 8 : 1;  // consecutive k per lane

  int pt[MAXPT], msk[MAXPT], oi[MAXPT], oj[MAXPT];
  int need = 0;  // l tiles this wave reads
#pragma unroll
  for (int q = 0; q < MAXPT; ++q) {
    pt[q] = wave_pt[gw * MAXPT + q];
    msk[q] = wave_mask[gw * MAXPT + q];
    need |= msk[q];
    const int si = pt[q] >= 0 ? pair_s[(pt[q] * 16 + r16) * 2] : -1;
    const int sj = pt[q] >= 0 ? pair_s[(pt[q] * 16 + r16) * 2 + 1] : -1;
    // padded rows read the zero row
    oi[q] = (si >= 0 ? si : ns) * LDW + kq * KL;
    oj[q] = (sj >= 0 ? sj : ns) * LDW + kq * KL;
  }
  int ob[CT_MAXLT];  // B rows (shells), padded columns read the zero row
#pragma unroll
  for (int t = 0; t < CT_MAXLT; ++t) {
    const int s = t * 16 + r16;
    ob[t] = (s < ns) ? s : ns;
  }
  Acc acc[MAXPT][CT_MAXLT];
#pragma unroll
  for (int q = 0; q < MAXPT; ++q)
#pragma unroll
    for (int t = 0; t < CT_MAXLT; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[q][t][e] = 0;

  const int64_t x_begin = (int64_t)blockIdx.x * pts_per_block;
  const int64_t x_end = min(G, x_begin + pts_per_block);
  // Chunks are prefetched into registers one chunk ahead, so the HBM loads of
  // chunk c+1 are in flight while chunk c is contracted.  Each thread owns up
  // to NV groups of 4 consecutive values (ns <= 64 shells).
  constexpr int NG = KC / 4;                 // 4-value groups per shell row
  constexpr int NV = 64 * NG / CT_THREADS;   // groups per thread at ns = 64
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
#pragma unroll
            for (int e = 0; e < 4; ++e) pre[u][e] = src[e];
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
#pragma unroll
        for (int e = 0; e < 4; ++e) lds[s * LDW + k + e] = pre[u][e];
        if constexpr (BF) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const __bf16 h = (__bf16)(float)pre[u][e];
            ldh[s * LDWH + k + e] = bf16_bits(h);
            ldl[s * LDWH + k + e] = bf16_bits((__bf16)((float)pre[u][e] - (float)h));
          }
        }
      }
    }
  };
  if (x_begin < x_end) fetch(x_begin);
  for (int64_t x0 = x_begin; x0 < x_end; x0 += KC) {
    commit();
    __syncthreads();
    if (x0 + KC < x_end) fetch(x0 + KC);
    for (int k0 = 0; k0 < KC; k0 += KS) {
      if constexpr (BF) {
        bf8v bh[CT_MAXLT], bl[CT_MAXLT];
#pragma unroll
        for (int t = 0; t < CT_MAXLT; ++t) {
          if ((need >> t) & 1) {
            const us8v h = *reinterpret_cast<const us8v*>(ldh + ob[t] * LDWH + k0 + kq * 8);
            const us8v l = *reinterpret_cast<const us8v*>(ldl + ob[t] * LDWH + k0 + kq * 8);
            bh[t] = __builtin_bit_cast(bf8v, h);
            bl[t] = __builtin_bit_cast(bf8v, l);
          }
        }
#pragma unroll
        for (int q = 0; q < MAXPT; ++q) {
          if (pt[q] < 0) continue;
          const f4v* pa = reinterpret_cast<const f4v*>(lds + oi[q] + k0);
          const f4v* pb = reinterpret_cast<const f4v*>(lds + oj[q] + k0);
          const f4v a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
          float xa[8];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            xa[e] = a0[e] * b0[e];
            xa[4 + e] = a1[e] * b1[e];
          }
          bf8v ah, al;
          split_bf16(xa, ah, al);
#pragma unroll
          for (int t = 0; t < CT_MAXLT; ++t) {
            if (!((msk[q] >> t) & 1)) continue;
            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[q][t], 0, 0, 0);
            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[q][t], 0, 0, 0);
            acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[q][t], 0, 0, 0);
          }
        }
      } else {
        TD b[CT_MAXLT];
#pragma unroll
        for (int t = 0; t < CT_MAXLT; ++t)
          b[t] = ((need >> t) & 1) ? lds[ob[t] * LDW + k0 + kq] : TD(0);
#pragma unroll
        for (int q = 0; q < MAXPT; ++q) {
          if (pt[q] < 0) continue;
          const TD a = lds[oi[q] + k0] * lds[oj[q] + k0];
#pragma unroll
          for (int t = 0; t < CT_MAXLT; ++t) {
            if (!((msk[q] >> t) & 1)) continue;
            if constexpr (MODE == PSX_PREC_F64)
              acc[q][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[q][t], 0, 0, 0);
            else
              acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[t], acc[q][t], 0, 0, 0);
          }
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
    for (int t = 0; t < CT_MAXLT; ++t) {
      const int slot = wave_slot[(gw * MAXPT + q) * CT_MAXLT + t];
      if (slot < 0) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) out[(int64_t)slot * 256 + lane * 4 + e] = acc[q][t][e];
    }
  }
}

// Cross-workgroup reduction in float64, in a fixed order (deterministic):
// stage 1 sums every (group, slot, entry) over the workgroups with coalesced
// reads; stage 2 gathers each triangle's entry, tri_loc[t] = slot*256 +
// lane*4 + reg of its (row, col) in the MFMA output layout.
template <typename TP>
__global__ void k_contract_reduce(const TP* __restrict__ partials, int64_t nwg, int nslot,
                                  double* __restrict__ red) {
  const int slot = blockIdx.x, g = blockIdx.y, e = threadIdx.x;  // 256 threads
  const TP* p = partials + (int64_t)g * nwg * nslot * 256 + (int64_t)slot * 256 + e;
  double s = 0.0;
  for (int64_t w = 0; w < nwg; ++w) s += (double)p[w * nslot * 256];
  red[((int64_t)g * nslot + slot) * 256 + e] = s;
}

__global__ void k_contract_gather(const double* __restrict__ red, int nslot,
                                  const int* __restrict__ tri_loc, const int* __restrict__ tri_grp,
                                  int ntri, double* __restrict__ braw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  braw[t] = red[(int64_t)tri_grp[t] * nslot * 256 + tri_loc[t]];
}

template <int MODE>
static int contract_t(psx_ctx* ctx, const void* D, int N, int step, int ncut, int nmax,
                      double* braw) {
  typedef CtTraits<MODE> Tr;
  const ContractPlan* cp;
  PSX_TRY(get_contract_plan(ctx, MODE, step, ncut, nmax, &cp));
  PSX_REQUIRE(cp->maxpt == Tr::MAXPT, "contraction plan built for another precision");
  const int64_t G = (int64_t)N * N * N;
  const int ngroups = cp->waves / CT_WAVES;
  const int64_t nchunks = cdiv(G, Tr::KC);
  int64_t nwg = 1024;
  if (nwg > nchunks) nwg = nchunks;
  const int64_t pts_per_block = cdiv(nchunks, nwg) * Tr::KC;
  nwg = cdiv(G, pts_per_block);
  const int nslot = cp->nslot;
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "contract_partials",
                     (size_t)ngroups * nwg * nslot * 256 * sizeof(typename Tr::TP), &part));
  const size_t lds = ct_lds_bytes<MODE>(cp->ns);
  if (lds > 65536)
    PSX_HIP(hipFuncSetAttribute((const void*)k_contract<MODE>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_contract<MODE>, dim3((unsigned)nwg, ngroups), dim3(CT_THREADS), lds,
                     ctx->stream, (const typename Tr::TD*)D, G, cp->ns, cp->nlt, cp->d_wave_pt,
                     cp->d_wave_mask, cp->d_wave_slot, cp->d_pair_s, pts_per_block, nslot,
                     (typename Tr::TP*)part);
  PSX_CHECK_LAUNCH();
  void* red = nullptr;
  PSX_TRY(ctx_buffer(ctx, "contract_red", (size_t)ngroups * nslot * 256 * sizeof(double), &red));
  hipLaunchKernelGGL(k_contract_reduce<typename Tr::TP>, dim3((unsigned)nslot, ngroups), dim3(256),
                     0, ctx->stream, (const typename Tr::TP*)part, nwg, nslot, (double*)red);
  PSX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_contract_gather, dim3((unsigned)cdiv(cp->ntri, 128)), dim3(128), 0,
                     ctx->stream, (const double*)red, nslot, cp->d_tri_loc, cp->d_tri_grp,
                     cp->ntri, braw);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

int contract_maxpt(int prec) {
  if (prec == PSX_PREC_F64) return CtTraits<PSX_PREC_F64>::MAXPT;
  if (prec == PSX_PREC_F32) return CtTraits<PSX_PREC_F32>::MAXPT;
  return CtTraits<PSX_PREC_BF16X3>::MAXPT;
}

int launch_contract(psx_ctx* ctx, const void* D, int N, int step, int ncut, int nmax, int prec,
                    double* braw) {
  if (prec == PSX_PREC_F64) return contract_t<PSX_PREC_F64>(ctx, D, N, step, ncut, nmax, braw);
  if (prec == PSX_PREC_F32) return contract_t<PSX_PREC_F32>(ctx, D, N, step, ncut, nmax, braw);
  return contract_t<PSX_PREC_BF16X3>(ctx, D, N, step, ncut, nmax, braw);
}

// ----------------------------------------------------------- finalize -----
// _Bk_periodic (pyspectrum.py:418-445) and Bk_periodic (:346-355) in float64.
// tri: [ntri][3] shell indices (i, j, l); p0k[j-1]; cnt[t] = the normalisation
// sum_x I_i I_j I_l.  Zero-count triangles produce zeros before the unit
// conversion, as in the reference.
struct BkConst {
  double G, nbar, a3, b3, a6, b6;
};

__global__ void k_bk_finalize(const double* __restrict__ braw, const double* __restrict__ p0k,
                              const double* __restrict__ cnt, const int* __restrict__ tri, int ntri,
                              BkConst c, double* __restrict__ out /* [7][ntri] */) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
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

int launch_bk_finalize(psx_ctx* ctx, const double* braw, const double* p0k, const double* cnt,
                       const int* tri, int ntri, double G, double nbar, double a3, double b3,
                       double a6, double b6, double* out) {
  BkConst c{G, nbar, a3, b3, a6, b6};
  hipLaunchKernelGGL(k_bk_finalize, dim3((unsigned)cdiv(ntri, 128)), dim3(128), 0, ctx->stream,
                     braw, p0k, cnt, tri, ntri, c, out);
  PSX_CHECK_LAUNCH();
  return PSX_OK;
}

}  // namespace psx
