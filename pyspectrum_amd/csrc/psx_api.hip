// C ABI, context, workspace and plan construction for libpsx.so.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

#include "psx_ctx.h"
#include "psx_plan.h"

namespace psx {

// stage launchers from the other translation units
int launch_fcomb(psx_ctx* ctx, const void* F, void* Hh, int N, double sumw, int prec,
                 const void* rec, const float* Wf, const double* Wd);
int launch_herm_half(psx_ctx* ctx, const void* F, int in_prec, int64_t sx, int64_t sy, int64_t sz,
                     void* Hh, int N, int prec);
int launch_reflect(psx_ctx* ctx, const void* Hh, void* F, int N, int prec);
int launch_bin_power(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                     int prec, double* out3, int nbox = 1);
int launch_bin_power_noalloc(psx_ctx* ctx, const void* half, int N, int step, int nbins, double kf,
                             int prec, double* out3);
int launch_pk_rsd(psx_ctx* ctx, const void* half, int N, int irsd, int lbox, int nbin, int nmu,
                  int prec, double* gout);
int launch_contract(psx_ctx* ctx, const void* D, int N, int step, int ncut, int nmax, int prec,
                    double* braw);
int contract_maxpt(int prec, int nlt);
int contract_waves_per_group(int prec, int nlt);
int launch_bk_finalize(psx_ctx* ctx, const double* braw, const double* p0k, const double* cnt,
                       const int* tri, int ntri, double G, double nbar, double a3, double b3,
                       double a6, double b6, double* out);
int launch_p0k(psx_ctx* ctx, const double* bins3, int nmax, int j0, double* p0k, int nbox = 1);
int launch_mock_gaussian(psx_ctx* ctx, int M, double lbox, uint64_t seed, double slope, double kcut,
                         float* out, double* sigma_host);
int launch_mock_sample(psx_ctx* ctx, const float* g, int M, double lbox, int64_t np, uint64_t seed,
                       double sigma_v, double fac, int los, float* xyz);
int launch_mock_velocities(psx_ctx* ctx, int64_t np, uint64_t seed, double sigma, float* v);
int launch_apply_rsd(psx_ctx* ctx, const float* xyz, const float* v, int64_t np, double fac,
                     double lbox, int los, float* out);

// Serialises the entry points on one context (host-side plan and workspace maps).
#define PSX_CTX_LOCK(c) std::lock_guard<std::recursive_mutex> psx_ctx_lock_((c)->mu)

std::string& last_error_string() {
  static thread_local std::string s;
  return s;
}

int set_error(int code, const std::string& msg) {
  last_error_string() = msg;
  return code;
}

static int prec_ok(int prec) {
  return prec == PSX_PREC_F64 || prec == PSX_PREC_F32 || prec == PSX_PREC_BF16X3;
}
static size_t cbytes(int prec) { return prec == PSX_PREC_F64 ? 16 : 8; }
static size_t rbytes(int prec) { return prec == PSX_PREC_F64 ? 8 : 4; }
// The FFT/shell precision of a pipeline precision.
static int fft_prec(int prec) { return prec == PSX_PREC_F64 ? PSX_PREC_F64 : PSX_PREC_F32; }

static int check_ngrid(int N) {
  FftShape sh;
  if (N < 4 || N > 1024 || (N & 1) || !fft_factor(N, &sh))
    return set_error(PSX_ERR_UNSUPPORTED,
                     "Ngrid must be even, in [4, 1024] and a product of 2, 3 and 5");
  return PSX_OK;
}

int ctx_buffer(psx_ctx* ctx, const char* name, size_t bytes, void** out) {
  DevBuf& b = ctx->bufs[name];
  if (b.bytes < bytes) {
    if (b.ptr) {
      PSX_HIP(hipStreamSynchronize(ctx->stream));
      PSX_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      return set_error(PSX_ERR_NOMEM, std::string("workspace '") + name + "' of " +
                                          std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    b.bytes = bytes;
  }
  *out = b.ptr;
  return PSX_OK;
}

int lds_attribute_needed(const void* kernel, size_t bytes) {
  static std::mutex m;
  static std::map<std::pair<int, const void*>, size_t> set;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  std::lock_guard<std::mutex> g(m);
  size_t& have = set[{dev, kernel}];
  if (have >= bytes) return 0;
  have = bytes;
  return 1;
}

int ctx_alloc_owned(psx_ctx* ctx, size_t bytes, void** out) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) return set_error(PSX_ERR_NOMEM, std::string("plan allocation: ") + hipGetErrorString(e));
  ctx->owned.push_back(p);
  *out = p;
  return PSX_OK;
}

template <typename V>
static int upload(psx_ctx* ctx, const std::vector<V>& h, V** d) {
  void* p;
  PSX_TRY(ctx_alloc_owned(ctx, h.size() * sizeof(V), &p));
  if (!h.empty()) PSX_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
  *d = (V*)p;
  return PSX_OK;
}

// ------------------------------------------------------------- FFT plan ----
// Twiddles exp(-2 pi i t/N) in long double, rounded once per precision.
// fcomb tables follow estimator.f:607-645: tpi is a float32 parameter,
// tpiL = tpi/float(N) a float32 quotient stored in double, recx =
// cmplx(dcos(piL), dsin(piL)) a single-precision complex, xrec = recx^n by a
// double recurrence; the window argument rk = tpiL*n is rounded to float32.
int get_fft_plan(psx_ctx* ctx, int N, const FftPlan** out) {
  auto it = ctx->fft_plans.find(N);
  if (it != ctx->fft_plans.end()) { *out = &it->second; return PSX_OK; }
  FftPlan p;
  if (!fft_factor(N, &p.shape)) return set_error(PSX_ERR_UNSUPPORTED, "unsupported FFT length");
  std::vector<c64> full(N), t64(N);
  std::vector<c32> t32(N);
  for (int t = 0; t < N; ++t) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)N;
    full[t] = mk((double)cosl(a), (double)sinl(a));
  }
  // stage-major rows of the FFT engine (psx_fft.h)
  fft_stage_twiddles(p.shape, full.data(), t64.data(), mk(1.0, 0.0));
  for (int t = 0; t < N; ++t) t32[t] = mk((float)t64[t].re, (float)t64[t].im);
  PSX_TRY(upload(ctx, t64, (c64**)&p.tw64));
  PSX_TRY(upload(ctx, t32, (c32**)&p.tw32));
  const int Nh = N / 2 + 1;
  const float tpi32 = 6.283185307f;
  const double tpiL = (double)(float)(tpi32 / (float)N);
  const double piL = -tpiL / 2.0;
  const c64 rec1 = mk((double)(float)cos(piL), (double)(float)sin(piL));
  std::vector<c64> rec(Nh);
  c64 x = mk(1.0, 0.0);
  for (int n = 0; n < Nh; ++n) { rec[n] = x; x = x * rec1; }
  std::vector<float> Wf(Nh);
  std::vector<double> Wd(Nh);
  for (int n = 0; n < Nh; ++n) {
    const float rk = (float)(tpiL * n);
    if (rk != 0.0f) {
      const float h = rk / 2.0f;
      const float q = sinf(h) / h;
      const float q2 = q * q;
      Wf[n] = q2 * q2;
      const double hd = (double)rk / 2.0;
      const double qd = sin(hd) / hd;
      Wd[n] = (qd * qd) * (qd * qd);
    } else {
      Wf[n] = 1.0f;
      Wd[n] = 1.0;
    }
  }
  PSX_TRY(upload(ctx, rec, (c64**)&p.rec));
  PSX_TRY(upload(ctx, Wf, &p.Wf));
  PSX_TRY(upload(ctx, Wd, &p.Wd));
  ctx->fft_plans[N] = p;
  *out = &ctx->fft_plans[N];
  return PSX_OK;
}

// ----------------------------------------------------------- shell plan ----

int get_shell_plan(psx_ctx* ctx, int N, int step, int ncut, int nmax, int kt, int xpairs,
                   const ShellPlan** out) {
  auto key = std::make_tuple(N, step, ncut, nmax, kt, xpairs);
  auto it = ctx->shell_plans.find(key);
  if (it != ctx->shell_plans.end()) { *out = it->second.get(); return PSX_OK; }
  auto sp = std::make_unique<ShellPlan>();
  sp->N = N; sp->step = step; sp->ncut = ncut; sp->nmax = nmax; sp->kt = kt;
  sp->j0 = ncut / step;
  sp->ns = nmax - sp->j0 + 1;
  const int Nh = N / 2 + 1, hN = N / 2;
  const int64_t mgrid = 3 * (int64_t)hN * hN;
  std::vector<int> meta;
  std::vector<Tile3> zt, yt, xt;
  int64_t off = 0;
  for (int s = 0; s < sp->ns; ++s) {
    const int j = sp->j0 + s;
    int64_t mm = std::min(shell_mmax(j, step), mgrid);
    int rmax = (int)std::min<int64_t>(isqrt64(mm), hN);
    const int xlim = std::min(rmax + 1, Nh);
    const int kxw = (int)(cdiv(xlim, kt) * kt);
    sp->mmax.push_back(mm);
    sp->rmax.push_back(rmax);
    sp->kxw.push_back(kxw);
    sp->off.push_back(off);
    meta.insert(meta.end(), {j, rmax, kxw, 0});
    off += (int64_t)N * N * kxw;
    for (int ky = 0; ky < N; ++ky) {
      const int kys = ky <= hN ? ky : ky - N;
      if (std::abs(kys) > rmax) continue;
      for (int kx0 = 0; kx0 < xlim; kx0 += kt)
        if ((int64_t)kx0 * kx0 + (int64_t)kys * kys <= mm) zt.push_back({s, ky, kx0, 0});
    }
    // the z pass wrote the rows kx0^2 + ky^2 <= mm of kx tile kx0: |ky| <= kyl
    for (int z = 0; z < N; ++z)
      for (int kx0 = 0; kx0 < xlim; kx0 += kt)
        yt.push_back({s, z, kx0, (int)isqrt64(mm - (int64_t)kx0 * kx0)});
    for (int z = 0; z < N; ++z)
      for (int y0 = 0; y0 < N; y0 += 2 * xpairs) xt.push_back({s, z, y0, 0});
  }
  sp->t_elems = off;
  // fused x pass: line p holds shells 2p (real part) and 2p+1 (imaginary part);
  // kx >= min(rmax + 1, N/2 + 1) is zero in T2
  {
    std::vector<int> items, lkxl;
    sp->nxlines = (sp->ns + 1) / 2;
    for (int p = 0; p < sp->nxlines; ++p) {
      int k = 0;
      for (int s = 2 * p; s < std::min(2 * p + 2, sp->ns); ++s) k = std::max(k, std::min(sp->rmax[s] + 1, Nh));
      lkxl.push_back(k);
      for (int kx = 0; kx < k; ++kx) items.insert(items.end(), {p, kx});
    }
    sp->nxitems = (int)(items.size() / 2);
    PSX_TRY(upload(ctx, items, &sp->d_xitems));
    PSX_TRY(upload(ctx, lkxl, &sp->d_xline_kxl));
  }
  sp->nz_tiles = (int)zt.size();
  sp->ny_tiles = (int)yt.size();
  sp->nx_tiles = (int)xt.size();
  PSX_TRY(upload(ctx, zt, &sp->d_ztiles));
  PSX_TRY(upload(ctx, yt, &sp->d_ytiles));
  PSX_TRY(upload(ctx, xt, &sp->d_xtiles));
  PSX_TRY(upload(ctx, meta, &sp->d_meta));
  PSX_TRY(upload(ctx, sp->off, &sp->d_off));
  *out = sp.get();
  ctx->shell_plans[key] = std::move(sp);
  return PSX_OK;
}

// triangle_list, triangle_cover: psx_plan.h


int get_contract_plan(psx_ctx* ctx, int prec, int step, int ncut, int nmax,
                      const ContractPlan** out);

// Split-BF16 contraction: the LDS serves the 16-byte A reads of 8 rows of a
// pair tile per cycle, conflict-free when their distinct shells differ
// modulo 8 (float rows 8 words apart modulo 64 banks).  Swap rows between
// the two 8-row halves of each tile while that lowers the number of shells
// that share a bank group with another shell of their half (i and j
// columns).  Row order inside a tile does not change any result.
static void arrange_half_tiles(std::vector<int>* prow, int nsh, int j0) {
  auto half_cost = [&](const int* r) {
    int c = 0;
    for (int col = 0; col < 2; ++col) {
      int seen[64], nseen = 0, grp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int k = 0; k < 8; ++k) {
        if (r[k] < 0) continue;
        const int sh = (col == 0 ? r[k] / nsh : r[k] % nsh) - j0;
        bool dup = false;
        for (int m = 0; m < nseen; ++m) dup |= seen[m] == sh;
        if (dup) continue;
        seen[nseen++] = sh;
        c += grp[sh & 7]++ > 0 ? 1 : 0;
      }
    }
    return c;
  };
  for (size_t t = 0; t + 16 <= prow->size(); t += 16) {
    int* r = prow->data() + t;
    int cur = half_cost(r) + half_cost(r + 8);
    for (bool improved = true; improved;) {
      improved = false;
      for (int a = 0; a < 8; ++a)
        for (int b = 8; b < 16; ++b) {
          std::swap(r[a], r[b]);
          const int c2 = half_cost(r) + half_cost(r + 8);
          if (c2 < cur) { cur = c2; improved = true; }
          else std::swap(r[a], r[b]);
        }
    }
  }
}

int get_contract_plan(psx_ctx* ctx, int prec, int step, int ncut, int nmax,
                      const ContractPlan** out) {
  const int cprec = prec;                         // plan code (precision + layout)
  const bool fused = (prec & CT_LAYOUT_FUSED) != 0;
  prec &= CT_LAYOUT_FUSED - 1;
  auto key = std::make_tuple(cprec, step, ncut, nmax);
  auto it = ctx->contract_plans.find(key);
  if (it != ctx->contract_plans.end()) { *out = it->second.get(); return PSX_OK; }
  auto cp = std::make_unique<ContractPlan>();
  const int j0 = ncut / step;
  const int ns = nmax - j0 + 1;
  cp->j0 = j0; cp->ns = ns; cp->nmax = nmax;
  cp->nlt = (int)cdiv(ns, 16);
  if (cp->nlt > 4)
    return set_error(PSX_ERR_UNSUPPORTED, "at most 64 shells (Nmax - Ncut//step + 1 <= 64)");
  std::vector<int> ijl;
  triangle_list(step, ncut, nmax, &ijl);
  cp->ntri = (int)(ijl.size() / 3);
  std::vector<int> tpair, tshell;
  triangle_cover(ijl, nmax, j0, &tpair, &tshell);
  for (int t = 0; t < cp->ntri; ++t)
    PSX_REQUIRE(tpair[t] >= 0, "contraction plan: a triangle without a chosen pair");
  // pair tiles: pairs grouped by their set of column tiles, whole tiles of one
  // set first, the remainders sorted by (tiles used, set) and cut into tiles
  std::map<int, int> pmask;  // pair key -> column tiles it reads
  for (int t = 0; t < cp->ntri; ++t) pmask[tpair[t]] |= 1 << ((tshell[t] - j0) / 16);
  std::map<int, std::vector<int>> by_mask;
  for (const auto& pm : pmask) by_mask[pm.second].push_back(pm.first);
  std::vector<int> prow;  // pair key of each row, -1 = padding
  std::vector<std::pair<int, int>> rest;
  for (const auto& bm : by_mask) {
    const size_t whole = bm.second.size() / 16 * 16;
    prow.insert(prow.end(), bm.second.begin(), bm.second.begin() + whole);
    for (size_t k = whole; k < bm.second.size(); ++k) rest.push_back({bm.first, bm.second[k]});
  }
  std::stable_sort(rest.begin(), rest.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    const int pa = __builtin_popcount(a.first), pb = __builtin_popcount(b.first);
    return pa != pb ? pa < pb : a.first < b.first;
  });
  for (const auto& r : rest) prow.push_back(r.second);
  cp->npairs = (int)prow.size();
  cp->npt = (int)cdiv(cp->npairs, 16);
  prow.resize((size_t)cp->npt * 16, -1);
  const int nsh = nmax + 1;
  if (prec != PSX_PREC_F64) arrange_half_tiles(&prow, nsh, j0);
  std::vector<int> pair_s(2 * prow.size(), -1);
  std::map<int, int> prow_of;
  for (size_t r = 0; r < prow.size(); ++r) {
    if (prow[r] < 0) continue;
    pair_s[2 * r] = prow[r] / nsh - j0;
    pair_s[2 * r + 1] = prow[r] % nsh - j0;
    prow_of[prow[r]] = (int)r;
  }
  std::vector<int> ptmask(cp->npt, 0);
  std::vector<int> tpt(cp->ntri), trow(cp->ntri), tlt(cp->ntri), tcol(cp->ntri);
  for (int t = 0; t < cp->ntri; ++t) {
    const int p = prow_of.at(tpair[t]);
    tpt[t] = p / 16; trow[t] = p % 16;
    tlt[t] = (tshell[t] - j0) / 16; tcol[t] = (tshell[t] - j0) % 16;
    ptmask[tpt[t]] |= 1 << tlt[t];
  }
  cp->nblk = 0;
  for (int m : ptmask) cp->nblk += __builtin_popcount(m);
  // assign pair tiles to waves: groups of WPG waves x maxpt slots, greedy by load
  const int maxpt = fused ? xc_maxpt() : contract_maxpt(prec, cp->nlt);
  const int WPG = fused ? xc_waves_per_group() : contract_waves_per_group(prec, cp->nlt);
  const int ngroups = (int)cdiv(cp->npt, WPG * maxpt);
  const int nw = ngroups * WPG;
  std::vector<int> order(cp->npt);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return __builtin_popcount(ptmask[a]) > __builtin_popcount(ptmask[b]);
  });
  std::vector<int> load(nw, 0), cnt(nw, 0);
  std::vector<int> wave_pt(nw * maxpt, -1), wave_mask(nw * maxpt, 0);
  std::vector<int> pt_wave(cp->npt), pt_q(cp->npt);
  for (int pt : order) {
    int best = -1;
    for (int w = 0; w < nw; ++w)
      if (cnt[w] < maxpt && (best < 0 || load[w] < load[best])) best = w;
    wave_pt[best * maxpt + cnt[best]] = pt;
    wave_mask[best * maxpt + cnt[best]] = ptmask[pt];
    pt_wave[pt] = best; pt_q[pt] = cnt[best];
    cnt[best]++;
    // cost of a tile: forming and splitting its products (about three
    // blocks' worth of MFMA time) plus its output blocks
    load[best] += 3 + __builtin_popcount(ptmask[pt]);
  }
  // Fused layout (one workgroup per CU): the waves of the workgroup are
  // spread over the CU's four SIMDs (wave w on SIMD w % 4), and a SIMD issues
  // its waves' VALU and MFMA work one at a time, so what bounds the group is
  // the busiest SIMD.  Swap tiles (or a tile and an empty slot) between waves
  // of different SIMDs while that lowers the group's (largest SIMD load, sum
  // of squared SIMD loads).  Which wave runs a tile changes no result: its
  // sums are formed in the same order.  (Measured at Ngrid 360, 40 shells:
  // k_xcontract 6.13 -> 5.95 ms; the staged k_contract, several workgroups
  // per CU, ran 4-6 % slower with it and keeps the per-wave balance.)
  if (fused && WPG % 4 == 0) {
    auto tcost = [&](int pt) { return pt < 0 ? 0 : 3 + __builtin_popcount(ptmask[pt]); };
    for (int g = 0; g < ngroups; ++g) {
      const int w0 = g * WPG;
      auto score = [&](long long* sq) {
        int sl[4] = {0, 0, 0, 0};
        for (int w = w0; w < w0 + WPG; ++w) sl[(w - w0) % 4] += load[w];
        *sq = 0;
        int mx = 0;
        for (int s = 0; s < 4; ++s) { mx = std::max(mx, sl[s]); *sq += (long long)sl[s] * sl[s]; }
        return mx;
      };
      long long sq;
      int mx = score(&sq);
      for (bool improved = true; improved;) {
        improved = false;
        for (int a = w0 * maxpt; a < (w0 + WPG) * maxpt; ++a)
          for (int b = a + 1; b < (w0 + WPG) * maxpt; ++b) {
            const int wa = a / maxpt, wb = b / maxpt;
            if ((wa - w0) % 4 == (wb - w0) % 4 || wave_pt[a] == wave_pt[b]) continue;
            const int d = tcost(wave_pt[b]) - tcost(wave_pt[a]);
            load[wa] += d; load[wb] -= d;
            long long sq2;
            const int mx2 = score(&sq2);
            if (mx2 < mx || (mx2 == mx && sq2 < sq)) {
              std::swap(wave_pt[a], wave_pt[b]);
              std::swap(wave_mask[a], wave_mask[b]);
              mx = mx2; sq = sq2; improved = true;
            } else {
              load[wa] -= d; load[wb] += d;
            }
          }
      }
    }
    // slots in use first within each wave (the kernels count a wave's tiles)
    for (int w = 0; w < nw; ++w)
      for (int q = 0; q < maxpt; ++q)
        for (int r = q + 1; r < maxpt && wave_pt[w * maxpt + q] < 0; ++r)
          if (wave_pt[w * maxpt + r] >= 0) {
            std::swap(wave_pt[w * maxpt + q], wave_pt[w * maxpt + r]);
            std::swap(wave_mask[w * maxpt + q], wave_mask[w * maxpt + r]);
          }
    for (int w = 0; w < nw; ++w)
      for (int q = 0; q < maxpt; ++q) {
        const int pt = wave_pt[w * maxpt + q];
        if (pt >= 0) { pt_wave[pt] = w; pt_q[pt] = q; }
      }
  }
  // output slots per group: the blocks that hold a triangle
  std::vector<int> wave_slot(nw * maxpt * 4, -1);
  std::vector<int> gslots(ngroups, 0);
  for (int w = 0; w < nw; ++w)
    for (int q = 0; q < maxpt; ++q) {
      const int pt = wave_pt[w * maxpt + q];
      for (int t = 0; t < 4; ++t)
        if (pt >= 0 && ((ptmask[pt] >> t) & 1)) wave_slot[(w * maxpt + q) * 4 + t] = gslots[w / WPG]++;
    }
  cp->nslot = *std::max_element(gslots.begin(), gslots.end());
  std::vector<int> tri_loc(cp->ntri), tri_grp(cp->ntri);
  for (int t = 0; t < cp->ntri; ++t) {
    const int w = pt_wave[tpt[t]], q = pt_q[tpt[t]];
    const int slot = wave_slot[(w * maxpt + q) * 4 + tlt[t]];
    const int row = trow[t], col = tcol[t];
    // 16x16 MFMA output layout (measured on gfx950): f32 outputs (f32 and bf16
    // inputs) hold row r in lane col + 16*(r/4), register r%4; the f64 MFMA
    // holds it in lane col + 16*(r%4), register r/4.
    const bool f64 = prec == PSX_PREC_F64;
    const int lane = col + 16 * (f64 ? row % 4 : row / 4);
    const int reg = f64 ? row / 4 : row % 4;
    tri_loc[t] = slot * 256 + lane * 4 + reg;
    tri_grp[t] = w / WPG;
  }
  cp->waves = nw;
  cp->maxpt = maxpt;
  PSX_TRY(upload(ctx, wave_pt, &cp->d_wave_pt));
  PSX_TRY(upload(ctx, wave_mask, &cp->d_wave_mask));
  PSX_TRY(upload(ctx, wave_slot, &cp->d_wave_slot));
  PSX_TRY(upload(ctx, pair_s, &cp->d_pair_s));
  PSX_TRY(upload(ctx, tri_loc, &cp->d_tri_loc));
  PSX_TRY(upload(ctx, tri_grp, &cp->d_tri_grp));
  PSX_TRY(upload(ctx, ijl, &cp->d_tri_ijl));
  *out = cp.get();
  ctx->contract_plans[key] = std::move(cp);
  return PSX_OK;
}

// ------------------------------------------------------------ pipelines ----
static int check_bk_args(int N, int step, int ncut, int nmax) {
  PSX_TRY(check_ngrid(N));
  PSX_REQUIRE(step >= 1 && ncut >= 0 && nmax >= 1, "step >= 1, Ncut >= 0, Nmax >= 1 required");
  PSX_REQUIRE(ncut / step <= nmax, "Ncut//step must not exceed Nmax");
  PSX_REQUIRE(nmax - ncut / step + 1 <= 64, "at most 64 shells (Nmax - Ncut//step + 1 <= 64)");
  return PSX_OK;
}

// Assignment, backward FFT and fcomb into the half field (FFT_periodic and the
// grid part of FFT_survey_mono share it; fcomb_survey is fcomb_periodic with
// cf = 1/(6^3*4), i.e. sumw = 1).
static int mesh_to_half(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int N,
                        PcsMap pm, double sumw, double wsum_abs, void* half, int prec) {
  const int fp = fft_prec(prec);
  const int64_t G = (int64_t)N * N * N;
  void* grid;
  PSX_TRY(ctx_buffer(ctx, "fft_grid", G * cbytes(fp), &grid));
  const double scale = ldexp(1.0, fixed_point_exponent(wsum_abs));
  // assignment with the x pass of the backward FFT fused, the y pass, then
  // the z pass with fcomb as its epilogue (writes the half field)
  PSX_TRY(launch_assign_grid(ctx, xyz, w, np, N, pm, scale, grid, fp, true));
  PSX_TRY(launch_fft_axis_batch(ctx, grid, N, 1, fp, 1, 1));
  return launch_fft_zcomb_batch(ctx, grid, half, N, 1, &sumw, fp);
}

static int fft_periodic_impl(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int N,
                             double lbox, double sumw, double wsum_abs, void* half, int prec) {
  PcsMap pm;
  pm.kf_ks = (float)((double)N / lbox);           // np.float32(float(Ngrid)/Lbox)
  pm.offset = 0.0f;
  pm.lo = 0.0f;                                   // np.clip(xyz, 0, Lbox*(1-1e-6)), float32
  pm.hi = (float)(lbox * (1.0 - 1e-6));
  return mesh_to_half(ctx, xyz, w, np, N, pm, sumw, wsum_abs, half, prec);
}

// The side stream: fork_side orders it after the work queued on the main
// stream so far; join_side makes the main stream wait for what was queued
// on it.  Both are stream operations only (a HIP graph capture of the main
// stream follows the fork and the join).
static int fork_side(psx_ctx* ctx) {
  if (!ctx->side) {
    PSX_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    PSX_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    PSX_HIP(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
  }
  PSX_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
  PSX_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
  return PSX_OK;
}

static int join_side(psx_ctx* ctx) {
  PSX_HIP(hipEventRecord(ctx->join_ev, ctx->side));
  PSX_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
  return PSX_OK;
}

// The shell powers p0k of nbox half fields read only the half fields: they
// run on the side stream beside the shell passes and the contraction (a
// latency-bound kernel next to HBM- and issue-bound ones).  The caller joins
// with join_side before p0k is read (on an error here it is joined already).
static int p0k_on_side(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                       int fp, double* p0k, int nbox) {
  void* bins;
  PSX_TRY(ctx_buffer(ctx, "p0k_bins", (size_t)nbox * 3 * nmax * sizeof(double), &bins));
  PSX_TRY(fork_side(ctx));
  const hipStream_t main_stream = ctx->stream;
  ctx->stream = ctx->side;
  int rc = launch_bin_power(ctx, half, N, step, nmax, 1.0, fp, (double*)bins, nbox);
  if (rc == PSX_OK) rc = launch_p0k(ctx, (const double*)bins, nmax, ncut / step, p0k, nbox);
  ctx->stream = main_stream;
  if (rc != PSX_OK) (void)join_side(ctx);
  return rc;
}

static int bk_raw_main(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                       double* braw, int prec, int fp, int ns, int64_t G);

static int bk_raw_impl(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                       double* braw, double* p0k, int prec) {
  const int fp = fft_prec(prec);
  const int j0 = ncut / step;
  const int ns = nmax - j0 + 1;
  const int64_t G = (int64_t)N * N * N;
  if (p0k) PSX_TRY(p0k_on_side(ctx, half, N, step, ncut, nmax, fp, p0k, 1));
  const int rc = bk_raw_main(ctx, half, N, step, ncut, nmax, braw, prec, fp, ns, G);
  if (p0k) PSX_TRY(join_side(ctx));
  return rc;
}

static int bk_raw_main(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                       double* braw, int prec, int fp, int ns, int64_t G) {
  if (xcontract_supported(ctx, N, step, ncut, nmax, prec) == PSX_OK) {
    // split-BF16 with the shell x pass fused into the contraction: the shell
    // fields never go through HBM
    PSX_TRY(launch_xcontract(ctx, half, N, step, ncut, nmax, braw));
  } else {
    void* D;
    PSX_TRY(ctx_buffer(ctx, "shells", (size_t)ns * G * rbytes(fp), &D));
    PSX_TRY(launch_shell_fields(ctx, half, N, step, ncut, nmax, fp, D));
    PSX_TRY(launch_contract(ctx, D, N, step, ncut, nmax, prec, braw));
  }
  return PSX_OK;
}

}  // namespace psx

using namespace psx;

bool psx::timing_events(psx_ctx* ctx, const char* name, hipEvent_t* start, hipEvent_t* stop) {
  if (!ctx->timing) return false;
  if (ctx->kev_used == ctx->kev.size()) {
    psx_ctx::KernelEvents e{name, nullptr, nullptr};
    if (hipEventCreate(&e.start) != hipSuccess) return false;
    if (hipEventCreate(&e.stop) != hipSuccess) {
      (void)hipEventDestroy(e.start);
      return false;
    }
    ctx->kev.push_back(e);
  }
  psx_ctx::KernelEvents& e = ctx->kev[ctx->kev_used++];
  e.name = name;
  *start = e.start;
  *stop = e.stop;
  return true;
}

// ================================================================= C ABI ===
extern "C" {

int psx_abi_version(void) { return 2; }

const char* psx_last_error(void) { return last_error_string().c_str(); }

int psx_ctx_create(psx_ctx** out, int device) {
  PSX_REQUIRE(out != nullptr, "null ctx pointer");
  int n = 0;
  PSX_HIP(hipGetDeviceCount(&n));
  PSX_REQUIRE(device >= 0 && device < n, "device index out of range");
  PSX_HIP(hipSetDevice(device));
  psx_ctx* c = new psx_ctx();
  c->device = device;
  const hipError_t e = hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    return set_error(PSX_ERR_HIP, std::string("hipEventCreateWithFlags: ") + hipGetErrorString(e));
  }
  *out = c;
  return PSX_OK;
}

int psx_ctx_destroy(psx_ctx* ctx) {
  if (!ctx) return PSX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->bufs) if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (void* p : ctx->owned) (void)hipFree(p);
  if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  for (auto& e : ctx->kev) {
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  delete ctx;
  return PSX_OK;
}

int psx_ctx_set_stream(psx_ctx* ctx, void* stream) {
  PSX_REQUIRE(ctx, "null ctx");
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  const hipStream_t s = (hipStream_t)stream;
  if (s != ctx->stream) {
    // work queued on the old stream (which may use the workspace) completes
    // before anything the new stream runs after this point
    PSX_HIP(hipEventRecord(ctx->order_ev, ctx->stream));
    PSX_HIP(hipStreamWaitEvent(s, ctx->order_ev, 0));
    ctx->stream = s;
  }
  return PSX_OK;
}

int psx_ctx_synchronize(psx_ctx* ctx) {
  PSX_REQUIRE(ctx, "null ctx");
  PSX_CTX_LOCK(ctx);
  PSX_HIP(hipStreamSynchronize(ctx->stream));
  return PSX_OK;
}

int psx_ctx_workspace_bytes(psx_ctx* ctx, int64_t* bytes) {
  PSX_REQUIRE(ctx && bytes, "null argument");
  PSX_CTX_LOCK(ctx);
  int64_t s = 0;
  for (auto& kv : ctx->bufs) s += (int64_t)kv.second.bytes;
  *bytes = s;
  return PSX_OK;
}

int psx_ctx_kernel_timing(psx_ctx* ctx, int enable) {
  PSX_REQUIRE(ctx, "null ctx");
  PSX_CTX_LOCK(ctx);
  PSX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timing = enable != 0;
  ctx->kev_used = 0;
  return PSX_OK;
}

int psx_ctx_kernel_time(psx_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms) {
  PSX_REQUIRE(ctx && kernel && launches && total_ms, "null argument");
  PSX_CTX_LOCK(ctx);
  int64_t n = 0;
  double ms = 0.0;
  for (size_t i = 0; i < ctx->kev_used; ++i) {
    const psx_ctx::KernelEvents& e = ctx->kev[i];
    if (std::strcmp(e.name, kernel) != 0) continue;
    PSX_HIP(hipEventSynchronize(e.stop));
    float t = 0.f;
    PSX_HIP(hipEventElapsedTime(&t, e.start, e.stop));
    ms += t;
    ++n;
  }
  *launches = n;
  *total_ms = ms;
  return PSX_OK;
}

int psx_ctx_release_workspace(psx_ctx* ctx) {
  PSX_REQUIRE(ctx, "null ctx");
  PSX_CTX_LOCK(ctx);
  PSX_HIP(hipSetDevice(ctx->device));
  PSX_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& kv : ctx->bufs) {
    if (kv.second.ptr) PSX_HIP(hipFree(kv.second.ptr));
    kv.second.ptr = nullptr;
    kv.second.bytes = 0;
  }
  ctx->bufs.clear();
  return PSX_OK;
}

int psx_assign_quad(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                    float kf_ks, float offset, float clip_lo, float clip_hi, double wsum_abs,
                    void* grid_out, int prec) {
  PSX_REQUIRE(ctx && grid_out && (xyz || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(np >= 0, "negative particle count");
  PSX_REQUIRE(clip_lo <= clip_hi && kf_ks > 0.0f, "bad clip range or kf_ks");
  PSX_TRY(check_ngrid(ngrid));
  // every clipped position must map to rx in [0, N + 1.5]: the cell indices
  // then wrap into the grid
  const double rlo = (double)kf_ks * clip_lo + 1.0 + offset, rhi = (double)kf_ks * clip_hi + 1.0 + offset;
  PSX_REQUIRE(rlo >= 0.0 && rhi <= ngrid + 1.5, "clip range maps outside the grid (kf_ks, offset)");
  const double scale = ldexp(1.0, fixed_point_exponent(wsum_abs));
  PcsMap pm;
  pm.kf_ks = kf_ks; pm.offset = offset; pm.lo = clip_lo; pm.hi = clip_hi;
  return launch_assign_grid(ctx, xyz, w, np, ngrid, pm, scale, grid_out, fft_prec(prec), false);
}

int psx_fft3d(psx_ctx* ctx, void* grid, int ngrid, int sign, int prec) {
  PSX_REQUIRE(ctx && grid, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(sign == 1 || sign == -1, "sign must be +1 or -1");
  PSX_TRY(check_ngrid(ngrid));
  return launch_fft3d(ctx, grid, ngrid, sign, fft_prec(prec));
}

int psx_fcomb_periodic(psx_ctx* ctx, const void* grid, void* half_out, int ngrid, double sumw,
                       int prec) {
  PSX_REQUIRE(ctx && grid && half_out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_ngrid(ngrid));
  const FftPlan* ft;
  PSX_TRY(get_fft_plan(ctx, ngrid, &ft));
  return launch_fcomb(ctx, grid, half_out, ngrid, sumw, fft_prec(prec), ft->rec, ft->Wf, ft->Wd);
}

int psx_fft_periodic(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                     double lbox, double sumw, double wsum_abs, void* half_out, int prec) {
  PSX_REQUIRE(ctx && half_out && (xyz || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(np >= 0 && lbox > 0, "bad particle count or box size");
  PSX_TRY(check_ngrid(ngrid));
  return fft_periodic_impl(ctx, xyz, w, np, ngrid, lbox, sumw, wsum_abs, half_out, prec);
}

int psx_quad_weights(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ia, int ib,
                     int ic, int id, float* w_out) {
  PSX_REQUIRE(ctx && w_out && (xyz || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(np >= 0, "bad particle count");
  const bool none = ia == 0 && ib == 0 && ic == 0 && id == 0;
  const bool qij = ia >= 1 && ia <= 3 && ib >= 1 && ib <= 3 && ic == 0 && id == 0;
  const bool qijkl = ia >= 1 && ia <= 3 && ib >= 1 && ib <= 3 && ic >= 1 && ic <= 3 && id >= 1 &&
                     id <= 3;
  PSX_REQUIRE(none || qij || qijkl, "ia, ib, ic, id: all 0, (i, j, 0, 0) or (i, j, k, l) in 1..3");
  return launch_quad_weights(ctx, xyz, w, np, ia, ib, ic, id, w_out);
}

int psx_five_delta2g_1(psx_ctx* ctx, void* dcgxx, const void* dcgyy, const void* dcgzz, int ngrid,
                       int unit_vectors, int prec) {
  PSX_REQUIRE(ctx && dcgxx && dcgyy && dcgzz, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(ngrid >= 2 && ngrid <= 4096, "Ngrid must be in [2, 4096]");
  return launch_five_delta2_1(ctx, dcgxx, dcgyy, dcgzz, ngrid, unit_vectors ? 1 : 0, fft_prec(prec));
}

int psx_five_delta2g_2(psx_ctx* ctx, const void* dcg, void* dcgxx, const void* dcgxy,
                       const void* dcgyz, const void* dcgzx, int ngrid, int unit_vectors, int prec) {
  PSX_REQUIRE(ctx && dcg && dcgxx && dcgxy && dcgyz && dcgzx, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(ngrid >= 2 && ngrid <= 4096, "Ngrid must be in [2, 4096]");
  return launch_five_delta2_2(ctx, dcg, dcgxx, dcgxy, dcgyz, dcgzx, ngrid, unit_vectors ? 1 : 0,
                              fft_prec(prec));
}

int psx_build_quad(psx_ctx* ctx, const void* dclr1, void* dclr2, int irsd, int ngrid, int prec) {
  PSX_REQUIRE(ctx && dclr1 && dclr2, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(irsd >= 1 && irsd <= 3, "irsd must be 1, 2 or 3 (line of sight x, y, z)");
  PSX_REQUIRE(ngrid >= 2 && ngrid <= 4096, "Ngrid must be in [2, 4096]");
  return launch_build_quad(ctx, dclr1, dclr2, irsd, ngrid, fft_prec(prec));
}

int psx_survey_catalogue(psx_ctx* ctx, const double* radecz, const double* nb, const double* w,
                         int64_t n, const double* dist_nodes, const double* dist_slopes, int nseg,
                         double zmax, double h, double p0_fkp, float* xyz_out, float* w_out,
                         double* sums_out) {
  PSX_REQUIRE(ctx && sums_out && dist_nodes && dist_slopes, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(n == 0 || (radecz && nb && xyz_out && w_out), "null argument");
  PSX_REQUIRE(n >= 0 && nseg >= 1 && zmax >= 0.0, "bad object count, table size or zmax");
  return launch_survey_catalogue(ctx, radecz, nb, w, n, dist_nodes, dist_slopes, nseg, zmax, h,
                                 p0_fkp, xyz_out, w_out, sums_out);
}

int psx_fft_survey_mono(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                        double lbox, double wsum_abs, void* half_out, int prec) {
  PSX_REQUIRE(ctx && half_out && (xyz || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(np >= 0 && lbox > 0, "bad particle count or box size");
  PSX_TRY(check_ngrid(ngrid));
  PcsMap pm;
  pm.kf_ks = (float)((double)ngrid / lbox);       // np.float32(float(Ngrid)/Lbox)
  pm.offset = (float)(0.5 * ngrid);               // 0.5*Ngrid (pyspectrum.py:819)
  // positions are not clipped by the reference (|x| < Lbox/2 is asserted,
  // pyspectrum.py:785); the clip only keeps invalid input inside the grid
  pm.lo = (float)(-0.5 * lbox);
  pm.hi = (float)(0.5 * lbox);
  return mesh_to_half(ctx, xyz, w, np, ngrid, pm, 1.0, wsum_abs, half_out, prec);
}

int psx_reflect_delta(psx_ctx* ctx, const void* half, void* full_out, int ngrid, int prec) {
  PSX_REQUIRE(ctx && half && full_out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_ngrid(ngrid));
  return launch_reflect(ctx, half, full_out, ngrid, fft_prec(prec));
}

int psx_hermitian_half(psx_ctx* ctx, const void* full, int in_prec, int64_t sx, int64_t sy,
                       int64_t sz, void* half_out, int ngrid, int prec) {
  PSX_REQUIRE(ctx && full && half_out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec) && prec_ok(in_prec), "unknown precision");
  PSX_TRY(check_ngrid(ngrid));
  return launch_herm_half(ctx, full, fft_prec(in_prec), sx, sy, sz, half_out, ngrid, fft_prec(prec));
}

int psx_bin_power(psx_ctx* ctx, const void* half, int ngrid, int step, int nbins, double kf,
                  double* out3, int prec) {
  PSX_REQUIRE(ctx && half && out3, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(step >= 1, "step >= 1 required");
  PSX_TRY(check_ngrid(ngrid));
  return launch_bin_power(ctx, half, ngrid, step, nbins, kf, fft_prec(prec), out3);
}

int psx_bin_power_noalloc(psx_ctx* ctx, const void* half, int ngrid, int step, int nbins,
                          double kf, double* out3, int prec) {
  PSX_REQUIRE(ctx && half && out3, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(step >= 1, "step >= 1 required");
  PSX_TRY(check_ngrid(ngrid));
  return launch_bin_power_noalloc(ctx, half, ngrid, step, nbins, kf, fft_prec(prec), out3);
}

int psx_mock_gaussian(psx_ctx* ctx, int mesh, double lbox, uint64_t seed, double slope,
                      double kcut, float* g_out, double* sigma_out) {
  PSX_REQUIRE(ctx && g_out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_TRY(check_ngrid(mesh));
  PSX_REQUIRE(lbox > 0 && kcut > 0, "lbox and kcut must be positive");
  return launch_mock_gaussian(ctx, mesh, lbox, seed, slope, kcut, g_out, sigma_out);
}

int psx_mock_sample(psx_ctx* ctx, const float* g, int mesh, double lbox, int64_t np,
                    uint64_t seed, double sigma_v, double rsd_fac, int los, float* xyz_out) {
  PSX_REQUIRE(ctx && g && (xyz_out || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_TRY(check_ngrid(mesh));
  PSX_REQUIRE(np >= 0 && lbox > 0 && sigma_v >= 0, "np >= 0, lbox > 0 and sigma_v >= 0 required");
  PSX_REQUIRE(los >= 0 && los <= 2, "los must be 0, 1 or 2");
  return launch_mock_sample(ctx, g, mesh, lbox, np, seed, sigma_v, rsd_fac, los, xyz_out);
}

int psx_mock_velocities(psx_ctx* ctx, int64_t np, uint64_t seed, double sigma, float* v_out) {
  PSX_REQUIRE(ctx && (v_out || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(np >= 0, "negative count");
  return launch_mock_velocities(ctx, np, seed, sigma, v_out);
}

int psx_apply_rsd(psx_ctx* ctx, const float* xyz, const float* v, int64_t np, double rsd_fac,
                  double lbox, int los, float* xyz_out) {
  PSX_REQUIRE(ctx && ((xyz && v && xyz_out) || np == 0), "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(np >= 0 && lbox > 0, "np >= 0 and lbox > 0 required");
  PSX_REQUIRE(los >= 0 && los <= 2, "los must be 0, 1 or 2");
  return launch_apply_rsd(ctx, xyz, v, np, rsd_fac, lbox, los, xyz_out);
}

int psx_pk_pbox_rsd(psx_ctx* ctx, const void* half, int irsd, int lbox, int nbin, int nmu,
                    int ngrid, double* out, int prec) {
  PSX_REQUIRE(ctx && half && out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_ngrid(ngrid));
  return launch_pk_rsd(ctx, half, ngrid, irsd, lbox, nbin, nmu, fft_prec(prec), out);
}

int psx_bk_num_triangles(int step, int ncut, int nmax, int64_t* ntri) {
  PSX_REQUIRE(ntri, "null argument");
  PSX_REQUIRE(step >= 1 && ncut >= 0 && nmax >= 0, "bad step/Ncut/Nmax");
  std::vector<int> ijl;
  triangle_list(step, ncut, nmax, &ijl);
  *ntri = (int64_t)(ijl.size() / 3);
  return PSX_OK;
}

int psx_bk_triangles(int step, int ncut, int nmax, int32_t* out) {
  PSX_REQUIRE(out, "null argument");
  PSX_REQUIRE(step >= 1 && ncut >= 0 && nmax >= 0, "bad step/Ncut/Nmax");
  std::vector<int> ijl;
  triangle_list(step, ncut, nmax, &ijl);
  std::memcpy(out, ijl.data(), ijl.size() * sizeof(int32_t));
  return PSX_OK;
}

int psx_bk_shell_fields(psx_ctx* ctx, const void* half, int ngrid, int step, int ncut, int nmax,
                        void* d_out, int prec) {
  PSX_REQUIRE(ctx && half && d_out, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  return launch_shell_fields(ctx, half, ngrid, step, ncut, nmax, fft_prec(prec), d_out);
}

int psx_bk_contract(psx_ctx* ctx, const void* d, int ngrid, int step, int ncut, int nmax,
                    double* b_raw, int prec) {
  PSX_REQUIRE(ctx && d && b_raw, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  return launch_contract(ctx, d, ngrid, step, ncut, nmax, prec, b_raw);
}

int psx_bk_raw(psx_ctx* ctx, const void* half, int ngrid, int step, int ncut, int nmax,
               double* b_raw, double* p0k, int prec) {
  PSX_REQUIRE(ctx && half && b_raw, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  return bk_raw_impl(ctx, half, ngrid, step, ncut, nmax, b_raw, p0k, prec);
}

int psx_bk_workspace_bytes(psx_ctx* ctx, int ngrid, int step, int ncut, int nmax, int64_t np,
                           int prec, int64_t* bytes) {
  PSX_REQUIRE(ctx && bytes, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_REQUIRE(np >= 0, "negative particle count");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  const int fp = fft_prec(prec);
  const int64_t N = ngrid, G = N * N * N, NN = N * N, Nh = N / 2 + 1, hN = N / 2;
  const int j0 = ncut / step, ns = nmax - j0 + 1;
  int64_t b = G * (int64_t)cbytes(fp) + Nh * NN * (int64_t)cbytes(fp);   // fft_grid, bk_half
  b += 8 * (NN + 1) + 24 * np;                                             // assignment sort
  // shell intermediates T1, T2: per shell N^2 rows of kxw values (the plan's
  // kx extent rounded to the z/y tile width)
  const int64_t kt = fp == PSX_PREC_F64 ? TileCfg<double>::LINES : TileCfg<float>::LINES;
  const int64_t mgrid = 3 * hN * hN;
  int64_t tel = 0;
  for (int s = 0; s < ns; ++s) {
    const int64_t mm = std::min(shell_mmax(j0 + s, step), mgrid);
    const int64_t xlim = std::min(std::min<int64_t>(isqrt64(mm), hN) + 1, Nh);
    tel += NN * cdiv(xlim, kt) * kt;
  }
  b += 2 * tel * (int64_t)cbytes(fp);
  const bool fused = xcontract_supported(ctx, ngrid, step, ncut, nmax, prec) == PSX_OK;
  if (!fused) b += (int64_t)ns * G * (int64_t)rbytes(fp);                  // shell fields
  const ContractPlan* cp;
  PSX_TRY(get_contract_plan(ctx, fused ? prec | CT_LAYOUT_FUSED : prec, step, ncut, nmax, &cp));
  int cus = 0;
  PSX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int64_t ngroups =
      cp->waves / (fused ? xc_waves_per_group() : contract_waves_per_group(prec, cp->nlt));
  b += (8 * (int64_t)cus + ngroups) * cp->nslot * 256 * 8;                // partials, reduction
  b += (int64_t)(cp->ntri + 1) * 8 + 513 * 3 * (int64_t)nmax * 8;          // braw, p0k bins
  *bytes = b;
  return PSX_OK;
}

int psx_bk_counts(psx_ctx* ctx, int ngrid, int step, int ncut, int nmax, double* counts) {
  PSX_REQUIRE(ctx && counts, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  const int64_t Nh = ngrid / 2 + 1;
  const int64_t nh = Nh * ngrid * ngrid;
  void* ones;
  PSX_TRY(ctx_buffer(ctx, "unit_half", nh * 16, &ones));
  PSX_TRY(launch_fill_unit(ctx, (double*)ones, nh));
  return bk_raw_impl(ctx, ones, ngrid, step, ncut, nmax, counts, nullptr, PSX_PREC_F64);
}

int psx_bk_finalize(psx_ctx* ctx, const double* b_raw, const double* p0k, const double* counts,
                    const int32_t* ijl, int ntri, int ngrid, double nbar, double a3, double b3,
                    double a6, double b6, double* out7) {
  PSX_REQUIRE(ctx && b_raw && p0k && counts && ijl && out7, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(ntri >= 0, "negative triangle count");
  if (ntri == 0) return PSX_OK;
  const double G = (double)ngrid * (double)ngrid * (double)ngrid;
  return launch_bk_finalize(ctx, b_raw, p0k, counts, ijl, ntri, G, nbar, a3, b3, a6, b6, out7);
}

int psx_bk_periodic_batch(psx_ctx* ctx, int nbox, const psx_box* boxes, int ngrid, int step,
                          int ncut, int nmax, const double* counts, const int32_t* ijl, double a3,
                          double a6, double* out7, double* p0k, void* half_out, int prec) {
  PSX_REQUIRE(ctx && boxes && counts && ijl && out7 && p0k, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(nbox >= 1 && nbox <= PSX_MAX_BATCH, "nbox must be in [1, 16]");
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  const int N = ngrid, fp = fft_prec(prec);
  const int64_t G = (int64_t)N * N * N, nh = (int64_t)(N / 2 + 1) * N * N;
  std::vector<const float*> xyz(nbox), w(nbox);
  std::vector<int64_t> np(nbox);
  std::vector<PcsMap> pm(nbox);
  std::vector<double> scale(nbox), sumw(nbox), nbar(nbox), b3(nbox), b6(nbox);
  for (int b = 0; b < nbox; ++b) {
    const psx_box& x = boxes[b];
    PSX_REQUIRE(x.np >= 0 && (x.xyz || x.np == 0) && x.lbox > 0, "bad box");
    xyz[b] = x.xyz; w[b] = x.w; np[b] = x.np;
    // the position map of FFT_periodic (pyspectrum.py:920-951)
    pm[b].kf_ks = (float)((double)N / x.lbox);
    pm[b].offset = 0.0f;
    pm[b].lo = 0.0f;
    pm[b].hi = (float)(x.lbox * (1.0 - 1e-6));
    scale[b] = ldexp(1.0, fixed_point_exponent(x.wsum_abs));
    sumw[b] = x.sumw; nbar[b] = x.nbar; b3[b] = x.b3; b6[b] = x.b6;
  }
  void *grid, *half = half_out, *braw;
  PSX_TRY(ctx_buffer(ctx, "fft_grid", (size_t)nbox * G * cbytes(fp), &grid));
  if (!half) PSX_TRY(ctx_buffer(ctx, "bk_half", (size_t)nbox * nh * cbytes(fp), &half));
  std::vector<int> tl;
  triangle_list(step, ncut, nmax, &tl);
  const int ntri = (int)(tl.size() / 3);
  PSX_TRY(ctx_buffer(ctx, "bk_braw", (size_t)nbox * std::max(ntri, 1) * sizeof(double), &braw));
  // FFT_periodic for every box: assignment with the x pass fused, the y
  // pass, the z pass with fcomb fused
  PSX_TRY(launch_assign_grid_batch(ctx, nbox, xyz.data(), w.data(), np.data(), pm.data(),
                                   scale.data(), N, grid, fp, true));
  PSX_TRY(launch_fft_axis_batch(ctx, grid, N, 1, fp, 1, nbox));
  PSX_TRY(launch_fft_zcomb_batch(ctx, grid, half, N, nbox, sumw.data(), fp));
  // shell powers p0k on the side stream; shells and contraction here
  PSX_TRY(p0k_on_side(ctx, half, N, step, ncut, nmax, fp, p0k, nbox));
  const int rc = [&]() -> int {
    if (xcontract_supported(ctx, N, step, ncut, nmax, prec) == PSX_OK)
      return launch_xcontract(ctx, half, N, step, ncut, nmax, (double*)braw, nbox);
    const int ns = nmax - ncut / step + 1;
    void* D;
    PSX_TRY(ctx_buffer(ctx, "shells", (size_t)ns * G * rbytes(fp), &D));
    for (int b = 0; b < nbox; ++b) {
      PSX_TRY(launch_shell_fields(ctx, (const char*)half + (size_t)b * nh * cbytes(fp), N, step,
                                  ncut, nmax, fp, D));
      PSX_TRY(launch_contract(ctx, D, N, step, ncut, nmax, prec, (double*)braw + (size_t)b * ntri));
    }
    return PSX_OK;
  }();
  PSX_TRY(join_side(ctx));
  PSX_TRY(rc);
  // units, shot noise and Q
  if (ntri == 0) return PSX_OK;
  return launch_bk_finalize_batch(ctx, nbox, (const double*)braw, p0k, nmax, counts, ijl, ntri,
                                  (double)G, nbar.data(), a3, b3.data(), a6, b6.data(), out7);
}

int psx_bk_periodic(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                    double lbox, double sumw, double wsum_abs, int step, int ncut, int nmax,
                    const double* counts, const int32_t* ijl, double nbar, double a3, double b3,
                    double a6, double b6, double* out7, double* p0k, int prec) {
  PSX_REQUIRE(ctx && (xyz || np == 0) && counts && ijl && out7 && p0k, "null argument");
  PSX_CTX_LOCK(ctx);
  PSX_REQUIRE(prec_ok(prec), "unknown precision");
  PSX_TRY(check_bk_args(ngrid, step, ncut, nmax));
  const int fp = fft_prec(prec);
  const int64_t nh = (int64_t)(ngrid / 2 + 1) * ngrid * ngrid;
  void *half, *braw;
  PSX_TRY(ctx_buffer(ctx, "bk_half", nh * cbytes(fp), &half));
  std::vector<int> tl;
  triangle_list(step, ncut, nmax, &tl);
  const int ntri = (int)(tl.size() / 3);
  PSX_TRY(ctx_buffer(ctx, "bk_braw", (size_t)std::max(ntri, 1) * sizeof(double), &braw));
  PSX_TRY(fft_periodic_impl(ctx, xyz, w, np, ngrid, lbox, sumw, wsum_abs, half, prec));
  PSX_TRY(bk_raw_impl(ctx, half, ngrid, step, ncut, nmax, (double*)braw, p0k, prec));
  const double G = (double)ngrid * (double)ngrid * (double)ngrid;
  if (ntri == 0) return PSX_OK;
  return launch_bk_finalize(ctx, (const double*)braw, p0k, counts, ijl, ntri, G, nbar, a3, b3, a6,
                            b6, out7);
}

}  // extern "C"
