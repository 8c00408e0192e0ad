// Library context: device, stream, cached plans and a grow-only workspace.
#pragma once
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include <hip/hip_ext.h>

#include "psx_common.h"
#include "psx_fft.h"

struct psx_ctx;

// Boxes of one shape that one batched launch per stage processes
// (psx_bk_periodic_batch); per-box scalars travel as kernel arguments.
#define PSX_MAX_BATCH 16

namespace psx {

// Device buffer owned by the context (freed at ctx destroy).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Per-N twiddle table, both precisions.
struct FftPlan {
  FftShape shape;
  void* tw32 = nullptr;  // cplx<float>[N]
  void* tw64 = nullptr;  // cplx<double>[N]
  // fcomb tables, n = 0..N/2: twiddle recurrence (cplx<double>) and PCS window
  void* rec = nullptr;
  float* Wf = nullptr;
  double* Wd = nullptr;
};

// Row pairs (complex FFT lines) per tile of the shell fields' x pass.  Small
// tiles keep several workgroups per CU to hide the gather latency.  With NP
// pairs a line gets 256/NP lanes; the radix stages take ceil((N/R)/lanes)
// passes each, so the lane count decides how many lanes idle: at N = 360
// (radices 5, 9, 8: 72, 40, 45 butterflies) 25 lanes (10 pairs) waste far
// less than 32 (8 pairs), at N = 512 (8, 8, 8: 64 each) 32 lanes fit exactly.
// Float tiles take whichever of 8 or 10 pairs the cost model below prefers
// (measured: 10 pairs 2.5 % faster shell fields at 360, 13 % slower at 512);
// double tiles (LDS) and the runtime-N path use 8.  PSX_SHELL_X_PAIRS fixes
// the float choice for every N.
#ifndef PSX_SHELL_X_PAIRS_F64
#define PSX_SHELL_X_PAIRS_F64 8
#endif
// VALU issue of one line's radix stages with `lanes` lanes per line
// (approximate VALU per butterfly by radix).
__host__ __device__ constexpr int fft_line_cost(int N, int lanes) {
  const FftShape sh = fft_shape_of(N);
  int c = 0;
  for (int s = 0; s < sh.nstages; ++s) {
    const int R = sh.radix[s];
    const int passes = (N / R + lanes - 1) / lanes;
    const int cr = R == 9 ? 140 : R == 8 ? 90 : R == 5 ? 60 : R == 4 ? 35 : R == 3 ? 25 : 10;
    c += passes * cr;
  }
  return c * lanes;
}
// NC: compile-time grid size, 0 for the runtime-N path.
template <typename T> __host__ __device__ constexpr int shell_x_pairs(int NC) {
  if (sizeof(T) != 4) return PSX_SHELL_X_PAIRS_F64;
#ifdef PSX_SHELL_X_PAIRS
  return PSX_SHELL_X_PAIRS;
#else
  if (NC <= 0) return 8;
  return fft_line_cost(NC, FFT_THREADS / 10) < fft_line_cost(NC, FFT_THREADS / 8) ? 10 : 8;
#endif
}
// LDS row padding (complex values) of the x pass lines.
#ifndef PSX_XPAD
#define PSX_XPAD 1
#endif

// One pruned-pass tile: (shell slot, a, b) with meaning per pass.
struct Tile3 {
  int s, a, b, c;  // c: y pass, the largest |ky| of the rows the z pass wrote
};

// Shell-field plan for (Ngrid, step, Ncut, Nmax).
struct ShellPlan {
  int N = 0, step = 0, ncut = 0, nmax = 0, kt = 16;  // kt: kx tile width of the z/y passes
  int j0 = 0, ns = 0;                   // shells j0..nmax, ns of them
  std::vector<int64_t> mmax;            // per slot: max |k|^2 of the shell
  std::vector<int> rmax;                // per slot: max |k component|
  std::vector<int> kxw;                 // per slot: stored kx width (multiple of 16)
  std::vector<int64_t> off;             // per slot: element offset in the T buffers
  int64_t t_elems = 0;                  // total elements of T1 (= T2)
  int nz_tiles = 0, ny_tiles = 0, nx_tiles = 0;
  Tile3* d_ztiles = nullptr;
  Tile3* d_ytiles = nullptr;
  Tile3* d_xtiles = nullptr;
  int* d_meta = nullptr;                // per slot: j, rmax, kxw, (int)mmax -- device copy
  int64_t* d_off = nullptr;
  // x pass fused with the contraction: shells 2p and 2p+1 share FFT line p;
  // one item (p, kx) per nonzero input column kx < kxl of line p
  int nxlines = 0, nxitems = 0;
  int* d_xitems = nullptr;              // [nxitems][2] (p, kx)
  int* d_xline_kxl = nullptr;           // [nxlines] kxl of line p
};

// MFMA contraction plan (pairs x shells, 16x16 blocks restricted to triangles).
struct ContractPlan {
  int ns = 0, j0 = 0, nmax = 0;
  int npairs = 0, npt = 0, nlt = 0, nblk = 0, ntri = 0;
  int waves = 0, maxpt = 0;             // total waves (groups x waves/WG), pair tiles per wave
  int nslot = 0;                        // valid output blocks per group (max over groups)
  int* d_wave_pt = nullptr;             // [waves][maxpt] pair tile (-1 = none)
  int* d_wave_mask = nullptr;           // [waves][maxpt] bitmask of l tiles
  int* d_wave_slot = nullptr;           // [waves][maxpt][4] output slot of each block (-1 = none)
  int* d_pair_s = nullptr;              // [npt*16][2] shell slots (i-j0, j-j0) per row (-1 pad)
  int* d_tri_loc = nullptr;             // [ntri] slot*256 + lane*4 + reg
  int* d_tri_grp = nullptr;             // [ntri] group of the triangle's pair tile
  int* d_tri_ijl = nullptr;             // [ntri][3] shell indices i, j, l (reference loop order)
};

}  // namespace psx

struct psx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Entry points hold the lock; a stream switch makes the new stream wait on
  // an event recorded on the old one, so the workspace is never in use on two
  // streams at once.
  std::recursive_mutex mu;
  hipEvent_t order_ev = nullptr;
  // Side stream for independent work beside the main sequence (the shell
  // powers beside the shell passes): forked from and joined back into
  // `stream` with an event each, within one entry point.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::map<int, psx::FftPlan> fft_plans;
  std::map<std::tuple<int, int, int, int, int, int>, std::unique_ptr<psx::ShellPlan>> shell_plans;
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<psx::ContractPlan>> contract_plans;
  std::map<std::string, psx::DevBuf> bufs;
  std::vector<void*> owned;  // plan allocations
  // Kernel timing (psx_ctx_kernel_timing): event pairs recorded around the
  // main launches, reused from one enable to the next.
  struct KernelEvents {
    const char* name;
    hipEvent_t start, stop;
  };
  bool timing = false;
  std::vector<KernelEvents> kev;
  size_t kev_used = 0;
};

namespace psx {
// Grow-only named workspace buffer.
int ctx_buffer(psx_ctx* ctx, const char* name, size_t bytes, void** out);
int ctx_alloc_owned(psx_ctx* ctx, size_t bytes, void** out);
int get_fft_plan(psx_ctx* ctx, int N, const FftPlan** out);
int get_shell_plan(psx_ctx* ctx, int N, int step, int ncut, int nmax, int kt, int xpairs,
                   const ShellPlan** out);
// Plan code: the pipeline precision, plus CT_LAYOUT_FUSED for the wave layout
// of the split-BF16 contraction fused with the shell x pass (psx_xcontract.hip).
constexpr int CT_LAYOUT_FUSED = 16;
int get_contract_plan(psx_ctx* ctx, int prec, int step, int ncut, int nmax,
                      const ContractPlan** out);
int xc_waves_per_group();
int xc_maxpt();

template <typename T> inline const cplx<T>* plan_tw(const FftPlan* p);
template <> inline const cplx<float>* plan_tw<float>(const FftPlan* p) { return (const cplx<float>*)p->tw32; }
template <> inline const cplx<double>* plan_tw<double>(const FftPlan* p) { return (const cplx<double>*)p->tw64; }

// Opt a kernel in to more than 64 KB of dynamic LDS.  The attribute is set
// once per (device, kernel) and size increase, so a warmed-up pipeline makes
// no runtime calls besides launches (it can be captured in a HIP graph).
int lds_attribute_needed(const void* kernel, size_t bytes);
template <typename K>
inline int set_lds(K kernel, size_t bytes) {
  if (bytes > 65536 && lds_attribute_needed((const void*)kernel, bytes)) {
    PSX_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
  }
  return PSX_OK;
}

// While kernel timing is on: the start/stop event pair for the next timed
// launch, recorded under `name` (a literal); false otherwise.
bool timing_events(psx_ctx* ctx, const char* name, hipEvent_t* start, hipEvent_t* stop);

// Launch `kern` on the context's stream.  While kernel timing is on the
// launch carries its start/stop event pair: hipExtLaunchKernelGGL has the
// dispatch itself record them, so no marker packets sit between kernels.
template <typename F, typename... Args>
inline void launch_timed(psx_ctx* ctx, const char* name, F kern, dim3 grid, dim3 block,
                         size_t lds, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (timing_events(ctx, name, &a, &b))
    hipExtLaunchKernelGGL(kern, grid, block, (uint32_t)lds, ctx->stream, a, b, 0u, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, lds, ctx->stream, args...);
}

// stage launchers (defined in the .hip files)
int launch_fft3d(psx_ctx* ctx, void* grid, int N, int sign, int prec);
// line passes along axes first_axis..2 only (0 = x, 1 = y, 2 = z)
int launch_fft3d_axes(psx_ctx* ctx, void* grid, int N, int sign, int prec, int first_axis);
// the same on nbox grids back to back (box stride N^3)
int launch_fft3d_axes_batch(psx_ctx* ctx, void* grid, int N, int sign, int prec, int first_axis,
                            int nbox);
// one axis of it
int launch_fft_axis_batch(psx_ctx* ctx, void* grid, int N, int sign, int prec, int axis, int nbox);
// The last (z) pass of the backward grid FFT fused with fcomb: nbox grids
// after their x and y passes -> nbox half fields (sumw[b] per box).
int launch_fft_zcomb_batch(psx_ctx* ctx, const void* F, void* Hh, int N, int nbox,
                           const double* sumw, int prec);
// Position -> grid map of the assignment: rx = kf_ks*x + 1 + offset (float32,
// estimator.f:302-304) after clipping x to [lo, hi].
struct PcsMap {
  float kf_ks, offset, lo, hi;
};
// PCS + interlacing assignment into the complex grid (re = grid A, im = grid B);
// fft_x: also transform every x line (backward, unnormalised) before the store.
int launch_assign_grid(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int N,
                       PcsMap pm, double scale, void* grid, int prec, bool fft_x);
// nbox boxes into nbox grids back to back (box stride N^3).
int launch_assign_grid_batch(psx_ctx* ctx, int nbox, const float* const* xyz,
                             const float* const* w, const int64_t* np, const PcsMap* pm,
                             const double* scale, int N, void* grid, int prec, bool fft_x);
// fcomb of nbox grids into nbox half fields (sumw[b] per box).
int launch_fcomb_batch(psx_ctx* ctx, const void* F, void* Hh, int N, int nbox, const double* sumw,
                       int prec, const void* rec, const float* Wf, const double* Wd);
int launch_bk_finalize_batch(psx_ctx* ctx, int nbox, const double* braw, const double* p0k,
                             int p0k_stride, const double* cnt, const int* tri, int ntri, double G,
                             const double* nbar, double a3, const double* b3, double a6,
                             const double* b6, double* out);
int fixed_point_exponent(double wsum_abs);
int launch_shell_fields(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                        int prec, void* dout);
// z and y passes of the FP32 shell fields only (the x pass is fused into the
// contraction): returns the plan and the intermediate T2.
// nbox same-shape half fields back to back (box stride N*N*(N/2+1)); T2
// holds nbox intermediates (box stride sp->t_elems).
int launch_shell_zy(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                    const ShellPlan** sp_out, const void** t2_out, int nbox = 1);
// Split-BF16 contraction with the shell x pass fused (PSX_ERR_UNSUPPORTED when
// the shape is outside what the fused kernel covers; the caller then runs the
// separate stages).
int xcontract_supported(psx_ctx* ctx, int N, int step, int ncut, int nmax, int prec);
// nbox half fields back to back -> braw [nbox][ntri]
int launch_xcontract(psx_ctx* ctx, const void* half, int N, int step, int ncut, int nmax,
                     double* braw, int nbox = 1);
int launch_fill_unit(psx_ctx* ctx, double* half, int64_t n);
int launch_survey_catalogue(psx_ctx* ctx, const double* radecz, const double* nb, const double* w,
                            int64_t n, const double* dist_nodes, const double* dist_slopes, int nseg,
                            double zmax, double h, double p0, float* xyz, float* wout,
                            double* sums);
int launch_five_delta2_1(psx_ctx* ctx, void* xx, const void* yy, const void* zz, int N, int unit,
                         int fp);
int launch_five_delta2_2(psx_ctx* ctx, const void* g, void* xx, const void* xy, const void* yz,
                         const void* zx, int N, int unit, int fp);
int launch_build_quad(psx_ctx* ctx, const void* d1, void* d2, int irsd, int N, int fp);
int launch_quad_weights(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ia, int ib,
                        int ic, int id, float* out);
// Fixed-order float64 reduction of contraction partials
// [nbox][ngroups][nwg][nslot][256] and the gather of each triangle's sum
// (reference loop order) into braw [nbox][ntri].
int launch_contract_reduce(psx_ctx* ctx, const ContractPlan* cp, const double* part, int64_t nwg,
                           int ngroups, double* braw, int nbox = 1);
}  // namespace psx
