/* psx.h -- C ABI of libpsx.so, the gfx950 periodic-box P(k)/B(k) estimator.
 *
 * Each entry point replaces one stage of pySpectrum's periodic-box path; the
 * comment above it cites the reference interface it stands in for
 * (pyspectrum/pyspectrum.py and the f2py module built from
 * pyspectrum/estimator.f).
 *
 * Conventions
 *  - Every function returns an int status (PSX_OK = 0); on failure the
 *    message is available from psx_last_error() (per calling thread).  No
 *    C++ exception or abort crosses this boundary.
 *  - Array arguments are DEVICE pointers unless marked [host].  The caller
 *    owns them; the library owns only its context (plans, workspace).
 *  - Grids are stored x fastest: element (x, y, z) of a full grid at
 *    (z*N + y)*N + x; the half field (x <= N/2, the reference's
 *    ifft_delta[:N//2+1]) at (z*N + y)*(N/2+1) + x.  Complex values are
 *    interleaved (re, im) of float (prec F32/BF16X3) or double (prec F64).
 *  - Work is enqueued on the context's stream (psx_ctx_set_stream); calls
 *    return before the device finishes unless stated otherwise.  The B(k)
 *    entry points also use a second stream the context owns, forked from
 *    and joined back into the context's stream by events within the call,
 *    so every call stays ordered on the context's stream.
 */
#ifndef PSX_H
#define PSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSX_OK 0
#define PSX_ERR_ARG 1
#define PSX_ERR_HIP 2
#define PSX_ERR_NOMEM 3
#define PSX_ERR_UNSUPPORTED 4

/* Pipeline precision. */
#define PSX_PREC_F64 0    /* FP64 FFTs, shells and MFMA contraction                */
#define PSX_PREC_F32 1    /* FP32 FFTs and shells, FP32 MFMA, FP64 tile reduction  */
#define PSX_PREC_BF16X3 2 /* FP32 FFTs and shells, split-BF16 3-term MFMA, FP32
                             accumulation, FP64 tile reduction                     */

typedef struct psx_ctx psx_ctx;

/* ABI version of this header (bumped on any signature change). */
int psx_abi_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* psx_last_error(void);

/* Context: one per device; plans and workspace are cached in it, work goes to
 * the stream last set.  Entry points lock the context, and setting a new
 * stream makes it wait for the work queued on the previous one, so threads
 * may share a context (set the stream and call under one caller-side lock
 * to keep each call on the caller's stream).
 * (No reference counterpart: the reference keeps no state between calls.) */
int psx_ctx_create(psx_ctx** ctx, int device);
int psx_ctx_destroy(psx_ctx* ctx);
int psx_ctx_set_stream(psx_ctx* ctx, void* hip_stream);
int psx_ctx_synchronize(psx_ctx* ctx);
/* Bytes of device workspace currently held by the context. */
int psx_ctx_workspace_bytes(psx_ctx* ctx, int64_t* bytes);
/* Free the context's workspace after the work queued on its stream has
 * finished (plans are kept; the next call allocates what it needs again).
 * A request loop calls it on an idle context to keep the device memory of
 * all its contexts within a budget.  (No reference counterpart.) */
int psx_ctx_release_workspace(psx_ctx* ctx);
/* Kernel timing for measurement tools (no reference counterpart): while
 * enabled, each launch of the main kernels (k_assign_lines, k_fft_lines,
 * k_fft_zcomb, k_shell_z, k_shell_y, k_shell_x, k_contract, k_xcontract) is
 * launched with a start/stop HIP event pair that the dispatch itself records
 * (hipExtLaunchKernel: no marker packets between kernels).  Enabling or
 * disabling waits for the stream and clears what was recorded. */
int psx_ctx_kernel_timing(psx_ctx* ctx, int enable);
/* Number of launches of `kernel` recorded since timing was enabled and their
 * summed device time in ms (waits for the recorded events). */
int psx_ctx_kernel_time(psx_ctx* ctx, const char* kernel, int64_t* launches /* [host] */,
                        double* total_ms /* [host] */);

/* Replaces estimator.f:284-512 assign_quad (ia=ib=ic=id=0 branch), called at
 * pyspectrum.py:951 (offset 0) and :819 (offset 0.5*Ngrid).  PCS assignment
 * with interlacing of np particles at grid coordinate kf_ks*x + 1 + offset.
 * xyz: float32 [3][np] (x block, y block, z block); w: float32 [np] or NULL
 * for unit weights; positions are first clipped to [clip_lo, clip_hi]
 * (FFT_periodic clips to [0, Lbox(1-1e-6)], pyspectrum.py:939-941), which
 * must map inside the grid.  grid_out: complex [N^3], re = grid A (the
 * reference's odd slots), im = grid B (even slots).  wsum_abs = sum |w| sets
 * the internal fixed-point scale (order-independent accumulation). */
int psx_assign_quad(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                    float kf_ks, float offset, float clip_lo, float clip_hi, double wsum_abs,
                    void* grid_out, int prec);

/* Replaces the transform of _FFT (pyspectrum.py:1060-1080, pyfftw ifftn
 * with normalise_idft=False): unnormalised 3D c2c FFT in place, exponent
 * sign +1 (backward) or -1 (forward). */
int psx_fft3d(psx_ctx* ctx, void* grid, int ngrid, int sign, int prec);

/* Replaces estimator.f:605-675 fcomb_periodic (pyspectrum.py:957) followed by
 * the half-field slice of pyspectrum.py:959: reads the transformed full
 * grid, writes the half field.  sumw = sum of weights (cast to float32 as
 * f2py does). */
int psx_fcomb_periodic(psx_ctx* ctx, const void* grid, void* half_out, int ngrid, double sumw,
                       int prec);

/* Replaces FFT_periodic (pyspectrum.py:909-959): assign + FFT + fcomb, as
 * three kernels: the assignment with the x pass of the backward FFT fused,
 * the y pass, and the z pass with fcomb fused (writes the half field). */
int psx_fft_periodic(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                     double lbox, double sumw, double wsum_abs, void* half_out, int prec);

/* Replaces the per-object work of FFT_survey_mono (pyspectrum.py:767-806)
 * with radecz_to_cartesian (util.py:27-51), in float64 on the device:
 * radecz: device float64 [3][n] (RA, Dec in degrees, redshift in
 * [0, zmax]); nb: device float64 [n]; w: device float64 [n] or NULL (unit
 * weights).  The comoving distance is the cubic Hermite interpolant of a
 * host table: dist_nodes[k] = D(k zmax/nseg) [Mpc] and dist_slopes[k] =
 * (zmax/nseg) dD/dz there (k = 0..nseg).  Writes xyz_out (float32 [3][n],
 * Mpc/h = D h) and w_out (float32 [n], w/(1 + nb p0_fkp)) and, on the host,
 * sums_out[PSX_SURVEY_SUMS] = Ntot, I12, I13, I22, I23, I33, sum|w_out|,
 * max x, max y, max z, max -x, max -y, max -z.  Synchronises the stream. */
#define PSX_SURVEY_SUMS 13
int psx_survey_catalogue(psx_ctx* ctx, const double* radecz, const double* nb, const double* w,
                         int64_t n, const double* dist_nodes, const double* dist_slopes, int nseg,
                         double zmax, double h, double p0_fkp, float* xyz_out, float* w_out,
                         double* sums_out);

/* Replaces the grid part of FFT_survey_mono (pyspectrum.py:817-822):
 * assign_quad with offset 0.5*Ngrid and kf_ks = float32(Ngrid/Lbox) (no
 * clipping of positions inside (-Lbox/2, Lbox/2)), the backward FFT and
 * estimator.f:677-745 fcomb_survey (fcomb_periodic with cf = 1/(6^3*4)), then
 * the half-field slice.  w: the float32 weights including the FKP factor. */
int psx_fft_survey_mono(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                        double lbox, double wsum_abs, void* half_out, int prec);

/* Quadrupole machinery (the reference's unfinished survey quadrupole,
 * pyspectrum.py:828-900 and :1083-1131, whose Python calls cannot run;
 * these replace the Fortran routines they call).  Half fields as everywhere
 * here: complex [N][N][N/2+1] (x fastest = the Fortran (N/2+1, N, N)).
 * The float path replays the Fortran's single-precision operations in order;
 * modes with |k| = 0 are left unchanged.
 *
 * psx_quad_weights replaces the weight branch of assign_quad
 * (estimator.f:291-299): w_out = w (ia..id all 0), w r_ia r_ib / |r|^2
 * (ic = id = 0) or w r_ia r_ib r_ic r_id / |r|^4 (components 1..3); feed
 * w_out to psx_assign_quad / psx_fft_periodic for the Q_ij fields.  w may
 * be NULL (unit weights). */
int psx_quad_weights(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ia, int ib,
                     int ic, int id, float* w_out);
/* estimator.f:514-540 FiveDelta2g_1: dcgxx = 7.5 (dcgxx kxh^2 + dcgyy kyh^2
 * + dcgzz kzh^2).  unit_vectors = 0 reproduces the Fortran, whose undeclared
 * kxh/kyh/kzh are INTEGER (float(k)/rk truncated: +-1 on the axes, else 0);
 * unit_vectors = 1 uses the real unit vectors k_i/|k|. */
int psx_five_delta2g_1(psx_ctx* ctx, void* dcgxx, const void* dcgyy, const void* dcgzz, int ngrid,
                       int unit_vectors, int prec);
/* estimator.f:542-572 FiveDelta2g_2: dcgxx = dcgxx + 7.5 (2 dcgxy kxh kyh +
 * 2 dcgyz kyh kzh + 2 dcgzx kzh kxh) - 2.5 dcg; unit_vectors as above. */
int psx_five_delta2g_2(psx_ctx* ctx, const void* dcg, void* dcgxx, const void* dcgxy,
                       const void* dcgyz, const void* dcgzx, int ngrid, int unit_vectors, int prec);
/* estimator.f:574-604 build_quad: dclr2 = (7.5 mu^2 - 2.5) dclr1 with
 * mu = k_los/|k|, line of sight irsd = 1, 2, 3 (x, y, z); the Fortran STOPs
 * on any other irsd, this returns PSX_ERR_ARG. */
int psx_build_quad(psx_ctx* ctx, const void* dclr1, void* dclr2, int irsd, int ngrid, int prec);

/* Replaces reflect_delta (pyspectrum.py:1134-1157): half field -> full
 * field written C-order [x][y][z] (z fastest), as the reference returns it. */
int psx_reflect_delta(psx_ctx* ctx, const void* half, void* full_out, int ngrid, int prec);

/* Entry of _Bk_periodic (pyspectrum.py:359) for an arbitrary full field:
 * writes the half field of its Hermitian part (delta(k)+conj delta(-k))/2,
 * whose shell fields equal the reference's Re(fftn(delta * mask)).
 * full: complex of precision in_prec with element strides sx, sy, sz. */
int psx_hermitian_half(psx_ctx* ctx, const void* full, int in_prec, int64_t sx, int64_t sy,
                       int64_t sz, void* half_out, int ngrid, int prec);

/* Replaces the binning loops of Pk_periodic (pyspectrum.py:691-714) and the
 * shell power of _Bk_periodic (:404, in Parseval form): over the full field
 * reflect_delta would build, bins b = int(|k|/step + 0.5) for b = 1..nbins
 * (1 <= nbins <= 1024); out3 = [3][nbins] float64 sums of (kf*|k|,
 * |delta|^2, 1), summed in a fixed order (bit-identical run to run). */
int psx_bin_power(psx_ctx* ctx, const void* half, int ngrid, int step, int nbins, double kf,
                  double* out3, int prec);

/* psx_bin_power with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments, kernels and results (bit
 * for bit) as psx_bin_power; its block partials live in a fixed 6 MiB device
 * array that the library holds on each device until the process ends, so
 * nbins is limited to 1..512 (PSX_ERR_ARG beyond).  Calls that overlap on one
 * device share that array and race: serialise them. */
int psx_bin_power_noalloc(psx_ctx* ctx, const void* half, int ngrid, int step, int nbins,
                          double kf, double* out3, int prec);

/* Replaces estimator.f:155-264 pk_pbox_rsd (pyspectrum.py:634).  out: float64
 * sums [5][nbin] (k, p0, p2, p4, nk) then [4][nbin][nmu] (km, mk, pkm, nkm),
 * before the division by counts and kf^3 (done by the caller). */
int psx_pk_pbox_rsd(psx_ctx* ctx, const void* half, int irsd, int lbox, int nbin, int nmu,
                    int ngrid, double* out, int prec);

/* Input side of the serving campaign (no reference counterpart; the
 * reference takes particle positions from files).  Random numbers are
 * Philox4x32-10 keyed by `seed`; every output depends only on (seed, index).
 *
 * psx_mock_gaussian: Gaussian field on a mesh^3 grid (mesh: an Ngrid the FFT
 * supports) with P(k) proportional to k^slope exp(-(k/kcut)^2), k in units
 * of 2 pi / lbox: white noise, forward FFT, amplitude filter, backward FFT,
 * divided by its sample standard deviation.  g_out: float32 [mesh^3], index
 * (i0 * mesh + i1) * mesh + i2; sigma_out [host, optional]: the standard
 * deviation before the division. */
int psx_mock_gaussian(psx_ctx* ctx, int mesh, double lbox, uint64_t seed, double slope,
                      double kcut, float* g_out, double* sigma_out /* [host] or NULL */);
/* psx_mock_sample: np points drawn from the lognormal density exp(g - 1/2)
 * of a field g (float32 [mesh^3]): each cell with probability proportional
 * to its weight (prefix sums in a fixed order, first cell whose cumulative
 * weight reaches u * total), placed uniformly inside it; xyz_out: float32
 * (3, np), axis a from index i_a, in [0, lbox).  With sigma_v > 0 the
 * coordinate on axis `los` is moved as psx_apply_rsd does with the velocity
 * psx_mock_velocities(seed, sigma_v) gives (the fused redshift-space shift,
 * util.py:54-75). */
int psx_mock_sample(psx_ctx* ctx, const float* g, int mesh, double lbox, int64_t np,
                    uint64_t seed, double sigma_v, double rsd_fac, int los, float* xyz_out);
/* Gaussian velocities: float32 (3, np), sigma * N(0, 1). */
int psx_mock_velocities(psx_ctx* ctx, int64_t np, uint64_t seed, double sigma, float* v_out);
/* Replaces util.py:54-75 applyRSD for float32 (3, np) positions and
 * velocities: on axis `los`, x <- (x + (rsd_fac v + lbox)) % lbox with the
 * sum formed in float64 and rounded to float32 before the float32 wrap
 * (numpy's in-place add and remainder); other axes are copied. */
int psx_apply_rsd(psx_ctx* ctx, const float* xyz, const float* v, int64_t np, double rsd_fac,
                  double lbox, int los, float* xyz_out);

/* Triangle list in the loop order of _Bk_periodic (pyspectrum.py:415-417). */
int psx_bk_num_triangles(int step, int ncut, int nmax, int64_t* ntri /* [host] */);
int psx_bk_triangles(int step, int ncut, int nmax, int32_t* ijl /* [host] [ntri][3] */);

/* Replaces the shell loop of _Bk_periodic (pyspectrum.py:387-400):
 * d_j(x) = Re FFT_forward[delta * 1(shell == j)] for j = Ncut//step..Nmax,
 * written as real [ns][N^3] (float for F32/BF16X3, double for F64). */
int psx_bk_shell_fields(psx_ctx* ctx, const void* half, int ngrid, int step, int ncut, int nmax,
                        void* d_out, int prec);

/* Replaces the triangle loop's einsum (pyspectrum.py:427-430):
 * b_raw[t] = sum_x d_i d_j d_l for every triangle t, float64. */
int psx_bk_contract(psx_ctx* ctx, const void* d, int ngrid, int step, int ncut, int nmax,
                    double* b_raw, int prec);

/* _Bk_periodic (pyspectrum.py:359-457) up to the normalisation: shell
 * fields (workspace) + contraction; p0k[j-1] = sum_x d_j^2 / N^3 / Nk[j]
 * ([nmax] float64, NaN for an empty shell, 0 below Ncut//step). */
int psx_bk_raw(psx_ctx* ctx, const void* half, int ngrid, int step, int ncut, int nmax,
               double* b_raw, double* p0k, int prec);

/* One box of a batch: device float32 positions (3, np) and weights (np, or
 * NULL), and the per-box scalars of psx_bk_periodic. */
typedef struct psx_box {
  const float* xyz;
  const float* w;
  int64_t np;
  double lbox, sumw, wsum_abs, nbar, b3, b6;
} psx_box;

/* psx_bk_periodic for 1..16 boxes of one shape (ngrid, step, ncut, nmax,
 * prec) with one launch per stage for all of them: assignment (per-box
 * counting sorts, tiles of every box in one grid), the backward FFT, fcomb,
 * the shell passes and the contraction (split-BF16: the fused kernel; FP32
 * and FP64 run their shell x pass and contraction box by box), the shell
 * powers and the normalisation.  Outputs per box back to back: out7
 * [nbox][7][ntri], p0k [nbox][nmax]; half_out (optional, NULL = internal)
 * [nbox][ngrid][ngrid][ngrid/2+1] complex of the pipeline precision.  Each
 * box's outputs are bit-identical to psx_bk_periodic on that box.  (The
 * reference runs boxes one at a time: run/glam/glam.py:98-116.) */
int psx_bk_periodic_batch(psx_ctx* ctx, int nbox, const psx_box* boxes /* [host] */, int ngrid,
                          int step, int ncut, int nmax, const double* counts, const int32_t* ijl,
                          double a3, double a6, double* out7, double* p0k, void* half_out,
                          int prec);

/* Upper bound of the device workspace (bytes) that psx_bk_periodic on np
 * particles allocates in a context, for admission by memory (the request
 * loop); the caller's own buffers are not included.  (No reference
 * counterpart: the reference allocates host arrays per call.) */
int psx_bk_workspace_bytes(psx_ctx* ctx, int ngrid, int step, int ncut, int nmax, int64_t np,
                           int prec, int64_t* bytes);

/* Replaces _counts_Bk123 (pyspectrum.py:962-1030, cache-miss path):
 * counts[t] = sum_x I_i I_j I_l of the unit shells in FP64 (= N^3 N_tri up to
 * round-off), in triangle order. */
int psx_bk_counts(psx_ctx* ctx, int ngrid, int step, int ncut, int nmax, double* counts);

/* Replaces pyspectrum.py:418-445 and :346-355: normalisation by counts,
 * unit conversion and shot noise, float64.  ijl: device int32 [ntri][3];
 * a3 = (2 pi)^3, b3 = kf^3, a6 = (2 pi)^6, b6 = kf^6 as the caller computes
 * them.  out7: [7][ntri] = p0k1, p0k2, p0k3, b123, q123, counts, b123_sn. */
int psx_bk_finalize(psx_ctx* ctx, const double* b_raw, const double* p0k, const double* counts,
                    const int32_t* ijl, int ntri, int ngrid, double nbar, double a3, double b3,
                    double a6, double b6, double* out7);

/* Replaces Bk_periodic (pyspectrum.py:285-356) end to end: positions ->
 * out7 (see psx_bk_finalize) and p0k [nmax].  counts/ijl as above. */
int psx_bk_periodic(psx_ctx* ctx, const float* xyz, const float* w, int64_t np, int ngrid,
                    double lbox, double sumw, double wsum_abs, int step, int ncut, int nmax,
                    const double* counts, const int32_t* ijl, double nbar, double a3, double b3,
                    double a6, double b6, double* out7, double* p0k, int prec);

#ifdef __cplusplus
}
#endif
#endif /* PSX_H */
