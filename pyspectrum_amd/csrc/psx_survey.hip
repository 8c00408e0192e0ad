// Survey catalogue preparation on the device: the per-object work of
// FFT_survey_mono (pyspectrum.py:767-806) and radecz_to_cartesian
// (util.py:27-51) -- comoving positions, FKP weights and the normalisation
// sums -- in float64, positions and weights stored as float32 for the
// assignment.
#include "psx_ctx.h"

namespace psx {

constexpr int CAT_THREADS = 256;
constexpr int CAT_SUMS = PSX_SURVEY_SUMS;  // see include/psx.h
constexpr int CAT_MAX_BLOCKS = 1024;

// One object: rad = h * D(z) by cubic Hermite interpolation of the node
// table (Dn at nseg+1 uniform nodes on [0, zmax], sl = node spacing x dD/dz),
// (RA, Dec) in degrees, x = rad cos(dec) cos(ra), y = rad cos(dec) sin(ra),
// z = rad sin(dec); w_fkp = w / (1 + nb P0).  Sums: Ntot = sum w, I12 =
// sum w_fkp^2, I13 = sum w_fkp^3, I22 = sum nb w_fkp^2, I23 = sum nb
// w_fkp^3, I33 = sum nb^2 w_fkp^3, sum |float32(w_fkp)|, then max x, y, z
// and max -x, -y, -z (the box check).  Per-block partial sums; the host adds
// them in block order.
__global__ __launch_bounds__(CAT_THREADS) void k_survey_catalogue(
    const double* __restrict__ radecz, const double* __restrict__ nb, const double* __restrict__ w,
    int64_t n, const double* __restrict__ Dn, const double* __restrict__ sl, int nseg, double zscale,
    double h, double p0, float* __restrict__ xyz, float* __restrict__ wout,
    double* __restrict__ part) {
  __shared__ double red[CAT_SUMS][CAT_THREADS / 64];
  double acc[CAT_SUMS];
#pragma unroll
  for (int c = 0; c < CAT_SUMS; ++c) acc[c] = c < 7 ? 0.0 : -INFINITY;
  const double deg = 3.141592653589793 / 180.;
  for (int64_t i = (int64_t)blockIdx.x * CAT_THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * CAT_THREADS) {
    const double ra = radecz[i] * deg, dec = radecz[n + i] * deg, zr = radecz[2 * n + i];
    const double s = zr * zscale;  // zr in [0, zmax] (checked by the caller)
    int k = (int)s;
    k = k < 0 ? 0 : (k > nseg - 1 ? nseg - 1 : k);
    const double t = s - (double)k, t2 = t * t, t3 = t2 * t;
    const double D = (2 * t3 - 3 * t2 + 1) * Dn[k] + (t3 - 2 * t2 + t) * sl[k] +
                     (-2 * t3 + 3 * t2) * Dn[k + 1] + (t3 - t2) * sl[k + 1];
    const double rad = D * h;
    const double cd = cos(dec);
    const double x = rad * cd * cos(ra), y = rad * cd * sin(ra), z = rad * sin(dec);
    xyz[i] = (float)x;
    xyz[n + i] = (float)y;
    xyz[2 * n + i] = (float)z;
    const double wi = w ? w[i] : 1.0, nbi = nb[i];
    const double wf = wi * (1. / (1. + nbi * p0));
    const float wf32 = (float)wf;
    wout[i] = wf32;
    const double w2 = wf * wf, w3 = w2 * wf;
    acc[0] += wi;
    acc[1] += w2;
    acc[2] += w3;
    acc[3] += nbi * w2;
    acc[4] += nbi * w3;
    acc[5] += (nbi * nbi) * w3;
    acc[6] += fabs((double)wf32);
    acc[7] = fmax(acc[7], x);
    acc[8] = fmax(acc[8], y);
    acc[9] = fmax(acc[9], z);
    acc[10] = fmax(acc[10], -x);
    acc[11] = fmax(acc[11], -y);
    acc[12] = fmax(acc[12], -z);
  }
  // wave reduction (fixed butterfly order), then the waves in order
#pragma unroll
  for (int c = 0; c < CAT_SUMS; ++c) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = c < 7 ? v + o : fmax(v, o);
    }
    if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < CAT_SUMS) {
    const int c = threadIdx.x;
    double v = red[c][0];
    for (int q = 1; q < CAT_THREADS / 64; ++q) v = c < 7 ? v + red[c][q] : fmax(v, red[c][q]);
    part[(int64_t)blockIdx.x * CAT_SUMS + c] = v;
  }
}

int launch_survey_catalogue(psx_ctx* ctx, const double* radecz, const double* nb, const double* w,
                            int64_t n, const double* dist_nodes, const double* dist_slopes, int nseg,
                            double zmax, double h, double p0, float* xyz, float* wout,
                            double* sums) {
  for (int c = 0; c < CAT_SUMS; ++c) sums[c] = c < 7 ? 0.0 : -INFINITY;
  if (n == 0) return PSX_OK;
  void* tab = nullptr;
  PSX_TRY(ctx_buffer(ctx, "catalogue_table", (size_t)2 * (nseg + 1) * sizeof(double), &tab));
  double* Dn = (double*)tab;
  double* sl = Dn + (nseg + 1);
  PSX_HIP(hipMemcpyAsync(Dn, dist_nodes, (size_t)(nseg + 1) * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  PSX_HIP(hipMemcpyAsync(sl, dist_slopes, (size_t)(nseg + 1) * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  const int nb_blocks = (int)std::min<int64_t>(cdiv(n, CAT_THREADS), CAT_MAX_BLOCKS);
  void* part = nullptr;
  PSX_TRY(ctx_buffer(ctx, "catalogue_partials", (size_t)nb_blocks * CAT_SUMS * sizeof(double), &part));
  const double zscale = zmax > 0 ? (double)nseg / zmax : 0.0;
  hipLaunchKernelGGL(k_survey_catalogue, dim3(nb_blocks), dim3(CAT_THREADS), 0, ctx->stream, radecz,
                     nb, w, n, Dn, sl, nseg, zscale, h, p0, xyz, wout, (double*)part);
  PSX_CHECK_LAUNCH();
  std::vector<double> hp((size_t)nb_blocks * CAT_SUMS);
  PSX_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  PSX_HIP(hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nb_blocks; ++b)
    for (int c = 0; c < CAT_SUMS; ++c) {
      const double v = hp[(size_t)b * CAT_SUMS + c];
      sums[c] = c < 7 ? sums[c] + v : std::max(sums[c], v);
    }
  return PSX_OK;
}

}  // namespace psx
