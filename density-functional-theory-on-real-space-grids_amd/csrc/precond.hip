// The Teter-Payne-Allan kinetic preconditioner of the LOBPCG cross-check (reference
// src/cdft/eigen/lobpcg.py:207-239, frames at lobpcg.py:28-41 and 129-133).
//
// On the real-FFT half spectrum of the box, K(x) = P(x) / (P(x) + 16 x^4) with P = 27 + 18x + 12x^2
// + 8x^3 and x = (G^2/2) / E_ref, E_ref the mean of G^2/2 over the half spectrum (at least 1e-6). A
// block of residual rows is filtered one row at a time: scale into the wavefunction frame (r f, with
// f = max(exp(-u), 1e-150), when the cusp factor u is given), a 3-D D2Z transform, the product with
// the stored filter (already divided by the transform size), the 3-D Z2D transform, and the scale
// back (/ f). The scaling is fused into the kernels that read and write the real rows. The
// transforms are rocFFT (through hipFFT); the caller owns every buffer.
#include "common.h"

#include <hipfft/hipfft.h>

namespace cdft {

struct TpaPlan {
  hipfftHandle forward;  // 3-D D2Z over (n0, n1, n2)
  hipfftHandle inverse;  // 3-D Z2D
  int n0, n1, n2;
};

__device__ __forceinline__ double freq(int i, int n, double h) {
  // numpy's fftfreq: i / (n h) for i < (n + 1) / 2, else (i - n) / (n h)
  return (i < (n + 1) / 2 ? (double)i : (double)(i - n)) / ((double)n * h);
}

__global__ void tpa_filter_kernel(int n0, int n1, int n2, double h, double e_ref, double norm,
                                  double* __restrict__ filter) {
  const int q2 = n2 / 2 + 1;
  const long total = (long)n0 * n1 * q2;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int k2 = (int)(e % q2);
  const int k1 = (int)((e / q2) % n1);
  const int k0 = (int)(e / ((long)q2 * n1));
  const double two_pi = 2.0 * M_PI;
  const double kx = two_pi * freq(k0, n0, h), ky = two_pi * freq(k1, n1, h);
  const double kz = two_pi * ((double)k2 / ((double)n2 * h));
  const double x = 0.5 * ((kx * kx + ky * ky) + kz * kz) / e_ref;
  const double x2 = x * x;
  const double poly = ((27.0 + 18.0 * x) + 12.0 * x2) + 8.0 * (x2 * x);
  filter[e] = poly / (poly + 16.0 * (x2 * x2)) * norm;
}

// out = in * f (u non-null) or in
__global__ void tpa_scale_in_kernel(const double* __restrict__ in, const double* __restrict__ u, long n,
                                    double* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double v = in[p];
  out[p] = u ? v * fmax(exp(-u[p]), 1.0e-150) : v;
}

__global__ void tpa_spectrum_kernel(double2* __restrict__ spec, const double* __restrict__ filter, long total) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const double k = filter[e];
  double2 v = spec[e];
  v.x *= k;
  v.y *= k;
  spec[e] = v;
}

// out = in / f (u non-null) or in
__global__ void tpa_scale_out_kernel(const double* __restrict__ in, const double* __restrict__ u, long n,
                                     double* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double v = in[p];
  out[p] = u ? v / fmax(exp(-u[p]), 1.0e-150) : v;
}

static unsigned tb(long n) { return (unsigned)((n + 255) / 256); }

// numpy's fftfreq / rfftfreq squared, summed over one axis
static double mean_square_freq(int n, double h, bool half) {
  const int count = half ? n / 2 + 1 : n;
  double s = 0.0;
  for (int i = 0; i < count; ++i) {
    const double f = (half || i < (n + 1) / 2 ? (double)i : (double)(i - n)) / ((double)n * h);
    s += f * f;
  }
  return s / count;
}

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_tpa_create(int n0, int n1, int n2, void** handle) {
  if (!handle || n0 < 1 || n1 < 1 || n2 < 1) return fail(CDFT_E_ARG, "tpa_create: bad sizes");
  TpaPlan* plan = new TpaPlan();
  plan->n0 = n0;
  plan->n1 = n1;
  plan->n2 = n2;
  if (hipfftPlan3d(&plan->forward, n0, n1, n2, HIPFFT_D2Z) != HIPFFT_SUCCESS) {
    delete plan;
    return fail(CDFT_E_LIBRARY, "hipfftPlan3d (D2Z, %d x %d x %d) failed", n0, n1, n2);
  }
  if (hipfftPlan3d(&plan->inverse, n0, n1, n2, HIPFFT_Z2D) != HIPFFT_SUCCESS) {
    hipfftDestroy(plan->forward);
    delete plan;
    return fail(CDFT_E_LIBRARY, "hipfftPlan3d (Z2D, %d x %d x %d) failed", n0, n1, n2);
  }
  *handle = plan;
  return CDFT_OK;
}

extern "C" int cdft_tpa_destroy(void* handle) {
  if (!handle) return CDFT_OK;
  TpaPlan* plan = (TpaPlan*)handle;
  hipfftDestroy(plan->forward);
  hipfftDestroy(plan->inverse);
  delete plan;
  return CDFT_OK;
}

// filter: device array n0 x n1 x (n2/2+1) doubles, written with K / (n0 n1 n2). e_ref (host, may be
// NULL) receives E_ref.
extern "C" int cdft_tpa_filter_f64(void* handle, double h, double* filter, double* e_ref, void* stream) {
  if (!handle || !filter || !(h > 0.0)) return fail(CDFT_E_ARG, "tpa_filter: bad argument");
  TpaPlan* pl = (TpaPlan*)handle;
  // the mean of (kx^2 + ky^2 + kz^2) / 2 over the tensor-product half spectrum is the sum of the
  // per-axis means of the squared frequencies times (2 pi)^2 / 2
  const double two_pi = 2.0 * M_PI;
  double reference = 0.5 * two_pi * two_pi *
                     ((mean_square_freq(pl->n0, h, false) + mean_square_freq(pl->n1, h, false)) +
                      mean_square_freq(pl->n2, h, true));
  if (reference < 1.0e-6) reference = 1.0e-6;
  if (e_ref) *e_ref = reference;
  const long total = (long)pl->n0 * pl->n1 * (pl->n2 / 2 + 1);
  const double norm = 1.0 / ((double)pl->n0 * pl->n1 * pl->n2);
  hipLaunchKernelGGL(tpa_filter_kernel, dim3(tb(total)), dim3(256), 0, (hipStream_t)stream, pl->n0, pl->n1, pl->n2, h,
                     reference, norm, filter);
  return check_launch("tpa_filter");
}

// out[i] = f^-1 K (f in[i]) for the k rows of `in` (u = NULL: K in[i]). real_work: n0 n1 n2 doubles;
// spectrum: n0 n1 (n2/2+1) complex doubles. `out` may alias `in`.
extern "C" int cdft_tpa_apply_f64(void* handle, const double* in, int k, const double* u, const double* filter,
                                  double* real_work, void* spectrum, double* out, void* stream) {
  if (!handle || !in || !filter || !real_work || !spectrum || !out || k < 0)
    return fail(CDFT_E_ARG, "tpa_apply: bad argument");
  TpaPlan* pl = (TpaPlan*)handle;
  hipStream_t s = (hipStream_t)stream;
  if (hipfftSetStream(pl->forward, s) != HIPFFT_SUCCESS || hipfftSetStream(pl->inverse, s) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftSetStream failed");
  const long n = (long)pl->n0 * pl->n1 * pl->n2;
  const long total = (long)pl->n0 * pl->n1 * (pl->n2 / 2 + 1);
  hipfftDoubleComplex* sp = (hipfftDoubleComplex*)spectrum;
  for (int i = 0; i < k; ++i) {
    hipLaunchKernelGGL(tpa_scale_in_kernel, dim3(tb(n)), dim3(256), 0, s, in + (long)i * n, u, n, real_work);
    int rc = check_launch("tpa scale_in");
    if (rc) return rc;
    if (hipfftExecD2Z(pl->forward, real_work, sp) != HIPFFT_SUCCESS) return fail(CDFT_E_LIBRARY, "hipfftExecD2Z failed");
    hipLaunchKernelGGL(tpa_spectrum_kernel, dim3(tb(total)), dim3(256), 0, s, (double2*)spectrum, filter, total);
    rc = check_launch("tpa spectrum");
    if (rc) return rc;
    if (hipfftExecZ2D(pl->inverse, sp, real_work) != HIPFFT_SUCCESS) return fail(CDFT_E_LIBRARY, "hipfftExecZ2D failed");
    hipLaunchKernelGGL(tpa_scale_out_kernel, dim3(tb(n)), dim3(256), 0, s, real_work, u, n, out + (long)i * n);
    rc = check_launch("tpa scale_out");
    if (rc) return rc;
  }
  return CDFT_OK;
}
