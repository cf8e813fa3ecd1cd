// K5: the Coulomb-cutoff free-space Poisson solve on a zero-padded box (reference
// src/cdft/operators/poisson.py:27-84): rocFFT (through hipFFT) D2Z / Z2D plans, plus hand-written
// kernels for the Green's function K(G) = 4 pi (1 - cos G R)/G^2, K(0) = 2 pi R^2 (stored divided by
// the transform size, so the inverse needs no separate normalisation pass), the pad-scatter of the
// density into the padded box and the gather of the interior potential.
//
// Buffers are owned by the caller: `padded_in` (P real, zero outside the box, never written by the
// transforms), `spectrum` (P0 x P1 x (P2/2+1) complex), `padded_out` (P real). The padded input
// region outside the box is zeroed once at setup; every solve rewrites only the interior.
#include "common.h"

#include <hipfft/hipfft.h>

namespace cdft {

struct PoissonPlan {
  hipfftHandle forward;
  hipfftHandle inverse;
  int n0, n1, n2;  // box
  int p0, p1, p2;  // padded
};

__global__ void green_kernel(int p0, int p1, int p2, double h, double radius, double norm,
                             double* __restrict__ kernel) {
  const int q2 = p2 / 2 + 1;
  const long total = (long)p0 * p1 * q2;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int k2 = (int)(e % q2);
  const int k1 = (int)((e / q2) % p1);
  const int k0 = (int)(e / ((long)q2 * p1));
  // fftfreq: index i of n maps to i/(n h) for i < (n+1)/2, else (i - n)/(n h); rfftfreq keeps i.
  const double two_pi = 2.0 * M_PI;
  const double f0 = (k0 < (p0 + 1) / 2 ? (double)k0 : (double)(k0 - p0)) / ((double)p0 * h);
  const double f1 = (k1 < (p1 + 1) / 2 ? (double)k1 : (double)(k1 - p1)) / ((double)p1 * h);
  const double f2 = (double)k2 / ((double)p2 * h);
  const double kx = two_pi * f0, ky = two_pi * f1, kz = two_pi * f2;
  const double g2 = (kx * kx + ky * ky) + kz * kz;
  double k;
  if (g2 > 0.0) {
    const double g = sqrt(g2);
    k = 4.0 * M_PI * (1.0 - cos(g * radius)) / g2;
  } else {
    k = 2.0 * M_PI * radius * radius;
  }
  kernel[e] = k * norm;
}

__global__ void pad_scatter_kernel(const double* __restrict__ box, int n0, int n1, int n2, int p1, int p2,
                                   double* __restrict__ padded) {
  const long n = (long)n0 * n1 * n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % n2);
  const int i1 = (int)((p / n2) % n1);
  const int i0 = (int)(p / ((long)n1 * n2));
  padded[((long)i0 * p1 + i1) * p2 + i2] = box[p];
}

__global__ void spectrum_multiply_kernel(double2* __restrict__ spec, const double* __restrict__ kernel, long total) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const double k = kernel[e];
  double2 v = spec[e];
  v.x *= k;
  v.y *= k;
  spec[e] = v;
}

__global__ void gather_kernel(const double* __restrict__ padded, int n0, int n1, int n2, int p1, int p2,
                              double* __restrict__ out) {
  const long n = (long)n0 * n1 * n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % n2);
  const int i1 = (int)((p / n2) % n1);
  const int i0 = (int)(p / ((long)n1 * n2));
  out[p] = padded[((long)i0 * p1 + i1) * p2 + i2];
}

static unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_poisson_create(int n0, int n1, int n2, int p0, int p1, int p2, void** handle) {
  if (!handle || n0 < 1 || n1 < 1 || n2 < 1 || p0 < n0 || p1 < n1 || p2 < n2)
    return fail(CDFT_E_ARG, "poisson_create: bad sizes");
  PoissonPlan* plan = new PoissonPlan();
  plan->n0 = n0; plan->n1 = n1; plan->n2 = n2;
  plan->p0 = p0; plan->p1 = p1; plan->p2 = p2;
  if (hipfftPlan3d(&plan->forward, p0, p1, p2, HIPFFT_D2Z) != HIPFFT_SUCCESS) {
    delete plan;
    return fail(CDFT_E_LIBRARY, "hipfftPlan3d D2Z (%d, %d, %d) failed", p0, p1, p2);
  }
  if (hipfftPlan3d(&plan->inverse, p0, p1, p2, HIPFFT_Z2D) != HIPFFT_SUCCESS) {
    hipfftDestroy(plan->forward);
    delete plan;
    return fail(CDFT_E_LIBRARY, "hipfftPlan3d Z2D (%d, %d, %d) failed", p0, p1, p2);
  }
  *handle = plan;
  return CDFT_OK;
}

extern "C" int cdft_poisson_destroy(void* handle) {
  if (!handle) return CDFT_OK;
  PoissonPlan* plan = (PoissonPlan*)handle;
  hipfftDestroy(plan->forward);
  hipfftDestroy(plan->inverse);
  delete plan;
  return CDFT_OK;
}

// kernel: device array P0 x P1 x (P2/2+1) doubles, written with K(G) / (P0 P1 P2).
extern "C" int cdft_poisson_kernel_f64(void* handle, double h, double radius, double* kernel, void* stream) {
  if (!handle || !kernel) return fail(CDFT_E_ARG, "poisson_kernel: null argument");
  PoissonPlan* pl = (PoissonPlan*)handle;
  const long total = (long)pl->p0 * pl->p1 * (pl->p2 / 2 + 1);
  const double norm = 1.0 / ((double)pl->p0 * pl->p1 * pl->p2);
  hipLaunchKernelGGL(green_kernel, dim3(nb(total)), dim3(256), 0, (hipStream_t)stream, pl->p0, pl->p1, pl->p2, h,
                     radius, norm, kernel);
  return check_launch("poisson_kernel");
}

extern "C" int cdft_poisson_solve_f64(void* handle, const double* density, double* padded_in,
                                      void* spectrum, double* padded_out, const double* kernel,
                                      double* potential, void* stream) {
  if (!handle || !density || !padded_in || !spectrum || !padded_out || !kernel || !potential)
    return fail(CDFT_E_ARG, "poisson_solve: null argument");
  PoissonPlan* pl = (PoissonPlan*)handle;
  hipStream_t s = (hipStream_t)stream;
  const long n = (long)pl->n0 * pl->n1 * pl->n2;
  const long total = (long)pl->p0 * pl->p1 * (pl->p2 / 2 + 1);
  hipLaunchKernelGGL(pad_scatter_kernel, dim3(nb(n)), dim3(256), 0, s, density, pl->n0, pl->n1, pl->n2, pl->p1,
                     pl->p2, padded_in);
  int rc = check_launch("poisson pad_scatter");
  if (rc) return rc;
  if (hipfftSetStream(pl->forward, s) != HIPFFT_SUCCESS || hipfftSetStream(pl->inverse, s) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftSetStream failed");
  if (hipfftExecD2Z(pl->forward, padded_in, (hipfftDoubleComplex*)spectrum) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecD2Z failed");
  hipLaunchKernelGGL(spectrum_multiply_kernel, dim3(nb(total)), dim3(256), 0, s, (double2*)spectrum, kernel, total);
  rc = check_launch("poisson spectrum_multiply");
  if (rc) return rc;
  if (hipfftExecZ2D(pl->inverse, (hipfftDoubleComplex*)spectrum, padded_out) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecZ2D failed");
  hipLaunchKernelGGL(gather_kernel, dim3(nb(n)), dim3(256), 0, s, padded_out, pl->n0, pl->n1, pl->n2, pl->p1, pl->p2,
                     potential);
  return check_launch("poisson gather");
}
