// K5: the Coulomb-cutoff free-space Poisson solve on a zero-padded box (reference
// src/cdft/operators/poisson.py:27-84): rocFFT (through hipFFT) plans, plus hand-written kernels for
// the Green's function K(G) = 4 pi (1 - cos G R)/G^2, K(0) = 2 pi R^2 (stored divided by the
// transform size, so the inverse needs no separate normalisation pass), the pad-scatter of the
// density and the gather of the interior potential.
//
// The transforms are pruned to the box. Only the first n0 of the p0 x-slabs of the padded density are
// non-zero, and only the first n0 x-slabs of the potential are kept, so the (y, z) transforms run on
// n0 slabs and only the x transform covers the whole padded spectrum:
//   forward: 2-D D2Z over (y, z) of the n0 slabs -> `slabs` (p0 x p1 x (p2/2+1) complex, slabs >= n0
//            permanently zero), then a strided 1-D Z2Z along x of every (ky, kz) column -> `spectrum`;
//   multiply by the kernel;
//   inverse: the 1-D Z2Z along x in place on `spectrum`, then a 2-D Z2D over (y, z) of its first n0
//            slabs -> `padded_out` (n0 x p1 x p2 real), of which the (n1, n2) corner is the potential.
// The zero rows y >= n1 and columns z >= n2 inside those slabs stay in the (y, z) transforms. The
// result is the full 3-D transform's up to rounding. Buffers are owned by the caller: `padded_in` (n0
// p1 p2 real, zero outside the box, never written by the transforms), `slabs`, `spectrum`, `padded_out`.
#include "common.h"

#include <hipfft/hipfft.h>

namespace cdft {

struct PoissonPlan {
  hipfftHandle yz_forward;  // 2-D D2Z over (p1, p2), n0 slabs
  hipfftHandle yz_inverse;  // 2-D Z2D over (p1, p2), n0 slabs
  hipfftHandle x_forward;   // 1-D Z2Z along x, stride p1 (p2/2+1), out of place
  hipfftHandle x_inverse;   // 1-D Z2Z along x, in place
  int n0, n1, n2;           // box
  int p0, p1, p2;           // padded
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
  const int q2 = p2 / 2 + 1;
  int yz[2] = {p1, p2};
  int yz_real[2] = {p1, p2};
  int yz_complex[2] = {p1, q2};
  int x[1] = {p0};
  int x_embed[1] = {p0};
  const int column = p1 * q2;
  int made = 0;
  hipfftHandle* handles[4] = {&plan->yz_forward, &plan->yz_inverse, &plan->x_forward, &plan->x_inverse};
  if (hipfftPlanMany(&plan->yz_forward, 2, yz, yz_real, 1, p1 * p2, yz_complex, 1, column, HIPFFT_D2Z, n0) == HIPFFT_SUCCESS) {
    ++made;
    if (hipfftPlanMany(&plan->yz_inverse, 2, yz, yz_complex, 1, column, yz_real, 1, p1 * p2, HIPFFT_Z2D, n0) ==
        HIPFFT_SUCCESS) {
      ++made;
      if (hipfftPlanMany(&plan->x_forward, 1, x, x_embed, column, 1, x_embed, column, 1, HIPFFT_Z2Z, column) ==
          HIPFFT_SUCCESS) {
        ++made;
        if (hipfftPlanMany(&plan->x_inverse, 1, x, x_embed, column, 1, x_embed, column, 1, HIPFFT_Z2Z, column) ==
            HIPFFT_SUCCESS)
          ++made;
      }
    }
  }
  if (made < 4) {
    for (int i = 0; i < made; ++i) hipfftDestroy(*handles[i]);
    delete plan;
    return fail(CDFT_E_LIBRARY, "hipfftPlanMany for the (%d, %d, %d) padded box failed (plan %d)", p0, p1, p2, made);
  }
  *handle = plan;
  return CDFT_OK;
}

extern "C" int cdft_poisson_destroy(void* handle) {
  if (!handle) return CDFT_OK;
  PoissonPlan* plan = (PoissonPlan*)handle;
  hipfftDestroy(plan->yz_forward);
  hipfftDestroy(plan->yz_inverse);
  hipfftDestroy(plan->x_forward);
  hipfftDestroy(plan->x_inverse);
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

extern "C" int cdft_poisson_solve_f64(void* handle, const double* density, double* padded_in, void* slabs,
                                      void* spectrum, double* padded_out, const double* kernel, double* potential,
                                      void* stream) {
  if (!handle || !density || !padded_in || !slabs || !spectrum || !padded_out || !kernel || !potential)
    return fail(CDFT_E_ARG, "poisson_solve: null argument");
  PoissonPlan* pl = (PoissonPlan*)handle;
  hipStream_t s = (hipStream_t)stream;
  const long n = (long)pl->n0 * pl->n1 * pl->n2;
  const long total = (long)pl->p0 * pl->p1 * (pl->p2 / 2 + 1);
  hipLaunchKernelGGL(pad_scatter_kernel, dim3(nb(n)), dim3(256), 0, s, density, pl->n0, pl->n1, pl->n2, pl->p1,
                     pl->p2, padded_in);
  int rc = check_launch("poisson pad_scatter");
  if (rc) return rc;
  for (hipfftHandle h : {pl->yz_forward, pl->yz_inverse, pl->x_forward, pl->x_inverse})
    if (hipfftSetStream(h, s) != HIPFFT_SUCCESS) return fail(CDFT_E_LIBRARY, "hipfftSetStream failed");
  hipfftDoubleComplex* sl = (hipfftDoubleComplex*)slabs;
  hipfftDoubleComplex* sp = (hipfftDoubleComplex*)spectrum;
  if (hipfftExecD2Z(pl->yz_forward, padded_in, sl) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecD2Z (y, z slabs) failed");
  if (hipfftExecZ2Z(pl->x_forward, sl, sp, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecZ2Z (x, forward) failed");
  hipLaunchKernelGGL(spectrum_multiply_kernel, dim3(nb(total)), dim3(256), 0, s, (double2*)spectrum, kernel, total);
  rc = check_launch("poisson spectrum_multiply");
  if (rc) return rc;
  if (hipfftExecZ2Z(pl->x_inverse, sp, sp, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecZ2Z (x, inverse) failed");
  if (hipfftExecZ2D(pl->yz_inverse, sp, padded_out) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecZ2D (y, z slabs) failed");
  hipLaunchKernelGGL(gather_kernel, dim3(nb(n)), dim3(256), 0, s, padded_out, pl->n0, pl->n1, pl->n2, pl->p1, pl->p2,
                     potential);
  return check_launch("poisson gather");
}
