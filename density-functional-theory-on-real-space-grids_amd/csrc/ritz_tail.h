// Device bodies of the small kernels that end a Rayleigh-Ritz pass, shared by their own launches
// (blas.hip) and by the fused residual-norm tails (hdiv.hip, local.hip), so all give bitwise the same values.
#pragma once

#include <hip/hip_runtime.h>

namespace cdft {

// The scaled Chebyshev rows of a CheFSI filter (reference src/cdft/eigen/chefsi.py:116-136):
// lower = evals[index] clamped into (smin + 1e-8, smax - 1e-8), upper = smax, smin = bounds[0],
// smax = bounds[1]; table row s = (c, a_s, b_s). The operations are those of the host routine in
// the same order, uncontracted, so the table is bitwise the host's. One thread.
__device__ __forceinline__ void cheb_table_rows(const double* __restrict__ evals, int index,
                                                const double* __restrict__ bounds, int degree,
                                                double* __restrict__ table) {
#pragma clang fp contract(off)
  const double smin = bounds[0], upper = bounds[1];
  double lower = evals[index];
  const double lo = smin + 1e-8, hi = upper - 1e-8;
  lower = lower >= lo ? lower : lo;
  lower = lower <= hi ? lower : hi;
  const double e = 0.5 * (upper - lower);
  const double c = 0.5 * (upper + lower);
  double sigma = e / (smin - c);
  const double sigma1 = sigma;
  table[0] = c;
  table[1] = sigma1 / e;
  table[2] = 0.0;
  for (int s = 1; s < degree; ++s) {
    const double sigma2 = 1.0 / (2.0 / sigma1 - sigma);
    table[3 * s] = c;
    table[3 * s + 1] = 2.0 * sigma2 / e;
    table[3 * s + 2] = sigma * sigma2;
    sigma = sigma2;
  }
}

// max |M - I| for a k x k matrix (orthonormality error), written to out[0]. One thread.
__device__ __forceinline__ void max_abs_minus_identity(const double* __restrict__ m, int k, double* __restrict__ out) {
  double worst = 0.0;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      const double d = fabs(m[i * k + j] - (i == j ? 1.0 : 0.0));
      worst = d > worst ? d : worst;
    }
  out[0] = worst;
}

}  // namespace cdft
