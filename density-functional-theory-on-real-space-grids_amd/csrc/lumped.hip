// Lumped far mass weights of an overlapping geometry (reference src/cdft/operators/quadrature.py:
// 421-497): the midpoint rule of far * f^2 on a lattice refined `refine` times per axis, moved to
// the grid by three one-dimensional transposed degree-7 Lagrange interpolations, then blended back
// to the plain weights h^3 far f^2 within a few spacings of the box faces.
//
// The reference contracts z first, one fine plane at a time in increasing order, then x and y with
// one einsum; here every output element of a pass sums its fine points in increasing order, which
// is the reference's z order (the x / y order of its einsum is a library detail).
#include "common.h"

namespace cdft {

constexpr int MAXSP = 8;

struct FarSpheres {
  int n;
  double cx[MAXSP], cy[MAXSP], cz[MAXSP], radius[MAXSP], centre[MAXSP], width[MAXSP], z[MAXSP];
};

// value = (volume * far) * exp(-2 u) at fine point (fx[i], fy[j], fz[l]): far = prod_a (1 - bump_a), the bump
// 0.5 erfc((s - centre) / width) inside the sphere and 0 outside, u = sum_a Z_a s_a.
__global__ void far_fine_kernel(FarSpheres sp, int f0, int f1, int f2, const double* __restrict__ fx,
                                const double* __restrict__ fy, const double* __restrict__ fz, double volume,
                                double* __restrict__ out) {
  const long n = (long)f0 * f1 * f2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int l = (int)(p % f2);
  const int j = (int)((p / f2) % f1);
  const int i = (int)(p / ((long)f1 * f2));
  const double x = fx[i], y = fy[j], z = fz[l];
  double far = 1.0, u = 0.0;
  for (int a = 0; a < sp.n; ++a) {
    const double dx = x - sp.cx[a], dy = y - sp.cy[a], dz = z - sp.cz[a];
    // separately rounded, as the reference's torch expression (s < radius is a threshold test)
    const double s = sqrt(rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz)));
    const double bump = s < sp.radius[a] ? 0.5 * erfc((s - sp.centre[a]) / sp.width[a]) : 0.0;
    far = far * (1.0 - bump);
    u = u + sp.z[a] * s;
  }
  out[p] = (volume * far) * exp(-2.0 * u);
}

// out[.., c, ..] = sum_{f = lo[c]}^{hi[c] - 1} w[f][c - first[f]] in[.., f, ..] along `axis` of a
// (d0, d1, d2) array; out has the axis replaced by n_coarse.
__global__ void interp_transpose_kernel(const double* __restrict__ in, int d0, int d1, int d2, int axis,
                                        int nc, const int* __restrict__ range, const int* __restrict__ first,
                                        const double* __restrict__ w, double* __restrict__ out) {
  const int e0 = axis == 0 ? nc : d0, e1 = axis == 1 ? nc : d1, e2 = axis == 2 ? nc : d2;
  const long n = (long)e0 * e1 * e2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % e2);
  const int i1 = (int)((p / e2) % e1);
  const int i0 = (int)(p / ((long)e1 * e2));
  const int c = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
  const long stride = axis == 0 ? (long)d1 * d2 : (axis == 1 ? (long)d2 : 1L);
  const long base = axis == 0 ? ((long)i1 * d2 + i2) : (axis == 1 ? ((long)i0 * d1 * d2 + i2) : ((long)i0 * d1 + i1) * d2);
  double acc = 0.0;
  for (int f = range[2 * c]; f < range[2 * c + 1]; ++f) acc = acc + in[base + (long)f * stride] * w[8 * f + (c - first[f])];
  out[p] = acc;
}

// out = plain + interior (lumped - plain), interior = 0.5 erfc((margin - face) / width), face the
// distance to the nearest box face.
__global__ void far_blend_kernel(int n0, int n1, int n2, double h, double ox, double oy, double oz,
                                 double margin, double width, const double* __restrict__ plain,
                                 const double* __restrict__ lumped, double* __restrict__ out) {
  const long n = (long)n0 * n1 * n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % n2);
  const int i1 = (int)((p / n2) % n1);
  const int i0 = (int)(p / ((long)n1 * n2));
  // Distances to the lower and upper faces, formed from the point coordinates x = origin + i h as
  // (x - origin) and (origin + (n - 1) h - x).
  const double x = rn_add(ox, rn_mul((double)i0, h)), y = rn_add(oy, rn_mul((double)i1, h)),
               z = rn_add(oz, rn_mul((double)i2, h));
  const double ux = rn_add(ox, rn_mul((double)(n0 - 1), h)), uy = rn_add(oy, rn_mul((double)(n1 - 1), h)),
               uz = rn_add(oz, rn_mul((double)(n2 - 1), h));
  const double lo = fmin(fmin(x - ox, y - oy), z - oz);
  const double hi = fmin(fmin(ux - x, uy - y), uz - z);
  const double face = fmin(lo, hi);
  const double interior = 0.5 * erfc((margin - face) / width);
  out[p] = plain[p] + interior * (lumped[p] - plain[p]);
}

static unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_far_fine_f64(int n_spheres, const double* charges, const double* spheres, const int* fine_shape,
                                 const double* fx, const double* fy, const double* fz, double volume, double* out,
                                 void* stream) {
  if (!charges || !spheres || !fine_shape || !fx || !fy || !fz || !out) return fail(CDFT_E_ARG, "far_fine: null argument");
  if (n_spheres < 1 || n_spheres > MAXSP) return fail(CDFT_E_ARG, "far_fine: 1..%d spheres supported", MAXSP);
  if (fine_shape[0] < 1 || fine_shape[1] < 1 || fine_shape[2] < 1) return fail(CDFT_E_ARG, "far_fine: bad shape");
  FarSpheres sp;
  sp.n = n_spheres;
  for (int a = 0; a < n_spheres; ++a) {
    sp.cx[a] = spheres[6 * a];
    sp.cy[a] = spheres[6 * a + 1];
    sp.cz[a] = spheres[6 * a + 2];
    sp.radius[a] = spheres[6 * a + 3];
    sp.centre[a] = spheres[6 * a + 4];
    sp.width[a] = spheres[6 * a + 5];
    sp.z[a] = charges[a];
  }
  const long n = (long)fine_shape[0] * fine_shape[1] * fine_shape[2];
  hipLaunchKernelGGL(far_fine_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, sp, fine_shape[0],
                     fine_shape[1], fine_shape[2], fx, fy, fz, volume, out);
  return check_launch("far_fine_kernel");
}

extern "C" int cdft_interp_transpose_axis_f64(const double* in, const int* dims, int axis, int n_coarse,
                                              const int* range, const int* first, const double* w, double* out,
                                              void* stream) {
  if (!in || !dims || !range || !first || !w || !out) return fail(CDFT_E_ARG, "interp_transpose: null argument");
  if (axis < 0 || axis > 2 || n_coarse < 1) return fail(CDFT_E_ARG, "interp_transpose: bad axis or size");
  if (in == out) return fail(CDFT_E_ARG, "interp_transpose: out must not alias in");
  const int e[3] = {axis == 0 ? n_coarse : dims[0], axis == 1 ? n_coarse : dims[1], axis == 2 ? n_coarse : dims[2]};
  const long n = (long)e[0] * e[1] * e[2];
  hipLaunchKernelGGL(interp_transpose_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, in, dims[0],
                     dims[1], dims[2], axis, n_coarse, range, first, w, out);
  return check_launch("interp_transpose_kernel");
}

extern "C" int cdft_far_blend_f64(const int* shape, double h, const double* origin, double margin, double width,
                                  const double* plain, const double* lumped, double* out, void* stream) {
  if (!shape || !origin || !plain || !lumped || !out) return fail(CDFT_E_ARG, "far_blend: null argument");
  const long n = (long)shape[0] * shape[1] * shape[2];
  hipLaunchKernelGGL(far_blend_kernel, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, shape[0], shape[1],
                     shape[2], h, origin[0], origin[1], origin[2], margin, width, plain, lumped, out);
  return check_launch("far_blend_kernel");
}
