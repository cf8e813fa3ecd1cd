// PBE exchange-correlation and the density-gradient (sigma) path of the Kohn-Sham potentials.
//
// pbe_point evaluates the spin-restricted PBE energy density per volume (Perdew, Burke, Ernzerhof,
// Phys. Rev. Lett. 77, 3865 (1996)) with its analytic derivatives in n and sigma = |grad n|^2, as the
// reference writes it (src/cdft/xc/gga.py:43-90 on the per-spin form of xc/base.py:115-189):
//   exchange    e_x = -C n^(4/3) F(s^2),  F = 1 + kappa - kappa / (1 + mu s^2 / kappa),
//               s^2 = sigma / (4 k_F^2 n^(8/3)),  C = (3/4)(3/pi)^(1/3),  k_F = (3 pi^2)^(1/3);
//   correlation e_c = n (eps_c(r_s) + H(t^2)),  eps_c the modified PW92 fit (xc/lda.py:64-108) at zeta = 0,
//               H = gamma ln(1 + (beta/gamma) t^2 (1 + A t^2) / (1 + A t^2 + A^2 t^4)),
//               A = (beta/gamma) / (exp(-eps_c / gamma) - 1),  t^2 = sigma pi / (16 k_F n^(7/3)).
// The reference differentiates by autograd; here the derivatives are written out. Its thresholds
// hold: a point with n < 1e-14 has zero energy and potentials (xc/base.py:35), and the per-spin
// sigma_ss = sigma / 4 is clamped at 1e-40 (xc/base.py:44), which at a point of vanishing gradient
// leaves exchange a zero sigma-derivative and correlation half of it (only the unclamped cross term
// sigma_ud carries it there).
#include "common.h"

namespace cdft {

namespace pbe {
constexpr double PI = 3.14159265358979323846;
constexpr double KAPPA = 0.804;
constexpr double MU = 0.2195149727645171;
constexpr double BETA = 0.06672455060314922;
constexpr double SIGMA_FLOOR = 1.0e-40;
constexpr double DENSITY_THRESHOLD = 1.0e-14;
// modified PW92, paramagnetic: A, alpha1, beta1..beta4
constexpr double PA = 0.0310907, PALPHA = 0.21370, PB1 = 7.5957, PB2 = 3.5876, PB3 = 1.6382, PB4 = 0.49294;
}  // namespace pbe

// PW92 G(r_s) of the paramagnetic fit and its derivative in r_s.
__device__ __forceinline__ void pw92_para(double rs, double* g, double* dg) {
  using namespace pbe;
  const double srs = sqrt(rs);
  const double q = 2.0 * PA * (PB1 * srs + PB2 * rs + PB3 * rs * srs + PB4 * rs * rs);
  const double dq = 2.0 * PA * (0.5 * PB1 / srs + PB2 + 1.5 * PB3 * srs + 2.0 * PB4 * rs);
  const double l = log1p(1.0 / q);
  *g = -2.0 * PA * (1.0 + PALPHA * rs) * l;
  // d/drs ln(1 + 1/q) = -q' / (q (q + 1))
  *dg = -2.0 * PA * PALPHA * l + 2.0 * PA * (1.0 + PALPHA * rs) * dq / (q * (q + 1.0));
}

__device__ __forceinline__ void pbe_point(double n, double sigma, double* e_out, double* vn_out, double* vs_out) {
  using namespace pbe;
  const double half = 0.5 * n;
  if (half + half < DENSITY_THRESHOLD) {
    *e_out = 0.0;
    *vn_out = 0.0;
    *vs_out = 0.0;
    return;
  }
  const double kf_c = cbrt(3.0 * PI * PI);
  const double slater = 0.75 * cbrt(3.0 / PI);
  const double gamma = (1.0 - log(2.0)) / (PI * PI);
  // the per-spin clamp of sigma / 4
  const bool floored = 0.25 * sigma < SIGMA_FLOOR;
  const double sx = floored ? 4.0 * SIGMA_FLOOR : sigma;                    // exchange's 4 sigma_ss
  const double sc = floored ? SIGMA_FLOOR + 0.5 * sigma + SIGMA_FLOOR : sigma;  // correlation's total
  const double dsx = floored ? 0.0 : 1.0;
  const double dsc = floored ? 0.5 : 1.0;

  // exchange
  const double n13 = cbrt(n);
  const double n43 = n * n13;
  const double n83 = n43 * n43;
  const double cs = 1.0 / (4.0 * kf_c * kf_c);
  const double s2 = cs * sx / n83;
  const double den = 1.0 + MU * s2 / KAPPA;
  const double fx = 1.0 + KAPPA - KAPPA / den;
  const double dfx = MU / (den * den);  // dF/ds^2
  const double x = -slater * n43;
  const double ex = x * fx;
  const double ex_s = x * dfx * (cs / n83) * dsx;
  const double ex_n = -(4.0 / 3.0) * slater * n13 * fx + x * dfx * (-(8.0 / 3.0) * s2 / n);

  // correlation
  const double rs = cbrt(3.0 / (4.0 * PI * n));
  double eps, deps_drs;
  pw92_para(rs, &eps, &deps_drs);
  const double deps = deps_drs * (-rs / (3.0 * n));  // d eps / dn
  const double b = BETA / gamma;
  const double ee = exp(-eps / gamma);
  const double a = b / (ee - 1.0);
  const double da = b * ee / (gamma * (ee - 1.0) * (ee - 1.0));  // dA / d eps
  const double ct = PI / (16.0 * kf_c);
  const double n73 = n * n * n13;  // n^(7/3)
  const double y = ct * sc / n73;  // t^2
  const double p = a * y;
  const double num = 1.0 + p;
  const double dd = 1.0 + p + p * p;
  const double r = y * num / dd;
  const double h = gamma * log1p(b * r);
  const double dh_dr = gamma * b / (1.0 + b * r);
  const double dr_dy = num / dd - p * p * (2.0 + p) / (dd * dd);
  const double dr_da = -y * y * p * (2.0 + p) / (dd * dd);
  const double dy_ds = ct / n73;
  const double dy_dn = -(7.0 / 3.0) * y / n;
  const double ec = n * (eps + h);
  const double ec_s = n * dh_dr * dr_dy * dy_ds * dsc;
  const double ec_n = (eps + h) + n * (deps + dh_dr * (dr_dy * dy_dn + dr_da * da * deps));

  *e_out = ex + ec;
  *vn_out = ex_n + ec_n;
  *vs_out = ex_s + ec_s;
}

__global__ void pbe_kernel(const double* __restrict__ n, const double* __restrict__ sigma, long len,
                           double* __restrict__ e, double* __restrict__ vn, double* __restrict__ vs) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  double a, b, c;
  pbe_point(n[p], sigma[p], &a, &b, &c);
  e[p] = a;
  vn[p] = b;
  vs[p] = c;
}

// The pointwise part of the sigma path at the nodes or at the grid points (reference
// src/cdft/scf/step.py:253-279, operators/quadrature.py:757-773): from rho = n / f^2, its gradient and
// grad u,
//   n = f^2 rho,  grad n = f^2 (grad rho - 2 rho grad u),  sigma = |grad n|^2,
//   (e, v_n, v_sigma) = PBE(n, sigma),  a = f^2 v_n,  flux = 2 v_sigma f^2 grad n,
//   cross = -2 flux . grad u,
// written as e, w_s a, w_s cross and w_f flux (w_s the node weights W P or the grid's h^3 far, w_f the
// node weights or the grid's far factor: the factors the lumped weights apply next).
__global__ void gga_points_kernel(const double* __restrict__ rho, const double* __restrict__ grad_rho,
                                  const double* __restrict__ f2, const double* __restrict__ grad_u,
                                  const double* __restrict__ w_s, const double* __restrict__ w_f, long len,
                                  double* __restrict__ e, double* __restrict__ aw, double* __restrict__ crossw,
                                  double* __restrict__ fluxw) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  const double r = rho[p], f = f2[p];
  const double two_r = 2.0 * r;
  double gn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) gn[d] = f * rn_sub(grad_rho[d * len + p], rn_mul(two_r, grad_u[d * len + p]));
  const double sigma = rn_add(rn_add(rn_mul(gn[0], gn[0]), rn_mul(gn[1], gn[1])), rn_mul(gn[2], gn[2]));
  double ep, vn, vs;
  pbe_point(f * r, sigma, &ep, &vn, &vs);
  e[p] = ep;
  const double a = f * vn;
  const double coef = rn_mul(2.0 * vs, f);
  double fl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) fl[d] = rn_mul(coef, gn[d]);
  const double dot = rn_add(rn_add(rn_mul(fl[0], grad_u[p]), rn_mul(fl[1], grad_u[len + p])),
                               rn_mul(fl[2], grad_u[2 * len + p]));
  const double cross = -2.0 * dot;
  const double ws = w_s[p], wf = w_f[p];
  aw[p] = ws * a;
  crossw[p] = ws * cross;
#pragma unroll
  for (int d = 0; d < 3; ++d) fluxw[d * len + p] = wf * fl[d];
}

// out = d0 F0 + d1 F1 + d2 F2 of a (3, N) field with the central order-2P first-derivative taps and a
// zero halo, each derivative accumulated over the taps in order (the reference's
// grid.partial_derivative, laplacian.py:111-130) and the three added in axis order.
template <int P>
__global__ void divergence_kernel(const double* __restrict__ field, int n0, int n1, int n2,
                                  const double* __restrict__ coef, double* __restrict__ out) {
  const long n = (long)n0 * n1 * n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % n2);
  const int i1 = (int)((p / n2) % n1);
  const int i0 = (int)(p / ((long)n1 * n2));
  const int idx[3] = {i0, i1, i2};
  const int dims[3] = {n0, n1, n2};
  const long stride[3] = {(long)n1 * n2, (long)n2, 1};
  double total = 0.0;
  for (int ax = 0; ax < 3; ++ax) {
    const double* src = field + (long)ax * n;
    double acc = 0.0;
    bool first = true;
    for (int j = 0; j < 2 * P + 1; ++j) {
      const double c = coef[j];
      if (c == 0.0) continue;
      const int o = idx[ax] + j - P;
      const double v = (o >= 0 && o < dims[ax]) ? src[p + (long)(j - P) * stride[ax]] : 0.0;
      acc = first ? v * c : acc + v * c;
      first = false;
    }
    total = ax == 0 ? acc : total + acc;
  }
  out[p] = total;
}

// The XC lumped weights of a GGA: ((a + (-near_div - h^3 far_div)) + cross), the order in which the
// reference adds the scalar part, the gradient part and the cross part (step.py:264-271,
// quadrature.py:640-650,718-729).
__global__ void gga_combine_kernel(const double* __restrict__ a, const double* __restrict__ near_div,
                                   const double* __restrict__ far_div, const double* __restrict__ cross, double h3,
                                   long len, double* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  const double grad = rn_sub(-near_div[p], rn_mul(h3, far_div[p]));
  out[p] = rn_add(rn_add(a[p], grad), cross[p]);
}

static unsigned gga_blocks(long len) { return (unsigned)((len + 255) / 256); }

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_pbe_f64(const double* n, const double* sigma, long len, double* e, double* vn, double* vs,
                            void* stream) {
  if (!n || !sigma || !e || !vn || !vs || len < 1) return fail(CDFT_E_ARG, "pbe: bad arguments");
  hipLaunchKernelGGL(pbe_kernel, dim3(gga_blocks(len)), dim3(256), 0, (hipStream_t)stream, n, sigma, len, e, vn, vs);
  return check_launch("pbe");
}

extern "C" int cdft_gga_points_f64(const double* rho, const double* grad_rho, const double* f2, const double* grad_u,
                                   const double* w_s, const double* w_f, long len, double* e, double* aw,
                                   double* crossw, double* fluxw, void* stream) {
  if (!rho || !grad_rho || !f2 || !grad_u || !w_s || !w_f || !e || !aw || !crossw || !fluxw || len < 1)
    return fail(CDFT_E_ARG, "gga_points: bad arguments");
  hipLaunchKernelGGL(gga_points_kernel, dim3(gga_blocks(len)), dim3(256), 0, (hipStream_t)stream, rho, grad_rho, f2,
                     grad_u, w_s, w_f, len, e, aw, crossw, fluxw);
  return check_launch("gga_points");
}

extern "C" int cdft_divergence_f64(const double* field, const int* shape, const double* coef, int order, double* out,
                                   void* stream) {
  if (!field || !shape || !coef || !out) return fail(CDFT_E_ARG, "divergence: bad arguments");
  const int n0 = shape[0], n1 = shape[1], n2 = shape[2];
  if (n0 < 1 || n1 < 1 || n2 < 1) return fail(CDFT_E_ARG, "divergence: bad shape");
  const long n = (long)n0 * n1 * n2;
  hipStream_t s = (hipStream_t)stream;
  switch (order) {
    case 2: hipLaunchKernelGGL(divergence_kernel<1>, dim3(gga_blocks(n)), dim3(256), 0, s, field, n0, n1, n2, coef, out); break;
    case 4: hipLaunchKernelGGL(divergence_kernel<2>, dim3(gga_blocks(n)), dim3(256), 0, s, field, n0, n1, n2, coef, out); break;
    case 6: hipLaunchKernelGGL(divergence_kernel<3>, dim3(gga_blocks(n)), dim3(256), 0, s, field, n0, n1, n2, coef, out); break;
    case 8: hipLaunchKernelGGL(divergence_kernel<4>, dim3(gga_blocks(n)), dim3(256), 0, s, field, n0, n1, n2, coef, out); break;
    default: return fail(CDFT_E_ARG, "divergence: unsupported order %d", order);
  }
  return check_launch("divergence");
}

extern "C" int cdft_gga_combine_f64(const double* a, const double* near_div, const double* far_div, const double* cross,
                                    double h3, long len, double* out, void* stream) {
  if (!a || !near_div || !far_div || !cross || !out || len < 1) return fail(CDFT_E_ARG, "gga_combine: bad arguments");
  hipLaunchKernelGGL(gga_combine_kernel, dim3(gga_blocks(len)), dim3(256), 0, (hipStream_t)stream, a, near_div,
                     far_div, cross, h3, len, out);
  return check_launch("gga_combine");
}
