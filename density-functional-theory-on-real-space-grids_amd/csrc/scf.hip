// Pointwise and reduction kernels of the Kohn-Sham step (float64): VWN LDA, density, the Hartree
// cusp split, the lumped-potential assembly, small solves and the SCF norms.
//
// LDA: Slater exchange + VWN5 correlation, spin-restricted, zeta = 0, with the density threshold
// 1e-14 (reference src/cdft/xc/base.py:35,171-210, xc/lda.py:115-159,201-217). The potential is
// the analytic derivative of the same energy density; at zeta = 0 the spin-interpolation terms and
// their derivatives vanish identically.
#include "common.h"

namespace cdft {

constexpr int RBS = 256;
constexpr int SC_MAX_PARTIALS = 1024;

__device__ __forceinline__ void lda_vwn(double n, double* e_out, double* v_out) {
  const double half = 0.5 * n;
  const double total = half + half;
  if (total < 1.0e-14) {
    *e_out = 0.0;
    *v_out = 0.0;
    return;
  }
  const double floor_ = 0.5 * 1.0e-14;
  const double ns = half > floor_ ? half : floor_;
  const double third = 1.0 / 3.0;
  const double slater = 0.75 * pow(3.0 / M_PI, third);
  const double factor = pow(2.0, third) * slater;
  const double ns43 = pow(ns, 4.0 / 3.0);
  const double ex = -factor * (ns43 + ns43);
  const double vx = -factor * (4.0 / 3.0) * pow(ns, third);
  const double nt = ns + ns;
  const double rs = pow(3.0 / (4.0 * M_PI * nt), third);
  const double x = sqrt(rs);
  const double A = 0.0310907, x0 = -0.10498, b = 3.72744, c = 12.9352;
  const double X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c;
  const double q = sqrt(4.0 * c - b * b);
  const double at = atan(q / (2.0 * x + b));
  const double para = A * (log(x * x / X) + (2.0 * b / q) * at -
                           (b * x0 / X0) * (log((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / q) * at));
  const double dpara = A * (2.0 / x - (2.0 * x + 2.0 * b) / X -
                            (b * x0 / X0) * (2.0 / (x - x0) - (2.0 * x + 2.0 * b + 2.0 * x0) / X));
  *e_out = ex + nt * para;
  *v_out = vx + (para - (x / 6.0) * dpara);
}

__global__ void lda_kernel(const double* __restrict__ n, long len, double* __restrict__ e,
                           double* __restrict__ v) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  double ee, vv;
  lda_vwn(n[p], &ee, &vv);
  e[p] = ee;
  v[p] = vv;
}

// Elementwise expressions over up to three arrays and a host scalar.
__global__ void ewise_kernel(int op, const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ c, double s, double* __restrict__ out, long len) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  double r = 0.0;
  switch (op) {
    case 0: r = a[p] * b[p]; break;                 // a b
    case 1: r = (a[p] * b[p]) * c[p]; break;        // (a b) c
    case 2: r = s / a[p]; break;                    // s / a
    case 3: r = a[p] / b[p]; break;                 // a / b
    case 4: r = a[p] * s; break;                    // a s
    case 5: r = (a[p] + b[p]) / c[p]; break;        // (a + b) / c
    case 6: r = a[p] - b[p]; break;                 // a - b
    case 7: r = (a[p] / s) * s; break;              // (a / s) s  (measure weight round trip)
    case 8: r = s; break;                           // fill
    case 9: r = a[p] + b[p]; break;                 // a + b
    case 10: r = a[p] * ((b[p] * c[p] + b[len + p] * c[len + p]) + b[2 * len + p] * c[2 * len + p]); break;  // a (b . c)
    case 11: r = (a[p] * b[p]) * c[p]; out[len + p] = (a[p] * b[len + p]) * c[p]; break;  // (a b_r) c, two rows
    default: r = 0.0;
  }
  out[p] = r;
}

// rho = sum_i occ_i phi_i^2 (occupied states only), n = rho f^2.
__global__ void density_kernel(const double* __restrict__ phi, int k, long len, const double* __restrict__ f2,
                               double o0, double o1, double o2, double o3, double* __restrict__ rho,
                               double* __restrict__ n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  const double occ[4] = {o0, o1, o2, o3};
  double r = 0.0;
  for (int i = 0; i < k; ++i) {
    const double v = phi[(long)i * len + p];
    r = r + occ[i] * (v * v);
  }
  rho[p] = r;
  if (n) n[p] = r * f2[p];
}

// out = base - sum_a c_a fields_a   (sign = -1) or base + sum_a c_a fields_a (sign = +1); the
// coefficients c live on the device. Sum over atoms first, in atom order.
__global__ void atom_combine_kernel(const double* __restrict__ base, const double* __restrict__ c,
                                    const double* __restrict__ fields, int na, long len, double sign,
                                    double* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  double s = c[0] * fields[p];
  for (int a = 1; a < na; ++a) s = s + c[a] * fields[(long)a * len + p];
  out[p] = sign > 0.0 ? s + base[p] : base[p] - s;
}

// Dense solve A x = nucleus_f2 * rho_nuclei for the Hartree split coefficients (na <= 8).
__global__ __launch_bounds__(64) void small_solve_kernel(const double* __restrict__ a_in, const double* __restrict__ scale,
                                   const double* __restrict__ b_in, int na, double* __restrict__ x) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a[64], b[8];
  for (int i = 0; i < na * na; ++i) a[i] = a_in[i];
  for (int i = 0; i < na; ++i) b[i] = scale[i] * b_in[i];
  for (int col = 0; col < na; ++col) {  // partial pivoting
    int piv = col;
    for (int r = col + 1; r < na; ++r)
      if (fabs(a[r * na + col]) > fabs(a[piv * na + col])) piv = r;
    if (piv != col) {
      for (int j = 0; j < na; ++j) {
        const double t = a[col * na + j];
        a[col * na + j] = a[piv * na + j];
        a[piv * na + j] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < na; ++r) {
      const double f = a[r * na + col] / a[col * na + col];
      for (int j = col; j < na; ++j) a[r * na + j] -= f * a[col * na + j];
      b[r] -= f * b[col];
    }
  }
  for (int i = na - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < na; ++j) s -= a[i * na + j] * x[j];
    x[i] = s / a[i * na + i];
  }
}

// Pulay/DIIS coefficients on the device (reference src/cdft/scf/mixing.py:86-162): from the Gram
// matrix G of the residual history (leading dimension ldg), the m active slots in age order
// (oldest first), G_ij = gram[slot_i][slot_j] + 1e-12 max_k G_kk delta_ij, solve
//   [G  -1] [c]   [0]
//   [1^T 0] [l] = [1]
// by Gaussian elimination with partial pivoting; coef[2i] = c_i, coef[2i+1] = alpha c_i.
struct Slots {
  int s[16];
};
__global__ __launch_bounds__(64) void pulay_kernel(const double* __restrict__ gram, int ldg, Slots slots, int m, double alpha,
                             double* __restrict__ coef) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = m + 1;
  double a[17 * 17], b[17];
  double scale = 0.0;
  for (int i = 0; i < m; ++i) {
    const double d = gram[slots.s[i] * ldg + slots.s[i]];
    scale = d > scale ? d : scale;
  }
  if (scale == 0.0) scale = 1.0;
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) a[i * n + j] = gram[slots.s[i] * ldg + slots.s[j]];
    a[i * n + i] += 1.0e-12 * scale;
    a[i * n + m] = -1.0;
    a[m * n + i] = 1.0;
    b[i] = 0.0;
  }
  a[m * n + m] = 0.0;
  b[m] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (fabs(a[r * n + col]) > fabs(a[piv * n + col])) piv = r;
    if (piv != col) {
      for (int j = 0; j < n; ++j) {
        const double t = a[col * n + j];
        a[col * n + j] = a[piv * n + j];
        a[piv * n + j] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < n; ++r) {
      const double f = a[r * n + col] / a[col * n + col];
      for (int j = col; j < n; ++j) a[r * n + j] -= f * a[col * n + j];
      b[r] -= f * b[col];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = b[i];
    for (int j = i + 1; j < n; ++j) t -= a[i * n + j] * b[j];
    b[i] = t / a[i * n + i];
  }
  for (int i = 0; i < m; ++i) {
    coef[2 * i] = b[i];
    coef[2 * i + 1] = alpha * b[i];
  }
}

// Reductions with block partials (no atomics): mode 0: sum a*b*w (w, b optional),
// mode 1: sum w*|a|, mode 2: max |a - b|.
__global__ __launch_bounds__(RBS) void reduce_partial_kernel(int mode, const double* __restrict__ a,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ w, long len,
                                                             double* __restrict__ partials) {
  __shared__ double sh[RBS / 64];
  double acc = 0.0;
  const long stride = (long)gridDim.x * RBS;
  for (long p = (long)blockIdx.x * RBS + threadIdx.x; p < len; p += stride) {
    if (mode == 0) {
      double t = a[p];
      if (b) t = t * b[p];
      if (w) t = t * w[p];
      acc += t;
    } else if (mode == 1) {
      acc += w[p] * fabs(a[p]);
    } else {
      const double d = fabs(a[p] - b[p]);
      acc = d > acc ? d : acc;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(acc, off, 64);
    acc = (mode == 2) ? (o > acc ? o : acc) : acc + o;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int i = 1; i < RBS / 64; ++i) r = (mode == 2) ? (sh[i] > r ? sh[i] : r) : r + sh[i];
    partials[blockIdx.x] = r;
  }
}

// out[slot] = scale * (reduction of partials) (+ out[slot] if accumulate).
// One wavefront: lane l combines partials l, l+64, ... in order, then a fixed butterfly.
__global__ __launch_bounds__(64) void reduce_final_kernel(int mode, const double* __restrict__ partials,
                                                          int nblocks, double scale, int accumulate,
                                                          double* __restrict__ out, int slot) {
  double r = (mode == 2) ? 0.0 : 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 64) r = (mode == 2) ? (partials[i] > r ? partials[i] : r) : r + partials[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(r, off, 64);
    r = (mode == 2) ? (o > r ? o : r) : r + o;
  }
  if (threadIdx.x == 0) {
    r = r * scale;
    out[slot] = accumulate ? out[slot] + r : r;
  }
}

static unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// Up to four dot products over one length in one pass (reduce mode 0 each: sum a b w, b and w
// optional). For each product the per-thread order, the block combination and the final pass are
// reduce_partial_kernel's and reduce_final_kernel's, so every result is bitwise the separate call's.
struct Dots {
  const double* a[4];
  const double* b[4];
  const double* w[4];
};

struct DotOut {
  double scale[4];
  int accumulate[4];
  int slot[4];
};

template <int ND>
__global__ __launch_bounds__(RBS) void dots_partial_kernel(Dots d, long len, double* __restrict__ partials) {
  __shared__ double sh[ND][RBS / 64];
  double acc[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) acc[i] = 0.0;
  const long stride = (long)gridDim.x * RBS;
  for (long p = (long)blockIdx.x * RBS + threadIdx.x; p < len; p += stride) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double t = d.a[i][p];
      if (d.b[i]) t = t * d.b[i][p];
      if (d.w[i]) t = t * d.w[i][p];
      acc[i] += t;
    }
  }
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    for (int off = 32; off > 0; off >>= 1) acc[i] = acc[i] + __shfl_down(acc[i], off, 64);
    if ((threadIdx.x & 63) == 0) sh[i][threadIdx.x >> 6] = acc[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double r = sh[i][0];
      for (int j = 1; j < RBS / 64; ++j) r = r + sh[i][j];
      partials[(long)i * gridDim.x + blockIdx.x] = r;
    }
  }
}

// One wavefront per product.
__global__ __launch_bounds__(256) void dots_final_kernel(int nd, const double* __restrict__ partials, int nblocks,
                                                        DotOut o, double* __restrict__ out) {
  const int i = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  if (i >= nd) return;
  const double* pp = partials + (long)i * nblocks;
  double r = 0.0;
  for (int j = lane; j < nblocks; j += 64) r = r + pp[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) r = r + __shfl_xor(r, off, 64);
  if (lane == 0) {
    r = r * o.scale[i];
    out[o.slot[i]] = o.accumulate[i] ? out[o.slot[i]] + r : r;
  }
}

static int red_blocks(long n) {
  long b = (n + 1023) / 1024;
  if (b < 1) b = 1;
  if (b > SC_MAX_PARTIALS) b = SC_MAX_PARTIALS;
  return (int)b;
}

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_lda_vwn_f64(const double* n, long len, double* e, double* v, void* stream) {
  if (!n || !e || !v || len < 1) return fail(CDFT_E_ARG, "lda_vwn: bad arguments");
  hipLaunchKernelGGL(lda_kernel, dim3(nblk(len)), dim3(256), 0, (hipStream_t)stream, n, len, e, v);
  return check_launch("lda_vwn");
}

extern "C" int cdft_ewise_f64(int op, const double* a, const double* b, const double* c, double s,
                              double* out, long len, void* stream) {
  if (!out || len < 1 || op < 0 || op > 11) return fail(CDFT_E_ARG, "ewise: bad arguments (op %d)", op);
  const bool need_a = op != 8,
             need_b = (op == 0 || op == 1 || op == 3 || op == 5 || op == 6 || op == 9 || op == 10 || op == 11);
  const bool need_c = (op == 1 || op == 5 || op == 10 || op == 11);
  if ((need_a && !a) || (need_b && !b) || (need_c && !c)) return fail(CDFT_E_ARG, "ewise: missing operand (op %d)", op);
  hipLaunchKernelGGL(ewise_kernel, dim3(nblk(len)), dim3(256), 0, (hipStream_t)stream, op, a, b, c, s, out, len);
  return check_launch("ewise");
}

// occ: host array of k <= 4 occupations.
extern "C" int cdft_density_f64(const double* phi, int k, long len, const double* f2, const double* occ,
                                double* rho, double* n, void* stream) {
  if (!phi || !occ || !rho || k < 1 || k > 4 || (n && !f2)) return fail(CDFT_E_ARG, "density: bad arguments");
  double o[4] = {0, 0, 0, 0};
  for (int i = 0; i < k; ++i) o[i] = occ[i];
  hipLaunchKernelGGL(density_kernel, dim3(nblk(len)), dim3(256), 0, (hipStream_t)stream, phi, k, len, f2, o[0], o[1],
                     o[2], o[3], rho, n);
  return check_launch("density");
}

extern "C" int cdft_atom_combine_f64(const double* base, const double* c, const double* fields, int na,
                                     long len, double sign, double* out, void* stream) {
  if (!base || !c || !fields || !out || na < 1) return fail(CDFT_E_ARG, "atom_combine: bad arguments");
  hipLaunchKernelGGL(atom_combine_kernel, dim3(nblk(len)), dim3(256), 0, (hipStream_t)stream, base, c, fields, na,
                     len, sign, out);
  return check_launch("atom_combine");
}

extern "C" int cdft_small_solve_f64(const double* a, const double* scale, const double* b, int na, double* x,
                                    void* stream) {
  if (!a || !scale || !b || !x || na < 1 || na > 8) return fail(CDFT_E_ARG, "small_solve: bad arguments");
  hipLaunchKernelGGL(small_solve_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, a, scale, b, na, x);
  return check_launch("small_solve");
}

// slots: host array of the m active history slots, oldest first (m <= 16, slots < ldg).
extern "C" int cdft_pulay_coefficients_f64(const double* gram, int ldg, const int* slots, int m, double alpha,
                                           double* coef, void* stream) {
  if (!gram || !slots || !coef || m < 1 || m > 16 || ldg < m) return fail(CDFT_E_ARG, "pulay: bad arguments (m=%d)", m);
  Slots sl;
  for (int i = 0; i < 16; ++i) sl.s[i] = 0;
  for (int i = 0; i < m; ++i) {
    if (slots[i] < 0 || slots[i] >= ldg) return fail(CDFT_E_ARG, "pulay: slot %d out of range", slots[i]);
    sl.s[i] = slots[i];
  }
  hipLaunchKernelGGL(pulay_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, gram, ldg, sl, m, alpha, coef);
  return check_launch("pulay");
}

extern "C" long cdft_reduce_workspace_doubles(long len) { return red_blocks(len); }

// out[slots[i]] = scale[i] * sum a_i b_i w_i (+= if accumulate[i]) for i < nd <= 4, in one pass over len
// values; a, b, w, scale, accumulate, slots are host arrays (b[i], w[i] may be NULL). Partials: nd x
// cdft_reduce_workspace_doubles(len) doubles. Bitwise nd cdft_reduce_f64 calls of mode 0.
extern "C" int cdft_reduce_dots_f64(int nd, const double* const* a, const double* const* b, const double* const* w,
                                    long len, const double* scale, const int* accumulate, const int* slots,
                                    double* partials, long partials_len, double* out, void* stream) {
  if (nd < 1 || nd > 4 || !a || !scale || !accumulate || !slots || !partials || !out || len < 1)
    return fail(CDFT_E_ARG, "reduce_dots: bad arguments (nd=%d)", nd);
  Dots d{};
  DotOut o{};
  for (int i = 0; i < nd; ++i) {
    if (!a[i]) return fail(CDFT_E_ARG, "reduce_dots: product %d has no first operand", i);
    d.a[i] = a[i];
    d.b[i] = b ? b[i] : nullptr;
    d.w[i] = w ? w[i] : nullptr;
    o.scale[i] = scale[i];
    o.accumulate[i] = accumulate[i];
    o.slot[i] = slots[i];
  }
  const int blocks = red_blocks(len);
  if (partials_len < (long)nd * blocks) return fail(CDFT_E_CAPACITY, "reduce_dots: partials workspace too small");
  hipStream_t s = (hipStream_t)stream;
  switch (nd) {
    case 1: hipLaunchKernelGGL(dots_partial_kernel<1>, dim3(blocks), dim3(RBS), 0, s, d, len, partials); break;
    case 2: hipLaunchKernelGGL(dots_partial_kernel<2>, dim3(blocks), dim3(RBS), 0, s, d, len, partials); break;
    case 3: hipLaunchKernelGGL(dots_partial_kernel<3>, dim3(blocks), dim3(RBS), 0, s, d, len, partials); break;
    default: hipLaunchKernelGGL(dots_partial_kernel<4>, dim3(blocks), dim3(RBS), 0, s, d, len, partials); break;
  }
  int rc = check_launch("reduce_dots_partial");
  if (rc) return rc;
  hipLaunchKernelGGL(dots_final_kernel, dim3(1), dim3(256), 0, s, nd, partials, blocks, o, out);
  return check_launch("reduce_dots_final");
}

// out[slot] = scale * reduce(mode, a, b, w) (+= if accumulate). mode 0: sum a b w; 1: sum w |a|;
// 2: max |a - b|.
extern "C" int cdft_reduce_f64(int mode, const double* a, const double* b, const double* w, long len,
                               double scale, int accumulate, double* partials, long partials_len,
                               double* out, int slot, void* stream) {
  if (!a || !out || !partials || len < 1 || mode < 0 || mode > 2) return fail(CDFT_E_ARG, "reduce: bad arguments");
  if ((mode == 1 && !w) || (mode == 2 && !b)) return fail(CDFT_E_ARG, "reduce: missing operand for mode %d", mode);
  const int blocks = red_blocks(len);
  if (partials_len < blocks) return fail(CDFT_E_CAPACITY, "reduce: partials workspace too small");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(reduce_partial_kernel, dim3(blocks), dim3(RBS), 0, s, mode, a, b, w, len, partials);
  int rc = check_launch("reduce_partial");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(64), 0, s, mode, partials, blocks, scale, accumulate, out, slot);
  return check_launch("reduce_final");
}
