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
#include <algorithm>

#include "common.h"

#include <hipfft/hipfft.h>

namespace cdft {

// The x transforms, the multiply by the kernel and the inverse x transforms of one solve as one
// launch: each block takes `cols` adjacent (ky, kz) columns into LDS (their first n0 x entries; the
// rest are the zero padding), runs a Stockham autosort FFT of length p0 over them (radix 4, 2, 3 and 5
// stages, twiddles exp(-2 pi i j / p0) from sincospi), multiplies by K, runs the inverse FFT and
// writes the first n0 x entries (the only ones the inverse (y, z) transforms read). Adjacent columns
// are adjacent in memory, so the loads and stores of a block are coalesced across its columns; the
// strided batched 1-D transforms it replaces read one column a transform.
constexpr int XFFT_MAX_STAGES = 24;
constexpr int XFFT_LDS_BYTES = 61440;

struct XFFTPlan {
  int n;       // transform length p0
  int n0;      // occupied (and kept) x entries of a column
  int ncols;   // columns: p1 (p2/2+1)
  int cols;    // columns per block
  int pitch;   // LDS doubles2 per column (n + 1: adjacent columns start in different banks)
  int stages;
  int radix[XFFT_MAX_STAGES];
};

struct PoissonPlan {
  hipfftHandle yz_forward;  // 2-D D2Z over (p1, p2), n0 slabs
  hipfftHandle yz_inverse;  // 2-D Z2D over (p1, p2), n0 slabs
  hipfftHandle x_forward;   // 1-D Z2Z along x, stride p1 (p2/2+1), out of place
  hipfftHandle x_inverse;   // 1-D Z2Z along x, in place
  int n0, n1, n2;           // box
  int p0, p1, p2;           // padded
  bool fused;               // the x transforms and the multiply run as one column kernel (xfft)
  XFFTPlan xfft;
};

// Which x path cdft_poisson_solve_f64 takes: 0 = the column kernel where the plan has one (default),
// 1 = the rocFFT strided transforms with a separate multiply (set by cdft_poisson_set_x_path).
static int g_x_path = 0;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double f) { return make_double2(a.x * f, a.y * f); }
// i^DIR z: DIR = -1 gives -i z, DIR = +1 gives i z
template <int DIR>
__device__ __forceinline__ double2 rot90(double2 z) {
  return DIR < 0 ? make_double2(z.y, -z.x) : make_double2(-z.y, z.x);
}

// One Stockham stage of radix R over `cols` columns: with n = N / s the current length and m = n / R,
// y[q + s (R p + k)] = w^(p k) sum_j x[q + s (p + j m)] exp(DIR 2 pi i j k / R), w = exp(DIR 2 pi i / n),
// for p < m, q < s; w^(p k) = tw[p k s] (conjugated for DIR = +1), p k s < N.
template <int DIR, int R>
__device__ void xfft_stage_r(const double2* __restrict__ src, double2* __restrict__ dst,
                             const double2* __restrict__ tw, int N, int pitch, int cols, int s) {
  const int m = N / s / R;
  const int per = N / R;
  for (int b = threadIdx.x; b < cols * per; b += blockDim.x) {
    const int c = b / per, t = b - c * per;
    const int p = t / s, q = t - p * s;
    const double2* x = src + c * pitch + q + s * p;
    double2* y = dst + c * pitch + q + s * R * p;
    double2 A[R];
    if constexpr (R == 4) {
      const double2 a0 = x[0], a1 = x[s * m], a2 = x[2 * s * m], a3 = x[3 * s * m];
      const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
      A[0] = cadd(t0, t2);
      A[2] = csub(t0, t2);
      A[1] = cadd(t1, rot90<DIR>(t3));
      A[3] = csub(t1, rot90<DIR>(t3));
    } else if constexpr (R == 2) {
      const double2 a0 = x[0], a1 = x[s * m];
      A[0] = cadd(a0, a1);
      A[1] = csub(a0, a1);
    } else if constexpr (R == 3) {
      const double2 a0 = x[0], a1 = x[s * m], a2 = x[2 * s * m];
      const double2 sm = cadd(a1, a2), df = csub(a1, a2);
      const double2 mid = csub(a0, cscale(sm, 0.5));
      const double2 e = rot90<DIR>(cscale(df, 0.86602540378443864676));  // sin(2 pi / 3)
      A[0] = cadd(a0, sm);
      A[1] = cadd(mid, e);
      A[2] = csub(mid, e);
    } else {
      static_assert(R == 5, "radices 2, 3, 4 and 5");
      const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2 pi/5), cos(4 pi/5)
      const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2 pi/5), sin(4 pi/5)
      const double2 a0 = x[0], a1 = x[s * m], a2 = x[2 * s * m], a3 = x[3 * s * m], a4 = x[4 * s * m];
      const double2 s14 = cadd(a1, a4), d14 = csub(a1, a4), s23 = cadd(a2, a3), d23 = csub(a2, a3);
      A[0] = cadd(a0, cadd(s14, s23));
      const double2 b1 = cadd(a0, cadd(cscale(s14, c1), cscale(s23, c2)));
      const double2 b2 = cadd(a0, cadd(cscale(s14, c2), cscale(s23, c1)));
      const double2 e1 = rot90<DIR>(cadd(cscale(d14, s1), cscale(d23, s2)));
      const double2 e2 = rot90<DIR>(csub(cscale(d14, s2), cscale(d23, s1)));
      A[1] = cadd(b1, e1);
      A[4] = csub(b1, e1);
      A[2] = cadd(b2, e2);
      A[3] = csub(b2, e2);
    }
    y[0] = A[0];
#pragma unroll
    for (int k = 1; k < R; ++k) {
      double2 w = tw[p * k * s];
      if (DIR > 0) w.y = -w.y;
      y[s * k] = cmul(A[k], w);
    }
  }
}

template <int DIR>
__device__ void xfft_stage(const double2* src, double2* dst, const double2* tw, int N, int pitch, int cols, int r,
                           int s) {
  switch (r) {
    case 4: xfft_stage_r<DIR, 4>(src, dst, tw, N, pitch, cols, s); break;
    case 2: xfft_stage_r<DIR, 2>(src, dst, tw, N, pitch, cols, s); break;
    case 3: xfft_stage_r<DIR, 3>(src, dst, tw, N, pitch, cols, s); break;
    default: xfft_stage_r<DIR, 5>(src, dst, tw, N, pitch, cols, s); break;
  }
}

template <int DIR>
__device__ double2* xfft_run(double2* a, double2* b, const double2* tw, const XFFTPlan& pl) {
  int s = 1;
  for (int st = 0; st < pl.stages; ++st) {
    const int r = pl.radix[st];
    xfft_stage<DIR>(a, b, tw, pl.n, pl.pitch, pl.cols, r, s);
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    s *= r;
  }
  return a;  // the buffer holding the result
}

__global__ __launch_bounds__(256) void xfft_convolve_kernel(const double2* __restrict__ slabs,
                                                            const double* __restrict__ kernel,
                                                            double2* __restrict__ spectrum, XFFTPlan pl) {
  extern __shared__ double2 lds[];
  const int N = pl.n, cols = pl.cols, pitch = pl.pitch;
  double2* tw = lds;
  double2* bufa = lds + N;
  double2* bufb = bufa + cols * pitch;
  const long col0 = (long)blockIdx.x * cols;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    double sn, cs;
    sincospi(-2.0 * (double)j / (double)N, &sn, &cs);
    tw[j] = make_double2(cs, sn);
  }
  // the first n0 entries of each column; the zero padding above them
  for (int e = threadIdx.x; e < cols * N; e += blockDim.x) {
    const int x = e / cols, c = e - x * cols;
    const long col = col0 + c;
    double2 v = make_double2(0.0, 0.0);
    if (x < pl.n0 && col < pl.ncols) v = slabs[(long)x * pl.ncols + col];
    bufa[c * pitch + x] = v;
  }
  __syncthreads();
  double2* f = xfft_run<-1>(bufa, bufb, tw, pl);
  double2* g = f == bufa ? bufb : bufa;
  for (int e = threadIdx.x; e < cols * N; e += blockDim.x) {
    const int x = e / cols, c = e - x * cols;
    const long col = col0 + c;
    if (col < pl.ncols) f[c * pitch + x] = cscale(f[c * pitch + x], kernel[(long)x * pl.ncols + col]);
  }
  __syncthreads();
  double2* h = xfft_run<1>(f, g, tw, pl);
  for (int e = threadIdx.x; e < cols * pl.n0; e += blockDim.x) {
    const int x = e / cols, c = e - x * cols;
    const long col = col0 + c;
    if (col < pl.ncols) spectrum[(long)x * pl.ncols + col] = h[c * pitch + x];
  }
}

// The radix sequence of a transform length (4s first, then 2, 3 and 5); false when another prime
// divides it or it needs more stages than the plan holds.
static bool xfft_factor(int n, XFFTPlan& pl) {
  pl.stages = 0;
  int r = n;
  for (int f : {4, 2, 3, 5}) {
    while (r % f == 0) {
      if (pl.stages == XFFT_MAX_STAGES) return false;
      pl.radix[pl.stages++] = f;
      r /= f;
    }
  }
  return r == 1;
}

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
  // the column kernel, where the length factors into radices 2-5 and one column fits the LDS budget;
  // four columns a block (a sweep of 1-16 columns and 64-256 threads on the 180^3, 216^3 and
  // 256 x 216^2 boxes: 114, 164 and 171 us a solve at four, against 133, 189 and 183 at sixteen)
  XFFTPlan& x_plan = plan->xfft;
  x_plan.n = p0;
  x_plan.n0 = n0;
  x_plan.ncols = column;
  x_plan.pitch = p0 + 1;
  const int budget = XFFT_LDS_BYTES / (int)sizeof(double2) - p0;  // after the twiddles
  x_plan.cols = budget >= 2 * x_plan.pitch ? std::min(4, budget / (2 * x_plan.pitch)) : 0;
  plan->fused = x_plan.cols >= 1 && xfft_factor(p0, x_plan);
  *handle = plan;
  return CDFT_OK;
}

extern "C" int cdft_poisson_set_x_path(int which) {
  if (which != 0 && which != 1) return fail(CDFT_E_ARG, "poisson_set_x_path: 0 (column kernel) or 1 (rocFFT)");
  g_x_path = which;
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
  if (pl->fused && g_x_path == 0) {
    const XFFTPlan& xp = pl->xfft;
    const unsigned blocks = (unsigned)((xp.ncols + xp.cols - 1) / xp.cols);
    const size_t lds = (size_t)(xp.n + 2 * xp.cols * xp.pitch) * sizeof(double2);
    hipLaunchKernelGGL(xfft_convolve_kernel, dim3(blocks), dim3(256), lds, s, (const double2*)sl, kernel, (double2*)sp,
                       xp);
    rc = check_launch("poisson xfft_convolve");
    if (rc) return rc;
  } else {
    if (hipfftExecZ2Z(pl->x_forward, sl, sp, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
      return fail(CDFT_E_LIBRARY, "hipfftExecZ2Z (x, forward) failed");
    hipLaunchKernelGGL(spectrum_multiply_kernel, dim3(nb(total)), dim3(256), 0, s, (double2*)spectrum, kernel, total);
    rc = check_launch("poisson spectrum_multiply");
    if (rc) return rc;
    if (hipfftExecZ2Z(pl->x_inverse, sp, sp, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
      return fail(CDFT_E_LIBRARY, "hipfftExecZ2Z (x, inverse) failed");
  }
  if (hipfftExecZ2D(pl->yz_inverse, sp, padded_out) != HIPFFT_SUCCESS)
    return fail(CDFT_E_LIBRARY, "hipfftExecZ2D (y, z slabs) failed");
  hipLaunchKernelGGL(gather_kernel, dim3(nb(n)), dim3(256), 0, s, padded_out, pl->n0, pl->n1, pl->n2, pl->p1, pl->p2,
                     potential);
  return check_launch("poisson gather");
}
