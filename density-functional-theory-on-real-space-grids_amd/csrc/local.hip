// K1 and K2 for the plain operator H = -1/2 lap + V of the analytic model systems (reference
// src/cdft/operators/hamiltonian.py:92-125, the Laplacian of src/cdft/grid.py:714-740 and
// src/cdft/operators/laplacian.py:85-105): central second-derivative weights of order 2-12 along each
// axis, accumulated axis 0, 1, 2 and, within an axis, from offset -R to +R (the reference's order), then
// -1/2 of the sum plus V phi. The halo follows the boundary mode (grid.py:808-827): 0 = zero
// (Dirichlet), 1 = odd reflection about walls one spacing outside the outermost points (the entry at
// the wall is zero, the entry m spacings beyond it is minus the interior point m spacings inside).
//
// The model systems are solved once each, not in the hot loop, so the kernel is a plain gather: each
// thread takes points of one vector in C order (z fastest, so a wavefront's loads of a tap are
// contiguous apart from the wrap at a row end) and reads its 6R + 1 neighbours through the caches.
// Modes as the divergence-form kernels: 0 = H phi, 1 = the first Chebyshev step (H y - c y) a,
// 2 = a fused step (H y - c y) a - b prev, 4 = per-block sums of q (H phi - e phi)^2 for the residual
// norms (combined in a fixed order by a second launch).
#include "common.h"
#include "ritz_tail.h"

namespace cdft {

constexpr int LOCAL_THREADS = 256;
constexpr int LOCAL_POINTS = 4;  // points per thread, one block's chunk = 1024 consecutive points
constexpr int LOCAL_MAX_REACH = 6;

struct LocalTaps {
  double c[2 * LOCAL_MAX_REACH + 1];  // c_j / h^2 for j = -R..R
};

// phi at index i along an axis of n points (base + i * stride when inside).
__device__ __forceinline__ double local_tap(const double* __restrict__ f, long base, long stride, int i, int n,
                                           int boundary) {
  if (i >= 0 && i < n) return f[base + (long)i * stride];
  if (boundary == 0 || i == -1 || i == n) return 0.0;
  const int src = i < 0 ? -i - 2 : 2 * n - i;
  return -f[base + (long)src * stride];
}

template <int R, int MODE>
__global__ __launch_bounds__(LOCAL_THREADS) void local_kernel(const double* __restrict__ phi,
                                                              const double* __restrict__ prev,
                                                              double* __restrict__ out,
                                                              const double* __restrict__ pot, int n0, int n1, int n2,
                                                              int k, int boundary, LocalTaps taps,
                                                              const double* __restrict__ table, int step) {
  const int v = (int)(blockIdx.x % (unsigned)k);
  const long tile = blockIdx.x / (unsigned)k;
  const long n = (long)n0 * n1 * n2;
  const long plane = (long)n1 * n2;
  const double* f = phi + (long)v * n;
  double cc = 0.0, ca = 0.0, cb = 0.0;
  if (MODE == 1 || MODE == 2) {
    cc = table[3 * step];
    ca = table[3 * step + 1];
    cb = table[3 * step + 2];
  }
  const double ev = MODE == 4 ? table[v] : 0.0;
  double racc = 0.0;
#pragma unroll
  for (int u = 0; u < LOCAL_POINTS; ++u) {
    const long p = (tile * LOCAL_POINTS + u) * LOCAL_THREADS + threadIdx.x;
    if (p >= n) break;
    const int x = (int)(p / plane);
    const long rem = p - (long)x * plane;
    const int y = (int)(rem / n2);
    const int z = (int)(rem - (long)y * n2);
    const long b0 = rem, b1 = (long)x * plane + z, b2 = (long)x * plane + (long)y * n2;
    double lap = local_tap(f, b0, plane, x - R, n0, boundary) * taps.c[0];
#pragma unroll
    for (int j = 1; j <= 2 * R; ++j) lap += local_tap(f, b0, plane, x - R + j, n0, boundary) * taps.c[j];
#pragma unroll
    for (int j = 0; j <= 2 * R; ++j) lap += local_tap(f, b1, n2, y - R + j, n1, boundary) * taps.c[j];
#pragma unroll
    for (int j = 0; j <= 2 * R; ++j) lap += local_tap(f, b2, 1, z - R + j, n2, boundary) * taps.c[j];
    const double pv = f[p];
    double val = lap * -0.5 + pv * pot[p];
    if (MODE == 1) val = (val - cc * pv) * ca;
    if (MODE == 2) val = (val - cc * pv) * ca - cb * prev[(long)v * n + p];
    if (MODE == 4) {
      const double r = val - ev * pv;
      racc += (r * r) * prev[p];  // prev holds q in this mode
    } else {
      out[(long)v * n + p] = val;
    }
  }
  if constexpr (MODE == 4) {
    // fixed-order block sum: a butterfly per wavefront, then the wavefronts in order
    __shared__ double waves[LOCAL_THREADS / 64];
    double sum = racc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((threadIdx.x & 63) == 0) waves[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double total = waves[0];
      for (int w = 1; w < LOCAL_THREADS / 64; ++w) total += waves[w];
      out[blockIdx.x] = total;
    }
  }
}

// Residual norms from the blocks' sums (one wavefront per vector: lane l sums the sums of tiles l,
// l + 64, ... in order, then a fixed butterfly); with a tail, block k also builds the next filter's
// Chebyshev table and block k + 1 writes max |G - I| (the divergence-form path's combine).
__global__ __launch_bounds__(64) void local_resnorm_kernel(const double* __restrict__ partials, int tiles, int k,
                                                           double* __restrict__ norms,
                                                           const double* __restrict__ evals, int index,
                                                           const double* __restrict__ bounds, int degree,
                                                           double* __restrict__ table,
                                                           const double* __restrict__ gram,
                                                           double* __restrict__ orth) {
  const int b = blockIdx.x;
  if (b < k) {
    double s = 0.0;
    for (int t = threadIdx.x; t < tiles; t += 64) s += partials[(long)t * k + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) norms[b] = sqrt(s > 0.0 ? s : 0.0);
    return;
  }
  if (threadIdx.x != 0) return;
  if (b == k)
    cheb_table_rows(evals, index, bounds, degree, table);
  else
    max_abs_minus_identity(gram, k, orth);
}

static long local_tiles(int n0, int n1, int n2) {
  const long n = (long)n0 * n1 * n2;
  const long chunk = (long)LOCAL_POINTS * LOCAL_THREADS;
  return (n + chunk - 1) / chunk;
}

static int local_validate(const double* phi, const void* out, const double* pot, int k, int n0, int n1, int n2,
                          const double* coef, int order, int boundary) {
  if (!phi || !out || !pot || !coef) return fail(CDFT_E_ARG, "local: null pointer argument");
  if (k < 1 || n0 < 1 || n1 < 1 || n2 < 1)
    return fail(CDFT_E_ARG, "local: sizes must be positive (k=%d, n=%d,%d,%d)", k, n0, n1, n2);
  if (order < 2 || order > 2 * LOCAL_MAX_REACH || order % 2)
    return fail(CDFT_E_ARG, "local: unsupported order %d (2, 4, ..., 12)", order);
  if (boundary != 0 && boundary != 1) return fail(CDFT_E_ARG, "local: boundary is 0 (zero) or 1 (odd reflection)");
  const int reach = order / 2;
  // an odd reflection mirrors at most reach - 1 interior points per side
  if (boundary == 1 && (n0 < reach || n1 < reach || n2 < reach))
    return fail(CDFT_E_ARG, "local: odd reflection needs at least %d points per axis", reach);
  if ((long double)local_tiles(n0, n1, n2) * k > 2.0e9L) return fail(CDFT_E_ARG, "local: block too large");
  if ((const void*)phi == out) return fail(CDFT_E_ARG, "local: out must not alias phi");
  return CDFT_OK;
}

template <int MODE>
static int local_launch(const double* phi, const double* prev, double* out, const double* pot, int k, int n0, int n1,
                        int n2, const double* coef, int order, int boundary, const double* table, int step,
                        hipStream_t s) {
  LocalTaps taps{};
  for (int j = 0; j <= order; ++j) taps.c[j] = coef[j];
  const unsigned blocks = (unsigned)(local_tiles(n0, n1, n2) * k);
#define CDFT_LOCAL_CASE(RR)                                                                                   \
  case RR:                                                                                                    \
    hipLaunchKernelGGL((local_kernel<RR, MODE>), dim3(blocks), dim3(LOCAL_THREADS), 0, s, phi, prev, out, pot, \
                       n0, n1, n2, k, boundary, taps, table, step);                                           \
    break;
  switch (order / 2) {
    CDFT_LOCAL_CASE(1)
    CDFT_LOCAL_CASE(2)
    CDFT_LOCAL_CASE(3)
    CDFT_LOCAL_CASE(4)
    CDFT_LOCAL_CASE(5)
    CDFT_LOCAL_CASE(6)
    default: return fail(CDFT_E_ARG, "local: unsupported order %d", order);
  }
#undef CDFT_LOCAL_CASE
  return check_launch("local_kernel");
}

}  // namespace cdft

using namespace cdft;

extern "C" long cdft_local_partials_len(int n0, int n1, int n2, int k) {
  if (k < 1 || n0 < 1 || n1 < 1 || n2 < 1) return -1;
  return local_tiles(n0, n1, n2) * k;
}

extern "C" int cdft_local_apply_f64(const double* phi, const double* pot, double* out, int k, int n0, int n1, int n2,
                                    const double* coef, int order, int boundary, void* stream) {
  int rc = local_validate(phi, out, pot, k, n0, n1, n2, coef, order, boundary);
  if (rc) return rc;
  return local_launch<0>(phi, nullptr, out, pot, k, n0, n1, n2, coef, order, boundary, nullptr, 0,
                         (hipStream_t)stream);
}

extern "C" int cdft_local_filter_f64(const double* x, double* out, double* work1, double* work2, const double* pot,
                                     int k, int n0, int n1, int n2, const double* coef, int order, int boundary,
                                     const double* table, int degree, void* stream) {
  int rc = local_validate(x, out, pot, k, n0, n1, n2, coef, order, boundary);
  if (rc) return rc;
  if (!work1 || !work2 || !table) return fail(CDFT_E_ARG, "local_filter: null workspace or table");
  if (degree < 1) return fail(CDFT_E_ARG, "local_filter: degree must be >= 1");
  if (out == work1 || out == work2 || work1 == work2 || x == work1 || x == work2)
    return fail(CDFT_E_ARG, "local_filter: x, out and the two work blocks must be distinct");
  // three rotating buffers; the step that writes buffer 0 (= out) is the last one
  double* bufs[3] = {out, work1, work2};
  const int offset = ((-(degree - 1)) % 3 + 3) % 3;
  const double* y = x;
  const double* prev = x;
  hipStream_t s = (hipStream_t)stream;
  for (int step = 0; step < degree; ++step) {
    double* dst = bufs[(step + offset) % 3];
    rc = step == 0 ? local_launch<1>(y, nullptr, dst, pot, k, n0, n1, n2, coef, order, boundary, table, step, s)
                   : local_launch<2>(y, prev, dst, pot, k, n0, n1, n2, coef, order, boundary, table, step, s);
    if (rc) return rc;
    prev = y;
    y = dst;
  }
  return CDFT_OK;
}

// norms[v] = ||H phi_v - evals[v] phi_v||_q. With index >= 0 the combine also writes the next filter's
// table (from evals[index] and bounds) and max |gram - I| to orth, as cdft_hdiv_resnorm_tail_f64.
extern "C" int cdft_local_resnorm_f64(const double* phi, const double* pot, const double* evals, const double* q, int k,
                                      int n0, int n1, int n2, const double* coef, int order, int boundary,
                                      double* partials, long partials_len, double* norms, int index,
                                      const double* bounds, int degree, double* table, const double* gram,
                                      double* orth, void* stream) {
  int rc = local_validate(phi, partials, pot, k, n0, n1, n2, coef, order, boundary);
  if (rc) return rc;
  if (!evals || !q || !norms) return fail(CDFT_E_ARG, "local_resnorm: null evals, q or norms");
  const bool tail = index >= 0;
  if (tail && (!bounds || !table || !gram || !orth || index >= k || degree < 1))
    return fail(CDFT_E_ARG, "local_resnorm: null bounds, table, gram or orth, or index/degree out of range");
  const long tiles = local_tiles(n0, n1, n2);
  if (partials_len < tiles * k)
    return fail(CDFT_E_CAPACITY, "local_resnorm: partials hold %ld doubles, the launch needs %ld", partials_len,
                tiles * k);
  hipStream_t s = (hipStream_t)stream;
  rc = local_launch<4>(phi, q, partials, pot, k, n0, n1, n2, coef, order, boundary, evals, 0, s);
  if (rc) return rc;
  hipLaunchKernelGGL(local_resnorm_kernel, dim3(tail ? k + 2 : k), dim3(64), 0, s, partials, (int)tiles, k, norms,
                     evals, tail ? index : 0, bounds, degree, table, gram, orth);
  return check_launch("local_resnorm");
}
