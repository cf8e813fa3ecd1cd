// K1 / K2: the cusp-factorised staggered divergence-form Hamiltonian apply and the scaled Chebyshev
// recurrence fused into its epilogue.
//
//   (A phi)_m = 0.5 * invw_m * sum_d sum_k c_k w_d(m-1-k) F_d(m-1-k)  +  V_m phi_m,
//   F_d(n)    = sum_j c_j phi(n+1+j)                        (face n sits at x_n + h/2)
//
// with c_k the staggered order-2P weights (already divided by h), a zero halo, and the face weights
// w_d stored on the box extended by 2P along axis d, indexed from face -P. Blocks are 4x16x16
// output tiles (x slowest, z contiguous). Each thread owns one (y, z) column of 4 x-planes: the
// x-faces come from a register pencil loaded straight from global memory (the x halo never enters
// LDS), the y- and z-faces from an LDS plane tile with a +-(2P-1) halo. Weighted faces are computed
// once per face and shared through LDS.
#include "common.h"

namespace cdft {

constexpr int TX = 4, TY = 16, TZ = 16, NT = TY * TZ;

template <int P>
struct Taps {
  double c[2 * P];  // c_k / h, k = -P .. P-1
};

// Epilogue modes: 0 = A phi; 1 = (A y - c y) a; 2 = (A y - c y) a - b prev.
//
// One block = one 4x16x16 tile of one vector (blockIdx.y is the vector). Every global load of a
// thread is independent and issued before the first barrier: its 4+2(2P-1) x-pencil, four y-halo
// values (threads with ty < 2P-1... 14 rows x 16 columns x 4 planes) and four z-halo values, plus
// the face weights. LDS holds the cross-shaped plane tiles (corners unused) and the weighted faces.
template <typename T, int P, int MODE>
__global__ __launch_bounds__(NT) void hdiv_kernel(
    const T* __restrict__ phi, const T* __restrict__ prev, T* __restrict__ out,
    const T* __restrict__ w0, const T* __restrict__ w1, const T* __restrict__ w2,
    const T* __restrict__ invw, const T* __restrict__ pot, int k, int n0, int n1, int n2,
    Taps<P> taps, const double* __restrict__ table, int step, int tiles_z, int tiles_y,
    int tiles_x) {
  constexpr int R = 2 * P - 1;  // stencil reach: 2P-1 points each side
  constexpr int LY = TY + 2 * R, LZ = TZ + 2 * R;
  static_assert(2 * R <= TY && 2 * R <= TZ, "halo rows must fit one pass of the block");
  __shared__ T sphi[TX][LY][LZ];
  __shared__ T syf[TX][TY + R][TZ];
  __shared__ T szf[TX][TY][TZ + R];

  const unsigned ntiles = (unsigned)tiles_z * tiles_y * tiles_x;
  const unsigned tile = xcd_tile(blockIdx.x, ntiles);
  const int v = blockIdx.y;
  const int bz = tile % tiles_z;
  const int by = (tile / tiles_z) % tiles_y;
  const int bx = tile / (tiles_z * tiles_y);
  const int X0 = bx * TX, Y0 = by * TY, Z0 = bz * TZ;
  const int tid = threadIdx.x;
  const int tz = tid % TZ, ty = tid / TZ;
  const int y = Y0 + ty, z = Z0 + tz;
  const bool yz_in = (y < n1) && (z < n2);
  const long plane = (long)n1 * n2;
  const long N = plane * n0;
  const int n1f = n1 + 2 * P, n2f = n2 + 2 * P;
  const T* ph = phi + (long)v * N;

  T c[2 * P];
#pragma unroll
  for (int i = 0; i < 2 * P; ++i) c[i] = (T)taps.c[i];

  // ---- global loads (all independent) ----
  T px[TX + 2 * R];
#pragma unroll
  for (int i = 0; i < TX + 2 * R; ++i) {
    const int gx = X0 - R + i;
    px[i] = (yz_in && gx >= 0 && gx < n0) ? ph[gx * plane + (long)y * n2 + z] : T(0);
  }
  // y halo: rows Y0-R .. Y0-1 (hy < R) and Y0+TY .. Y0+TY+R-1 (hy >= R), column z.
  T hyv[TX];
  const bool do_hy = ty < 2 * R;
  const int hy_row = ty < R ? Y0 - R + ty : Y0 + TY + (ty - R);
  const int hy_l = ty < R ? ty : TY + ty;  // local LDS row
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    const int gx = X0 + p;
    hyv[p] = (do_hy && gx < n0 && z < n2 && hy_row >= 0 && hy_row < n1) ? ph[gx * plane + (long)hy_row * n2 + z] : T(0);
  }
  // z halo: columns Z0-R .. Z0-1 and Z0+TZ .. Z0+TZ+R-1, row y.
  T hzv[TX];
  const bool do_hz = tz < 2 * R;
  const int hz_col = tz < R ? Z0 - R + tz : Z0 + TZ + (tz - R);
  const int hz_l = tz < R ? tz : TZ + tz;
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    const int gx = X0 + p;
    hzv[p] = (do_hz && gx < n0 && y < n1 && hz_col >= 0 && hz_col < n2) ? ph[gx * plane + (long)y * n2 + hz_col] : T(0);
  }
  // Face weights this thread applies: x faces of its pencil, y faces (rows fy = ty, ty+TY), z faces.
  T wx[TX + R];
#pragma unroll
  for (int fx = 0; fx < TX + R; ++fx) {
    const int a = X0 + fx;
    wx[fx] = (yz_in && a < n0 + 2 * P) ? w0[((long)a * n1 + y) * n2 + z] : T(0);
  }
  T wy0[TX], wy1[TX], wz0[TX], wz1[TX];
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    const int gx = X0 + p;
    const bool xin = gx < n0;
    const int ay0 = Y0 + ty, ay1 = Y0 + ty + TY;
    wy0[p] = (xin && z < n2 && ay0 < n1f) ? w1[((long)gx * n1f + ay0) * n2 + z] : T(0);
    wy1[p] = (xin && z < n2 && ty < R && ay1 < n1f) ? w1[((long)gx * n1f + ay1) * n2 + z] : T(0);
    const int az0 = Z0 + tz, az1 = Z0 + tz + TZ;
    wz0[p] = (xin && y < n1 && az0 < n2f) ? w2[((long)gx * n1 + y) * n2f + az0] : T(0);
    wz1[p] = (xin && y < n1 && tz < R && az1 < n2f) ? w2[((long)gx * n1 + y) * n2f + az1] : T(0);
  }

  // ---- stage the plane tiles ----
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    sphi[p][ty + R][tz + R] = px[R + p];
    if (do_hy) sphi[p][hy_l][tz + R] = hyv[p];
    if (do_hz) sphi[p][ty + R][hz_l] = hzv[p];
  }
  __syncthreads();

  // ---- weighted faces ----
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * sphi[p][ty + i][tz + R];
      syf[p][ty][tz] = f * wy0[p];
    }
    if (ty < R) {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * sphi[p][ty + TY + i][tz + R];
      syf[p][ty + TY][tz] = f * wy1[p];
    }
    {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * sphi[p][ty + R][tz + i];
      szf[p][ty][tz] = f * wz0[p];
    }
    if (tz < R) {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * sphi[p][ty + R][tz + TZ + i];
      szf[p][ty][tz + TZ] = f * wz1[p];
    }
  }
  T gxf[TX + R];
#pragma unroll
  for (int fx = 0; fx < TX + R; ++fx) {
    T f = 0;
#pragma unroll
    for (int i = 0; i < 2 * P; ++i) f += c[i] * px[fx + i];
    gxf[fx] = f * wx[fx];
  }
  T cc = 0, ca = 0, cb = 0;
  if (MODE >= 1) {
    cc = (T)table[3 * step];
    ca = (T)table[3 * step + 1];
    cb = (T)table[3 * step + 2];
  }
  __syncthreads();

  if (!yz_in) return;
#pragma unroll
  for (int p = 0; p < TX; ++p) {
    const int gx = X0 + p;
    if (gx < n0) {
      T ax = 0, ay = 0, az = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) ax += c[i] * gxf[p + R - i];
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) ay += c[i] * syf[p][ty + R - i][tz];
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) az += c[i] * szf[p][ty][tz + R - i];
      const T lap = (ax + ay) + az;
      const long idx = gx * plane + (long)y * n2 + z;
      const T pv = px[R + p];
      T val = lap * T(0.5) * invw[idx] + pv * pot[idx];
      if (MODE == 1) val = (val - cc * pv) * ca;
      if (MODE == 2) val = (val - cc * pv) * ca - cb * prev[(long)v * N + idx];
      out[(long)v * N + idx] = val;
    }
  }
}

struct Dims {
  int tiles_z, tiles_y, tiles_x;
  unsigned blocks;
};

static Dims dims_for(int n0, int n1, int n2) {
  Dims d;
  d.tiles_z = (n2 + TZ - 1) / TZ;
  d.tiles_y = (n1 + TY - 1) / TY;
  d.tiles_x = (n0 + TX - 1) / TX;
  d.blocks = (unsigned)d.tiles_z * d.tiles_y * d.tiles_x;
  return d;
}

template <typename T, int P, int MODE>
static int launch_p(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2,
                    const T* invw, const T* pot, int k, int n0, int n1, int n2, const double* coef,
                    const double* table, int step, hipStream_t stream) {
  Taps<P> taps;
  for (int i = 0; i < 2 * P; ++i) taps.c[i] = coef[i];
  const Dims d = dims_for(n0, n1, n2);
  hipLaunchKernelGGL((hdiv_kernel<T, P, MODE>), dim3(d.blocks, k), dim3(NT), 0, stream, phi, prev, out,
                     w0, w1, w2, invw, pot, k, n0, n1, n2, taps, table, step, d.tiles_z, d.tiles_y,
                     d.tiles_x);
  return check_launch("hdiv_kernel");
}

template <typename T, int MODE>
static int launch(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2,
                  const T* invw, const T* pot, int k, int n0, int n1, int n2, const double* coef,
                  int order, const double* table, int step, hipStream_t stream) {
  switch (order) {
    case 2: return launch_p<T, 1, MODE>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, table, step, stream);
    case 4: return launch_p<T, 2, MODE>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, table, step, stream);
    case 6: return launch_p<T, 3, MODE>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, table, step, stream);
    case 8: return launch_p<T, 4, MODE>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, table, step, stream);
    default: return fail(CDFT_E_ARG, "unsupported stencil order %d (2, 4, 6 or 8)", order);
  }
}

static int validate(const void* phi, const void* out, const void* w0, const void* w1,
                    const void* w2, const void* invw, const void* pot, int k, int n0, int n1,
                    int n2, const double* coef, int order) {
  if (!phi || !out || !w0 || !w1 || !w2 || !invw || !pot || !coef)
    return fail(CDFT_E_ARG, "hdiv: null pointer argument");
  if (k < 1 || n0 < 1 || n1 < 1 || n2 < 1)
    return fail(CDFT_E_ARG, "hdiv: sizes must be positive (k=%d, n=%d,%d,%d)", k, n0, n1, n2);
  if ((long double)n0 * n1 * n2 * k > 9.0e18L || k > 65535) return fail(CDFT_E_ARG, "hdiv: block too large");
  if (order != 2 && order != 4 && order != 6 && order != 8)
    return fail(CDFT_E_ARG, "hdiv: unsupported stencil order %d (2, 4, 6 or 8)", order);
  if (phi == out) return fail(CDFT_E_ARG, "hdiv: out must not alias phi");
  return CDFT_OK;
}

template <typename T>
static int filter_impl(const T* x, T* out, T* work1, T* work2, const T* w0, const T* w1,
                       const T* w2, const T* invw, const T* pot, int k, int n0, int n1, int n2,
                       const double* coef, int order, const double* table, int degree,
                       hipStream_t stream) {
  int rc = validate(x, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!work1 || !work2 || !table) return fail(CDFT_E_ARG, "cheb_filter: null workspace or table");
  if (degree < 1) return fail(CDFT_E_ARG, "cheb_filter: degree must be >= 1");
  if (out == work1 || out == work2 || work1 == work2 || x == work1 || x == work2)
    return fail(CDFT_E_ARG, "cheb_filter: x, out and the two work blocks must be distinct");
  // Three rotating buffers; the step that writes buffer 0 (= out) is the last one.
  T* bufs[3] = {out, work1, work2};
  const int offset = ((-(degree - 1)) % 3 + 3) % 3;
  const T* y = x;
  const T* prev = x;
  for (int s = 0; s < degree; ++s) {
    T* dst = bufs[(s + offset) % 3];
    if (s == 0) {
      rc = launch<T, 1>(y, nullptr, dst, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, s, stream);
    } else {
      rc = launch<T, 2>(y, prev, dst, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, s, stream);
    }
    if (rc) return rc;
    prev = y;
    y = dst;
  }
  return CDFT_OK;
}

}  // namespace cdft

using namespace cdft;

extern "C" int cdft_hdiv_apply_f64(const double* phi, const double* w0, const double* w1,
                                   const double* w2, const double* invw, const double* pot,
                                   double* out, int k, int n0, int n1, int n2, const double* coef,
                                   int order, void* stream) {
  int rc = validate(phi, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  return launch<double, 0>(phi, nullptr, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order,
                           nullptr, 0, (hipStream_t)stream);
}

extern "C" int cdft_hdiv_apply_f32(const float* phi, const float* w0, const float* w1,
                                   const float* w2, const float* invw, const float* pot, float* out,
                                   int k, int n0, int n1, int n2, const double* coef, int order,
                                   void* stream) {
  int rc = validate(phi, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  return launch<float, 0>(phi, nullptr, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order,
                          nullptr, 0, (hipStream_t)stream);
}

extern "C" int cdft_cheb_step_f64(const double* y, const double* prev, const double* w0,
                                  const double* w1, const double* w2, const double* invw,
                                  const double* pot, double* out, int k, int n0, int n1, int n2,
                                  const double* coef, int order, const double* table, int step,
                                  void* stream) {
  int rc = validate(y, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!table || step < 0) return fail(CDFT_E_ARG, "cheb_step: null table or negative step");
  if (step > 0 && !prev) return fail(CDFT_E_ARG, "cheb_step: step > 0 needs the previous block");
  if (step == 0)
    return launch<double, 1>(y, nullptr, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, 0, (hipStream_t)stream);
  return launch<double, 2>(y, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, step, (hipStream_t)stream);
}

extern "C" int cdft_cheb_step_f32(const float* y, const float* prev, const float* w0,
                                  const float* w1, const float* w2, const float* invw,
                                  const float* pot, float* out, int k, int n0, int n1, int n2,
                                  const double* coef, int order, const double* table, int step,
                                  void* stream) {
  int rc = validate(y, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!table || step < 0) return fail(CDFT_E_ARG, "cheb_step: null table or negative step");
  if (step > 0 && !prev) return fail(CDFT_E_ARG, "cheb_step: step > 0 needs the previous block");
  if (step == 0)
    return launch<float, 1>(y, nullptr, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, 0, (hipStream_t)stream);
  return launch<float, 2>(y, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, coef, order, table, step, (hipStream_t)stream);
}

extern "C" int cdft_cheb_filter_f64(const double* x, double* out, double* work1, double* work2,
                                    const double* w0, const double* w1, const double* w2,
                                    const double* invw, const double* pot, int k, int n0, int n1,
                                    int n2, const double* coef, int order, const double* table,
                                    int degree, void* stream) {
  return filter_impl<double>(x, out, work1, work2, w0, w1, w2, invw, pot, k, n0, n1, n2, coef,
                             order, table, degree, (hipStream_t)stream);
}

extern "C" int cdft_cheb_filter_f32(const float* x, float* out, float* work1, float* work2,
                                    const float* w0, const float* w1, const float* w2,
                                    const float* invw, const float* pot, int k, int n0, int n1,
                                    int n2, const double* coef, int order, const double* table,
                                    int degree, void* stream) {
  return filter_impl<float>(x, out, work1, work2, w0, w1, w2, invw, pot, k, n0, n1, n2, coef,
                            order, table, degree, (hipStream_t)stream);
}
