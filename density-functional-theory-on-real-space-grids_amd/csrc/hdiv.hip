// K1 / K2: the cusp-factorised staggered divergence-form Hamiltonian apply and the scaled Chebyshev
// recurrence fused into its epilogue.
//
//   (A phi)_m = 0.5 * invw_m * sum_d sum_k c_k w_d(m-1-k) F_d(m-1-k)  +  V_m phi_m,
//   F_d(n)    = sum_j c_j phi(n+1+j)                        (face n sits at x_n + h/2)
//
// with c_k the staggered order-2P weights (already divided by h), a zero halo, and the face weights
// w_d stored on the box extended by 2P along axis d, indexed from face -P.
//
// Layout: C order, x slowest, z contiguous. A block owns a (y, z) tile of at most 16 x 16 columns
// and marches along x through a segment of planes. Each thread keeps its column's x-direction state
// in registers -- the 2P values phi(m .. m+2P-1) and the 2P weighted x-faces F(m-P .. m+P-1) -- so
// the x stencil costs no LDS traffic and each phi value is read from memory once per segment (plus
// a 4P-2 plane warm-up). The y and z stencils of plane m go through an LDS plane tile with a
// +-(2P-1) halo; weighted y/z faces are computed once per face and shared through LDS. The global
// loads of plane m+1 are issued before the LDS work of plane m, so their latency overlaps it.
// Tile extents are chosen per grid (ceil(n / ceil(n / 16))), so n = 65 runs 13-wide tiles rather
// than four full tiles and a one-column tail.
#include <algorithm>
#include <atomic>
#include <iterator>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "ritz_tail.h"

namespace cdft {

constexpr int TMAX = 16;   // largest tile extent along y and z
constexpr int MAXH = 2;    // halo values staged per thread per plane (orders up to 8)
constexpr int NTMAX = 256; // largest block

// Halo values a column-kernel thread stages per plane: the 2R (TY + TZ) halo values of a tile over
// at most NTMAX threads (order 12, reach 11: 704 values at 16 x 16 tiles take three slots).
constexpr int halo_slots(int reach) { return reach <= 7 ? MAXH : 3; }

template <int P>
struct Taps {
  double c[2 * P];  // c_k / h, k = -P .. P-1
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int bytes) {
  // 0x00020000: the raw (untyped, stride 0) buffer format word of gfx9-family descriptors.
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// One geometry record (wx, wy, wz, invw) of four T at byte offset voff + soff; zeros when out of range.
// The 16-byte load is bit-cast as a whole to a vector of the element type: extracting single 32-bit
// lanes of it and bit-casting each (the obvious form) is compiled here into one 4-byte load read for
// every lane.
template <typename T>
__device__ __forceinline__ void bload_rec(__amdgpu_buffer_rsrc_t r, int voff, int soff, T& a, T& b, T& c, T& d);
template <>
__device__ __forceinline__ void bload_rec<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff, double& a, double& b,
                                                  double& c, double& d) {
  typedef __attribute__((ext_vector_type(2))) double d2;
  const d2 lo = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  const d2 hi = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16, soff, 0));
  a = lo[0];
  b = lo[1];
  c = hi[0];
  d = hi[1];
}
template <>
__device__ __forceinline__ void bload_rec<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff, float& a, float& b,
                                                 float& c, float& d) {
  typedef __attribute__((ext_vector_type(4))) float f4;
  const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  a = v[0];
  b = v[1];
  c = v[2];
  d = v[3];
}

template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ void bstore<double>(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void bstore<float>(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}

// Launch shape of the marching kernel (host-computed; see march_dims).
struct March {
  int ty, tz;            // tile extents (<= TMAX)
  int tiles_y, tiles_z;  // tiles per plane
  int seg, nseg;         // longest segment (planes), segments along x
  unsigned tiles;        // tiles_y * tiles_z * nseg
  int threads;           // block size: a multiple of 64 covering the tile and the halo staging
};

// Epilogue modes: 0 = A phi; 1 = (A y - c y) a; 2 = (A y - c y) a - b prev, where (c, a, b) is row
// `step` of the Chebyshev coefficient table; 3 = the transpose A^T x = 0.5 sum_d D_d^T[w_d D_d
// (invw x)] + V x, with phi = invw x and prev = x; 4 = the residual norm: nothing is stored, the
// block accumulates sum q (A phi - e_v phi)^2 over its points (prev = q, table = the eigenvalues e)
// and writes one partial to out[0] (the kernel points out at the block's slot). TZ is the tile extent along z (compile-time, so the LDS
// pitches below are constants and neighbouring reads pair into ds_read2). Threads are numbered
// row-major over the tile and every LDS row pitch is congruent to TZ modulo 16 doubles, so any 16
// consecutive lanes touch 16 consecutive bank pairs: the plane-tile, y-face and z-face accesses are
// free of bank conflicts. Consecutive blocks take the vectors of one tile, so the tile's geometry
// is reused from L2 by all of them.
template <typename T, int P, int TZ>
struct StencilLds {
  static constexpr int R = 2 * P - 1;
  static constexpr int LZ = TZ + (2 * R <= 16 ? 16 : 32);  // plane-tile pitch (>= TZ + 2R)
  static constexpr int YP = TZ;                            // y-face pitch
  static constexpr int ZP = TZ + 16;                       // z-face pitch (>= TZ + R)
  T phi[(TMAX + 2 * R) * LZ];
  T yf[(TMAX + R) * YP];
  T zf[TMAX * ZP];
};

// One block's work: x-segment `bs` of (y, z) tile (by, bz) of vector v (tile = bz + tiles_z (bs +
// nseg by)). Shared by the single-grid and the batched kernels.
template <typename T, int P, int MODE, int TZ>
__device__ __forceinline__ void stencil_block(
    const T* __restrict__ phi, const T* __restrict__ prev, T* __restrict__ out,
    const T* __restrict__ w0, const T* __restrict__ w1, const T* __restrict__ w2,
    const T* __restrict__ invw, const T* __restrict__ pot, const T* __restrict__ gp, int n0, int n1, int n2,
    const Taps<P>& taps, const double* __restrict__ table, int step, const March& md, unsigned tile, int v,
    StencilLds<T, P, TZ>& lds) {
  constexpr int R = 2 * P - 1;  // stencil reach: 2P-1 points each side
  constexpr int LZ = StencilLds<T, P, TZ>::LZ;  // plane-tile pitch (>= TZ + 2R)
  constexpr int YP = StencilLds<T, P, TZ>::YP;  // y-face pitch
  constexpr int ZP = StencilLds<T, P, TZ>::ZP;  // z-face pitch (>= TZ + R)
  constexpr int NH = halo_slots(R);             // halo slots per thread
  static_assert(LZ >= TZ + 2 * R && ZP >= TZ + R && (LZ - TZ) % 16 == 0, "LDS pitches");
  T* s_phi = lds.phi;
  T* s_yf = lds.yf;
  T* s_zf = lds.zf;

  const int TY = md.ty;
  // z-tiles fastest, then segments, then y-tiles: the XCD-contiguous block ranges then hold boxes of
  // whole z-rows of tiles over consecutive segments, so a segment's x warm-up planes (the previous
  // segment's) are mostly in the same XCD's L2 (about 6% faster at 65^3 than segments slowest).
  const int bz = (int)(tile % md.tiles_z);
  const int bs = (int)((tile / md.tiles_z) % md.nseg);
  const int by = (int)(tile / ((unsigned)md.tiles_z * md.nseg));
  // Segments split the planes evenly: segment s covers [s n0 / nseg, (s + 1) n0 / nseg).
  const int X0 = (int)((long)bs * n0 / md.nseg), X1 = (int)((long)(bs + 1) * n0 / md.nseg);
  const int Y0 = by * TY, Z0 = bz * TZ;
  const int t = threadIdx.x, nt = blockDim.x;
  const bool act = t < TY * TZ;
  const int ty = act ? t / TZ : 0, tz = act ? t - (t / TZ) * TZ : 0;
  const int y = Y0 + ty, z = Z0 + tz;
  const bool own = act && y < n1 && z < n2;
  // All loads are raw buffer loads: a per-array descriptor (base, size), a per-thread byte offset
  // that is the same in every plane, and the plane's byte offset as a uniform scalar. Lanes with
  // nothing to load (outside the grid or the tile) use an offset past the end of the buffer, which
  // the hardware returns as zero. Offsets are 32-bit (the host checks every array fits).
  constexpr int S = (int)sizeof(T);
  constexpr int NONE = (int)0x80000000;
  const int plane = n1 * n2;
  const int n1f = n1 + 2 * P, n2f = n2 + 2 * P;
  const int pl1 = n1f * n2, pl2 = n1 * n2f;
  const long vbase = (long)v * plane * n0;
  const int vbytes = plane * n0 * S;
  const auto r_phi = make_rsrc(phi + vbase, vbytes);
  const auto r_prev = make_rsrc(MODE == 4 ? prev : (MODE >= 2 ? prev + vbase : phi), vbytes);
  const auto r_out = make_rsrc(out + vbase, vbytes);
  const auto r_w0 = make_rsrc(w0, (n0 + 2 * P) * plane * S);
  const auto r_w1 = make_rsrc(w1, n0 * pl1 * S);
  const auto r_w2 = make_rsrc(w2, n0 * pl2 * S);
  // per-point geometry records (wx, wy, wz, invw) over the face-extended (y, z) rows
  const auto r_gp = make_rsrc(gp, n0 * n1f * n2f * 4 * S);
  const auto r_pt = make_rsrc(pot, n0 * plane * S);
  const int col = own ? (y * n2 + z) * S : NONE;

  T c[2 * P];
#pragma unroll
  for (int i = 0; i < 2 * P; ++i) c[i] = (T)taps.c[i];
  T cc = 0, ca = 0, cb = 0;
  if (MODE == 1 || MODE == 2) {
    cc = (T)table[3 * step];
    ca = (T)table[3 * step + 1];
    cb = (T)table[3 * step + 2];
  }
  const T ev = MODE == 4 ? (T)table[v] : T(0);
  T racc = 0;  // MODE 4: this thread's sum of q r^2 over its planes

  // Halo slots of this thread (the same in every plane): in-plane source byte offset (NONE =
  // outside the grid, staged as zero) and LDS destination (-1 = no slot).
  const int nhy = 2 * R * TZ, nhalo = nhy + 2 * R * TY;
  int hsrc[NH], hdst[NH];
#pragma unroll
  for (int s = 0; s < NH; ++s) {
    const int h = t + s * nt;
    hsrc[s] = NONE;
    hdst[s] = -1;
    if (h < nhy) {
      const int hr = h / TZ, hc = h - (h / TZ) * TZ;
      const int rl = hr < R ? hr : TY + hr;
      const int gy = Y0 - R + rl, gz = Z0 + hc;
      hdst[s] = rl * LZ + R + hc;
      if (gy >= 0 && gy < n1 && gz < n2) hsrc[s] = (gy * n2 + gz) * S;
    } else if (h < nhalo) {
      const int hh = h - nhy;
      const int r = hh / (2 * R), hc = hh - (hh / (2 * R)) * (2 * R);
      const int cl = hc < R ? hc : TZ + hc;
      const int gy = Y0 + r, gz = Z0 - R + cl;
      hdst[s] = (R + r) * LZ + cl;
      if (gy < n1 && gz >= 0 && gz < n2) hsrc[s] = (gy * n2 + gz) * S;
    }
  }
  const bool zin = act && z < n2, yin = act && y < n1;
  const int ay0 = Y0 + ty, ay1 = Y0 + ty + TY;
  const int az0 = Z0 + tz, az1 = Z0 + tz + TZ;
  // the record of column (y, z) (records exist for y < n1 + 2P, z < n2 + 2P; the wy / wz parts are
  // zero where the face is not this thread's, as the separate loads' out-of-range offsets were)
  const int rec = (act && y < n1f && z < n2f) ? (y * n2f + z) * 4 * S : NONE;
  const int wy1_off = (zin && ty < R && ay1 < n1f) ? (ay1 * n2 + z) * S : NONE;
  const int wz1_off = (yin && tz < R && az1 < n2f) ? (y * n2f + az1) * S : NONE;

  // ---- x warm-up: phi(X0-2P+1 .. X0+2P-1) and the faces F(X0-P .. X0+P-2) ----
  T qphi[2 * P], qf[2 * P];
  {
    T pen[4 * P - 1];
#pragma unroll
    for (int q = 0; q < 4 * P - 1; ++q) {
      const int gx = X0 - 2 * P + 1 + q;  // block-uniform
      pen[q] = (gx >= 0 && gx < n0) ? bload<T>(r_phi, col, gx * plane * S) : T(0);
    }
#pragma unroll
    for (int j = 0; j < 2 * P - 1; ++j) {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * pen[j + i];
      qf[j] = f * bload<T>(r_w0, col, (X0 + j) * plane * S);  // extended index of face X0-P+j
    }
    qf[2 * P - 1] = 0;
#pragma unroll
    for (int j = 0; j < 2 * P; ++j) qphi[j] = pen[2 * P - 1 + j];
  }

  // Per-plane operands used before the plane's first barrier, loaded one plane ahead (their latency
  // overlaps the LDS work of a plane).
  struct Plane {
    T wxf;        // w0 at the newest x-face of the plane (extended index m + 2P - 1)
    T hal[NH];  // y/z halo values
    T wy0, wz0, iw;
  };
  // Operands first used after the plane's first barrier: loaded at the start of the plane's own
  // body, ahead of the next plane's prefetch, so only one copy of them is ever live.
  struct Late {
    T phn;  // phi(m + 2P): enters the x queue after the plane
    T wy1, wz1;
    T pt, pr;
  };
  auto load_plane = [&](int m, Plane& L) {
    const int base = m * plane * S;
    bload_rec<T>(r_gp, rec, m * n1f * n2f * 4 * S, L.wxf, L.wy0, L.wz0, L.iw);
    if (MODE == 3) L.iw = T(0);
#pragma unroll
    for (int s = 0; s < NH; ++s) L.hal[s] = bload<T>(r_phi, hsrc[s], base);
  };
  auto load_late = [&](int m, Late& L) {
    const int base = m * plane * S;
    {
      // loaded unconditionally (a clamped plane, then a select): a load on only one side of a
      // branch makes the compiler's wait counting assume the worst and wait for every load
      const int mp = m + 2 * P < n0 ? m + 2 * P : n0 - 1;
      const T y = bload<T>(r_phi, col, mp * plane * S);
      L.phn = m + 2 * P < n0 ? y : T(0);
    }
    L.wy1 = bload<T>(r_w1, wy1_off, m * pl1 * S);
    L.wz1 = bload<T>(r_w2, wz1_off, m * pl2 * S);
    L.pt = bload<T>(r_pt, col, base);
    L.pr = MODE >= 2 ? bload<T>(r_prev, col, base) : T(0);
  };

  auto body = [&](auto rot, int m, const Plane& cur, const Late& late) __attribute__((always_inline)) {
    // queue entry i of this plane (i = 0 the oldest) lives in slot (SR + i) % 2P
    constexpr int SR = decltype(rot)::value;
#define QI(i) (((SR) + (i)) % (2 * P))

    // newest x-face F(m+P-1) and the x part of the divergence
    {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * qphi[QI(i)];
      qf[QI(2 * P - 1)] = f * cur.wxf;
    }
    T ax = 0;
#pragma unroll
    for (int i = 0; i < 2 * P; ++i) ax += c[i] * qf[QI(2 * P - 1 - i)];

    // stage plane m
    if (act) s_phi[(ty + R) * LZ + tz + R] = qphi[QI(0)];
#pragma unroll
    for (int s = 0; s < NH; ++s)
      if (hdst[s] >= 0) s_phi[hdst[s]] = cur.hal[s];
    __syncthreads();

    // weighted y and z faces
    if (act) {
      T f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(ty + i) * LZ + tz + R];
      s_yf[ty * YP + tz] = f * cur.wy0;
      f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(ty + R) * LZ + tz + i];
      s_zf[ty * ZP + tz] = f * cur.wz0;
      if (ty < R) {
        f = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(ty + TY + i) * LZ + tz + R];
        s_yf[(ty + TY) * YP + tz] = f * late.wy1;
      }
      if (tz < R) {
        f = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(ty + R) * LZ + tz + TZ + i];
        s_zf[ty * ZP + tz + TZ] = f * late.wz1;
      }
      // Tiles narrower than R (grids thinner than 2P-1 points) own further face rows/columns.
      for (int r = ty + 2 * TY; r < TY + R; r += TY) {
        f = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(r + i) * LZ + tz + R];
        const int ay = Y0 + r;
        const T w = (z < n2 && ay < n1f) ? w1[m * pl1 + ay * n2 + z] : T(0);
        s_yf[r * YP + tz] = f * w;
      }
      for (int q = tz + 2 * TZ; q < TZ + R; q += TZ) {
        f = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) f += c[i] * s_phi[(ty + R) * LZ + q + i];
        const int az = Z0 + q;
        const T w = (y < n1 && az < n2f) ? w2[m * pl2 + y * n2f + az] : T(0);
        s_zf[ty * ZP + q] = f * w;
      }
    }
    __syncthreads();

    if (own) {
      T ay = 0, az = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) ay += c[i] * s_yf[(ty + R - i) * YP + tz];
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) az += c[i] * s_zf[ty * ZP + tz + R - i];
      const T lap = (ax + ay) + az;
      const T pv = qphi[QI(0)];
      T val = MODE == 3 ? lap * T(0.5) + late.pt * late.pr : lap * T(0.5) * cur.iw + pv * late.pt;
      if (MODE == 1) val = (val - cc * pv) * ca;
      if (MODE == 2) val = (val - cc * pv) * ca - cb * late.pr;
      if (MODE == 4) {
        const T r = val - ev * pv;
        racc += (r * r) * late.pr;
      } else {
        bstore<T>(val, r_out, col, m * plane * S);
      }
    }

    // advance the x queues and the plane pipeline
    qphi[QI(0)] = late.phn;  // becomes logical index 2P-1 of the next plane
#undef QI
  };
  // Two plane buffers used in turn: the next plane's loads land in the other buffer and no register
  // with a load in flight is ever copied (a copy waits for the load). The prefetch past the segment's
  // last plane reloads that plane, so every iteration issues the same loads. Each plane first issues
  // its own late operands, then the next plane's prefetch: loads complete in issue order, so the wait
  // for the late operands at the end of the plane leaves the prefetch in flight. Keeping the late
  // operands out of the double buffer holds the float64 order-8 kernels at or below 128 VGPRs (four
  // waves per SIMD instead of three). The loop is unrolled over the x queues' period 2P, so the
  // queues rotate by renaming (the body's compile-time `rot`) instead of 4P-2 register moves per plane.
  static_assert(2 * P <= 12, "the unrolled plane loop covers queue periods up to 12");
  Plane pa, pb;
  load_plane(X0, pa);
  for (int m = X0; m < X1; m += 2 * P) {
    if constexpr (0 < 2 * P) {
      Late late;
      load_late(m + 0, late);
      load_plane(m + 1 < X1 ? m + 1 : X1 - 1, pb);
      body(std::integral_constant<int, 0>(), m + 0, pa, late);
    }
    if constexpr (1 < 2 * P) {
      if (m + 1 >= X1) break;
      Late late;
      load_late(m + 1, late);
      load_plane(m + 2 < X1 ? m + 2 : X1 - 1, pa);
      body(std::integral_constant<int, 1>(), m + 1, pb, late);
    }
    if constexpr (2 < 2 * P) {
      if (m + 2 >= X1) break;
      Late late;
      load_late(m + 2, late);
      load_plane(m + 3 < X1 ? m + 3 : X1 - 1, pb);
      body(std::integral_constant<int, 2>(), m + 2, pa, late);
    }
    if constexpr (3 < 2 * P) {
      if (m + 3 >= X1) break;
      Late late;
      load_late(m + 3, late);
      load_plane(m + 4 < X1 ? m + 4 : X1 - 1, pa);
      body(std::integral_constant<int, 3>(), m + 3, pb, late);
    }
    if constexpr (4 < 2 * P) {
      if (m + 4 >= X1) break;
      Late late;
      load_late(m + 4, late);
      load_plane(m + 5 < X1 ? m + 5 : X1 - 1, pb);
      body(std::integral_constant<int, 4>(), m + 4, pa, late);
    }
    if constexpr (5 < 2 * P) {
      if (m + 5 >= X1) break;
      Late late;
      load_late(m + 5, late);
      load_plane(m + 6 < X1 ? m + 6 : X1 - 1, pa);
      body(std::integral_constant<int, 5>(), m + 5, pb, late);
    }
    if constexpr (6 < 2 * P) {
      if (m + 6 >= X1) break;
      Late late;
      load_late(m + 6, late);
      load_plane(m + 7 < X1 ? m + 7 : X1 - 1, pb);
      body(std::integral_constant<int, 6>(), m + 6, pa, late);
    }
    if constexpr (7 < 2 * P) {
      if (m + 7 >= X1) break;
      Late late;
      load_late(m + 7, late);
      load_plane(m + 8 < X1 ? m + 8 : X1 - 1, pa);
      body(std::integral_constant<int, 7>(), m + 7, pb, late);
    }
    if constexpr (8 < 2 * P) {
      if (m + 8 >= X1) break;
      Late late;
      load_late(m + 8, late);
      load_plane(m + 9 < X1 ? m + 9 : X1 - 1, pb);
      body(std::integral_constant<int, 8>(), m + 8, pa, late);
    }
    if constexpr (9 < 2 * P) {
      if (m + 9 >= X1) break;
      Late late;
      load_late(m + 9, late);
      load_plane(m + 10 < X1 ? m + 10 : X1 - 1, pa);
      body(std::integral_constant<int, 9>(), m + 9, pb, late);
    }
    if constexpr (10 < 2 * P) {
      if (m + 10 >= X1) break;
      Late late;
      load_late(m + 10, late);
      load_plane(m + 11 < X1 ? m + 11 : X1 - 1, pb);
      body(std::integral_constant<int, 10>(), m + 10, pa, late);
    }
    if constexpr (11 < 2 * P) {
      if (m + 11 >= X1) break;
      Late late;
      load_late(m + 11, late);
      load_plane(m + 12 < X1 ? m + 12 : X1 - 1, pa);
      body(std::integral_constant<int, 11>(), m + 11, pb, late);
    }
  }
  if constexpr (MODE == 4) {
    // fixed-order block sum: a butterfly per wavefront, then the wavefronts in order
    T sum = racc;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    __syncthreads();  // every thread is done with the face tiles, which now hold the wave sums
    if ((t & 63) == 0) s_yf[t >> 6] = sum;
    __syncthreads();
    if (t == 0) {
      T total = s_yf[0];
      for (int w = 1; w < (nt + 63) / 64; ++w) total += s_yf[w];
      out[0] = total;
    }
  }
}

template <typename T, int P, int MODE, int TZ>
__global__ __launch_bounds__(NTMAX) void hdiv_kernel(
    const T* __restrict__ phi, const T* __restrict__ prev, T* __restrict__ out,
    const T* __restrict__ w0, const T* __restrict__ w1, const T* __restrict__ w2,
    const T* __restrict__ invw, const T* __restrict__ pot, const T* __restrict__ gp, int n0, int n1, int n2, int k,
    Taps<P> taps, const double* __restrict__ table, int step, March md) {
  __shared__ StencilLds<T, P, TZ> lds;
  const unsigned lin = xcd_tile(blockIdx.x, gridDim.x);
  const unsigned tile = lin / (unsigned)k;
  const int v = (int)(lin - tile * (unsigned)k);
  stencil_block<T, P, MODE, TZ>(phi, prev, MODE == 4 ? out + lin : out, w0, w1, w2, invw, pot, gp, n0, n1, n2, taps, table,
                                step, md, tile, v, lds);
}

// ---- the vector-column kernel: K1 and K2 (modes 0-2) at order 8 ----
//
// The same marching scheme with each thread owning CZ = 16 / sizeof(T) adjacent z-columns (two
// doubles or four floats) of a (TY, TZ) tile: every plane-tile, face-tile and global access of the
// plane loop is one 16-byte vector. The column kernel above moves 8-byte values through LDS, whose
// 8-byte reads run at half the bytes per clock of 16-byte reads on this part; a thread here reads
// 5 vectors for the z-stencil of its CZ columns where a column thread reads 8 values for one. Each
// output is computed with the operations of stencil_block in the same order (the z-stencil takes its
// operands from a window of whole vectors), so the two kernels give bitwise the same results.
template <typename T>
struct V16;
template <>
struct V16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <>
struct V16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

template <typename T>
__device__ __forceinline__ typename V16<T>::type vload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(typename V16<T>::type, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void vstore(typename V16<T>::type v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef __attribute__((ext_vector_type(4))) unsigned int u4;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, voff, soff, 0);
}

constexpr int VEC_TMAX_Y = 17;  // largest tile extent along y of the vector kernel

// LDS of one vector-kernel block. Columns are staged from z = Z0 - 8 (a whole number of vectors of
// left halo); every pitch is congruent to the tile width modulo one 256-byte bank row, so the lanes of
// a wave, numbered row-major over the tile's vectors, read consecutive 16-byte bank groups.
template <typename T, int TZ>
struct VecLds {
  static constexpr int HZ = 8;                          // left halo columns staged
  static constexpr int LZ = TZ + 256 / (int)sizeof(T);  // plane-tile pitch (>= TZ + 16)
  static constexpr int YP = TZ;                         // y-face pitch
  static constexpr int ZP = TZ + 256 / (int)sizeof(T);  // z-face pitch (>= TZ + 8)
  T phi[(VEC_TMAX_Y + 14) * LZ];
  T yf[(VEC_TMAX_Y + 7) * YP];
  T zf[VEC_TMAX_Y * ZP];
};

template <typename T, int MODE, int TZ>
__device__ __forceinline__ void stencil_block_vec(
    const T* __restrict__ phi, const T* __restrict__ prev, T* __restrict__ out,
    const T* __restrict__ w0, const T* __restrict__ w1, const T* __restrict__ w2,
    const T* __restrict__ invw, const T* __restrict__ pot, int n0, int n1, int n2,
    const Taps<4>& taps, const double* __restrict__ table, int step, const March& md, unsigned tile, int v,
    VecLds<T, TZ>& lds) {
  static_assert(MODE <= 2, "the vector kernel covers the apply and the Chebyshev steps");
  typedef typename V16<T>::type V;
  constexpr int P = 4, R = 7;                    // order 8: reach 2P - 1 = 7
  constexpr int CZ = V16<T>::n;
  constexpr int HZ = VecLds<T, TZ>::HZ;
  constexpr int LZ = VecLds<T, TZ>::LZ;
  constexpr int YP = VecLds<T, TZ>::YP;
  constexpr int ZP = VecLds<T, TZ>::ZP;
  constexpr int TZV = TZ / CZ;                   // vectors per tile row
  constexpr int HL = HZ / CZ;                    // left halo vectors per row
  constexpr int HRV = (R + CZ - 1) / CZ;         // right halo vectors per row (= threads owning edge z-faces)
  constexpr int NW = (8 + CZ + CZ - 1) / CZ;     // vectors of a z-face window (columns c - 8 .. c + CZ - 1)
  constexpr int NF = (R + CZ + CZ - 1) / CZ;     // vectors of a z-divergence window (faces c .. c + CZ + R - 1)
  static_assert(TZ % CZ == 0 && TZV >= HRV, "tile width must hold whole vectors and the edge faces");
  T* s_phi = lds.phi;
  T* s_yf = lds.yf;
  T* s_zf = lds.zf;

  const int TY = md.ty;
  const int bz = (int)(tile % md.tiles_z);
  const int bs = (int)((tile / md.tiles_z) % md.nseg);
  const int by = (int)(tile / ((unsigned)md.tiles_z * md.nseg));
  const int X0 = (int)((long)bs * n0 / md.nseg), X1 = (int)((long)(bs + 1) * n0 / md.nseg);
  const int Y0 = by * TY, Z0 = bz * TZ;
  const int t = threadIdx.x, nt = blockDim.x;
  const bool act = t < TY * TZV;
  const int ty = act ? t / TZV : 0, pz = act ? t - (t / TZV) * TZV : 0;
  const int y = Y0 + ty, z = Z0 + CZ * pz;  // this thread's columns are z .. z + CZ - 1
  const int nz = act && y < n1 ? (n2 - z < CZ ? n2 - z : CZ) : 0;  // how many of them are grid points
  constexpr int S = (int)sizeof(T);
  constexpr int NONE = (int)0x80000000;
  const int plane = n1 * n2;
  const int n1f = n1 + 2 * P, n2f = n2 + 2 * P;
  const int pl1 = n1f * n2, pl2 = n1 * n2f;
  const long vbase = (long)v * plane * n0;
  const int vbytes = plane * n0 * S;
  const auto r_phi = make_rsrc(phi + vbase, vbytes);
  const auto r_prev = make_rsrc(MODE == 2 ? prev + vbase : phi, vbytes);
  const auto r_out = make_rsrc(out + vbase, vbytes);
  const auto r_w0 = make_rsrc(w0, (n0 + 2 * P) * plane * S);
  const auto r_w1 = make_rsrc(w1, n0 * pl1 * S);
  const auto r_w2 = make_rsrc(w2, n0 * pl2 * S);
  const auto r_iw = make_rsrc(invw, n0 * plane * S);
  const auto r_pt = make_rsrc(pot, n0 * plane * S);
  // A vector whose first column is inside the grid is loaded whole; its components past n2 (the
  // next row's values) are zeroed where they are staged as phi and otherwise only reach outputs
  // that are never stored.
  const int col = nz > 0 ? (y * n2 + z) * S : NONE;

  T c[2 * P];
#pragma unroll
  for (int i = 0; i < 2 * P; ++i) c[i] = (T)taps.c[i];
  T cc = 0, ca = 0, cb = 0;
  if (MODE == 1 || MODE == 2) {
    cc = (T)table[3 * step];
    ca = (T)table[3 * step + 1];
    cb = (T)table[3 * step + 2];
  }
  auto keep = [&](V a, int valid) {  // zero the components from `valid` on
#pragma unroll
    for (int j = 0; j < CZ; ++j) a[j] = j < valid ? a[j] : T(0);
    return a;
  };

  // Halo vectors of this thread (the same in every plane): source byte offset (NONE = outside the
  // grid), LDS destination (-1 = none) and how many leading components are grid points.
  const int nhy = 2 * R * TZV, nhalo = nhy + (HL + HRV) * TY;
  int hsrc[MAXH], hdst[MAXH], hval[MAXH];
#pragma unroll
  for (int s = 0; s < MAXH; ++s) {
    const int h = t + s * nt;
    hsrc[s] = NONE;
    hdst[s] = -1;
    hval[s] = 0;
    if (h < nhy) {
      const int hr = h / TZV, hc = h - (h / TZV) * TZV;
      const int rl = hr < R ? hr : TY + hr;
      const int gy = Y0 - R + rl, gz = Z0 + CZ * hc;
      hdst[s] = rl * LZ + HZ + CZ * hc;
      if (gy >= 0 && gy < n1 && gz < n2) {
        hsrc[s] = (gy * n2 + gz) * S;
        hval[s] = n2 - gz;
      }
    } else if (h < nhalo) {
      const int hh = h - nhy;
      const int r = hh / (HL + HRV), hc = hh - (hh / (HL + HRV)) * (HL + HRV);
      const int cl = hc < HL ? CZ * hc : HZ + TZ + CZ * (hc - HL);  // LDS column
      const int gy = Y0 + r, gz = Z0 - HZ + cl;
      hdst[s] = (R + r) * LZ + cl;
      if (gy < n1 && gz >= 0 && gz < n2) {
        hsrc[s] = (gy * n2 + gz) * S;
        hval[s] = n2 - gz;
      }
    }
  }
  const bool zin = act && z < n2, yin = act && y < n1;
  const int ay1 = Y0 + ty + TY;       // the edge y-face row (extended index) of threads ty < R
  const int az1 = Z0 + TZ + CZ * pz;  // the edge z-face columns (extended index) of threads pz < HRV
  const int wy_off = (zin && y < n1f) ? (y * n2 + z) * S : NONE;
  const int wz_off = (yin && z < n2f) ? (y * n2f + z) * S : NONE;
  const int wy1_off = (zin && ty < R && ay1 < n1f) ? (ay1 * n2 + z) * S : NONE;
  const int wz1_off = (yin && pz < HRV && az1 < n2f) ? (y * n2f + az1) * S : NONE;

  // ---- x warm-up: phi(X0-7 .. X0+7) and the faces F(X0-4 .. X0+2) ----
  V qphi[2 * P], qf[2 * P];
  {
    V pen[4 * P - 1];
#pragma unroll
    for (int q = 0; q < 4 * P - 1; ++q) {
      const int gx = X0 - 2 * P + 1 + q;  // block-uniform
      pen[q] = (gx >= 0 && gx < n0) ? vload<T>(r_phi, col, gx * plane * S) : V(0);
    }
#pragma unroll
    for (int j = 0; j < 2 * P - 1; ++j) {
      V f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * pen[j + i];
      qf[j] = f * vload<T>(r_w0, col, (X0 + j) * plane * S);
    }
    qf[2 * P - 1] = 0;
#pragma unroll
    for (int j = 0; j < 2 * P; ++j) qphi[j] = pen[2 * P - 1 + j];
  }

  // Operands of a plane: those used before its first barrier loaded one plane ahead into two
  // alternating buffers; the rest at the start of the plane's own body (as in stencil_block).
  struct Plane {
    V wxf;        // w0 at the plane's newest x-face
    V hal[MAXH];  // halo vectors
  };
  struct Late {
    V wy0, wz0, iw;
    V phn;  // phi(m + 8)
    V wy1, wz1;
    V pt, pr;
  };
  auto load_plane = [&](int m, Plane& L) {
    const int base = m * plane * S;
    L.wxf = vload<T>(r_w0, col, (m + 2 * P - 1) * plane * S);
#pragma unroll
    for (int s = 0; s < MAXH; ++s) L.hal[s] = vload<T>(r_phi, hsrc[s], base);
  };
  auto load_late = [&](int m, Late& L) {
    const int base = m * plane * S;
    L.wy0 = vload<T>(r_w1, wy_off, m * pl1 * S);
    L.wz0 = vload<T>(r_w2, wz_off, m * pl2 * S);
    L.iw = vload<T>(r_iw, col, base);
    {
      const int mp = m + 2 * P < n0 ? m + 2 * P : n0 - 1;
      const V yv = vload<T>(r_phi, col, mp * plane * S);
      L.phn = m + 2 * P < n0 ? yv : V(0);
    }
    L.wy1 = vload<T>(r_w1, wy1_off, m * pl1 * S);
    L.wz1 = vload<T>(r_w2, wz1_off, m * pl2 * S);
    L.pt = vload<T>(r_pt, col, base);
    L.pr = MODE == 2 ? vload<T>(r_prev, col, base) : V(0);
  };
  auto lv = [](const T* p) { return *reinterpret_cast<const V*>(p); };
  auto sv = [](T* p, V a) { *reinterpret_cast<V*>(p) = a; };
  // the z-faces of CZ consecutive face columns from the plane-tile window starting at LDS column w0c
  auto zfaces = [&](int row, int w0c) {
    T win[NW * CZ];
#pragma unroll
    for (int u = 0; u < NW; ++u) {
      const V a = lv(&s_phi[row * LZ + w0c + CZ * u]);
#pragma unroll
      for (int j = 0; j < CZ; ++j) win[CZ * u + j] = a[j];
    }
    V g;
#pragma unroll
    for (int j = 0; j < CZ; ++j) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) s += c[i] * win[1 + j + i];
      g[j] = s;
    }
    return g;
  };

  auto body = [&](auto rot, int m, const Plane& cur, const Late& late) __attribute__((always_inline)) {
    constexpr int SR = decltype(rot)::value;
#define QI(i) (((SR) + (i)) % (2 * P))
    {
      V f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * qphi[QI(i)];
      qf[QI(2 * P - 1)] = f * cur.wxf;
    }
    V ax = 0;
#pragma unroll
    for (int i = 0; i < 2 * P; ++i) ax += c[i] * qf[QI(2 * P - 1 - i)];

    const V pv = qphi[QI(0)];
    if (act) sv(&s_phi[(ty + R) * LZ + HZ + CZ * pz], keep(pv, nz));
#pragma unroll
    for (int s = 0; s < MAXH; ++s)
      if (hdst[s] >= 0) sv(&s_phi[hdst[s]], keep(cur.hal[s], hval[s]));
    __syncthreads();

    if (act) {
      // y-face F_y(y - 4) of the thread's columns, from rows y - 7 .. y
      V f = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) f += c[i] * lv(&s_phi[(ty + i) * LZ + HZ + CZ * pz]);
      sv(&s_yf[ty * YP + CZ * pz], f * late.wy0);
      // z-faces F_z(z + j - 4), from columns z + j - 7 .. z + j
      sv(&s_zf[ty * ZP + CZ * pz], zfaces(ty + R, HZ - 8 + CZ * pz) * late.wz0);
      if (ty < R) {
        V e = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) e += c[i] * lv(&s_phi[(ty + TY + i) * LZ + HZ + CZ * pz]);
        sv(&s_yf[(ty + TY) * YP + CZ * pz], e * late.wy1);
      }
      if (pz < HRV) sv(&s_zf[ty * ZP + TZ + CZ * pz], zfaces(ty + R, HZ - 8 + TZ + CZ * pz) * late.wz1);
      // tiles narrower than R rows (grids thinner than 7 points) own further face rows
      for (int r = ty + 2 * TY; r < TY + R; r += TY) {
        V e = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) e += c[i] * lv(&s_phi[(r + i) * LZ + HZ + CZ * pz]);
        const int ay = Y0 + r;
        const V w = vload<T>(r_w1, (zin && ay < n1f) ? (ay * n2 + z) * S : NONE, m * pl1 * S);
        sv(&s_yf[r * YP + CZ * pz], e * w);
      }
    }
    __syncthreads();

    if (nz > 0) {
      V ay = 0;
#pragma unroll
      for (int i = 0; i < 2 * P; ++i) ay += c[i] * lv(&s_yf[(ty + R - i) * YP + CZ * pz]);
      // az_j = sum_i c_i F_z(z + j + 3 - i): face columns CZ pz + j + R - i
      T win[NF * CZ];
#pragma unroll
      for (int u = 0; u < NF; ++u) {
        const V a = lv(&s_zf[ty * ZP + CZ * pz + CZ * u]);
#pragma unroll
        for (int j = 0; j < CZ; ++j) win[CZ * u + j] = a[j];
      }
      V az;
#pragma unroll
      for (int j = 0; j < CZ; ++j) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < 2 * P; ++i) s += c[i] * win[j + R - i];
        az[j] = s;
      }
      const V lap = (ax + ay) + az;
      V val = lap * T(0.5) * late.iw + pv * late.pt;
      if (MODE == 1) val = (val - cc * pv) * ca;
      if (MODE == 2) val = (val - cc * pv) * ca - cb * late.pr;
      if (nz >= CZ) {
        vstore<T>(val, r_out, col, m * plane * S);
      } else {
#pragma unroll
        for (int j = 0; j < CZ; ++j)
          if (j < nz) bstore<T>(val[j], r_out, col + j * S, m * plane * S);
      }
    }
    qphi[QI(0)] = late.phn;
#undef QI
  };

  // the plane loop of stencil_block: two prefetch buffers, unrolled over the x queues' period
  Plane pa, pb;
  load_plane(X0, pa);
  for (int m = X0; m < X1; m += 2 * P) {
    {
      Late late;
      load_late(m + 0, late);
      load_plane(m + 1 < X1 ? m + 1 : X1 - 1, pb);
      body(std::integral_constant<int, 0>(), m + 0, pa, late);
    }
    if (m + 1 >= X1) break;
    {
      Late late;
      load_late(m + 1, late);
      load_plane(m + 2 < X1 ? m + 2 : X1 - 1, pa);
      body(std::integral_constant<int, 1>(), m + 1, pb, late);
    }
    if (m + 2 >= X1) break;
    {
      Late late;
      load_late(m + 2, late);
      load_plane(m + 3 < X1 ? m + 3 : X1 - 1, pb);
      body(std::integral_constant<int, 2>(), m + 2, pa, late);
    }
    if (m + 3 >= X1) break;
    {
      Late late;
      load_late(m + 3, late);
      load_plane(m + 4 < X1 ? m + 4 : X1 - 1, pa);
      body(std::integral_constant<int, 3>(), m + 3, pb, late);
    }
    if (m + 4 >= X1) break;
    {
      Late late;
      load_late(m + 4, late);
      load_plane(m + 5 < X1 ? m + 5 : X1 - 1, pb);
      body(std::integral_constant<int, 4>(), m + 4, pa, late);
    }
    if (m + 5 >= X1) break;
    {
      Late late;
      load_late(m + 5, late);
      load_plane(m + 6 < X1 ? m + 6 : X1 - 1, pa);
      body(std::integral_constant<int, 5>(), m + 5, pb, late);
    }
    if (m + 6 >= X1) break;
    {
      Late late;
      load_late(m + 6, late);
      load_plane(m + 7 < X1 ? m + 7 : X1 - 1, pb);
      body(std::integral_constant<int, 6>(), m + 6, pa, late);
    }
    if (m + 7 >= X1) break;
    {
      Late late;
      load_late(m + 7, late);
      load_plane(m + 8 < X1 ? m + 8 : X1 - 1, pa);
      body(std::integral_constant<int, 7>(), m + 7, pb, late);
    }
  }
}

template <typename T, int MODE, int TZ>
__global__ __launch_bounds__(NTMAX) void hdiv_vec_kernel(
    const T* __restrict__ phi, const T* __restrict__ prev, T* __restrict__ out,
    const T* __restrict__ w0, const T* __restrict__ w1, const T* __restrict__ w2,
    const T* __restrict__ invw, const T* __restrict__ pot, int n0, int n1, int n2, int k,
    Taps<4> taps, const double* __restrict__ table, int step, March md) {
  __shared__ VecLds<T, TZ> lds;
  const unsigned lin = xcd_tile(blockIdx.x, gridDim.x);
  const unsigned tile = lin / (unsigned)k;
  const int v = (int)(lin - tile * (unsigned)k);
  stencil_block_vec<T, MODE, TZ>(phi, prev, out, w0, w1, w2, invw, pot, n0, n1, n2, taps, table, step, md, tile, v,
                                 lds);
}

// ---- batched launches: many grids (one descriptor each) in one launch ----

// One grid of a batch. The filter rotates through out / work1 / work2 exactly as the single-grid
// filter does, so the batched steps are bitwise the single-grid steps.
template <typename T>
struct BatchGrid {
  const T* x;
  T* out;
  T* work1;
  T* work2;
  const T *w0, *w1, *w2, *invw, *pot;
  const T* gp;          // the grid's geometry records (cdft_hdiv_pack_*)
  const double* table;  // (degree, 3) Chebyshev rows of this grid
  double taps[8];
  int n0, n1, n2, k;
  March md;
  unsigned block_start;  // first block of this grid in the launch
  unsigned blocks;
};

template <typename T, int P, int MODE, int TZ>
__global__ __launch_bounds__(NTMAX) void hdiv_batched_kernel(const BatchGrid<T>* __restrict__ grids, int n_grids,
                                                             int step, int degree) {
  __shared__ StencilLds<T, P, TZ> lds;
  const unsigned lin = xcd_tile(blockIdx.x, gridDim.x);
  int lo = 0, hi = n_grids - 1;  // the grid whose block range holds lin (block-uniform search)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (grids[mid].block_start <= lin) lo = mid; else hi = mid - 1;
  }
  const BatchGrid<T>& g = grids[lo];
  const unsigned local = lin - g.block_start;
  const unsigned tile = local / (unsigned)g.k;
  const int v = (int)(local - tile * (unsigned)g.k);
  Taps<P> taps;
#pragma unroll
  for (int i = 0; i < 2 * P; ++i) taps.c[i] = g.taps[i];
  const T* y = g.x;
  const T* prev = nullptr;
  T* dst = g.out;
  if (MODE != 0) {
    // the rotation out / work1 / work2 by selects (an array indexed at run time would live in scratch)
    auto buf = [&](int i) { return i == 0 ? g.out : (i == 1 ? g.work1 : g.work2); };
    const int offset = ((-(degree - 1)) % 3 + 3) % 3;
    dst = buf((step + offset) % 3);
    y = step == 0 ? g.x : buf((step - 1 + offset) % 3);
    prev = step <= 1 ? g.x : buf((step - 2 + offset) % 3);
  }
  stencil_block<T, P, MODE, TZ>(y, prev, dst, g.w0, g.w1, g.w2, g.invw, g.pot, g.gp, g.n0, g.n1, g.n2, taps, g.table, step,
                                g.md, tile, v, lds);
}

// The order-8 batch on the vector kernel: the same grid search and buffer rotation, then exactly the
// single-grid vector kernel's block (bitwise its steps, and so the column kernel's).
template <typename T, int MODE, int TZ>
__global__ __launch_bounds__(NTMAX) void hdiv_batched_vec_kernel(const BatchGrid<T>* __restrict__ grids, int n_grids,
                                                                 int step, int degree) {
  __shared__ VecLds<T, TZ> lds;
  const unsigned lin = xcd_tile(blockIdx.x, gridDim.x);
  int lo = 0, hi = n_grids - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (grids[mid].block_start <= lin) lo = mid; else hi = mid - 1;
  }
  const BatchGrid<T>& g = grids[lo];
  const unsigned local = lin - g.block_start;
  const unsigned tile = local / (unsigned)g.k;
  const int v = (int)(local - tile * (unsigned)g.k);
  Taps<4> taps;
#pragma unroll
  for (int i = 0; i < 8; ++i) taps.c[i] = g.taps[i];
  const T* y = g.x;
  const T* prev = nullptr;
  T* dst = g.out;
  if (MODE != 0) {
    auto buf = [&](int i) { return i == 0 ? g.out : (i == 1 ? g.work1 : g.work2); };
    const int offset = ((-(degree - 1)) % 3 + 3) % 3;
    dst = buf((step + offset) % 3);
    y = step == 0 ? g.x : buf((step - 1 + offset) % 3);
    prev = step <= 1 ? g.x : buf((step - 2 + offset) % 3);
  }
  stencil_block_vec<T, MODE, TZ>(y, prev, dst, g.w0, g.w1, g.w2, g.invw, g.pot, g.n0, g.n1, g.n2, taps, g.table, step,
                                 g.md, tile, v, lds);
}

// The per-point geometry records the stencil kernels read each plane: for plane m and column (y, z) of
// the face-extended rows (y < n1 + 2P, z < n2 + 2P), (w0 at extended plane m + 2P - 1, w1 at (m, y, z),
// w2 at (m, y, z), invw), each part zero where the separate array has no such entry for this column.
template <typename T>
__global__ __launch_bounds__(256) void pack_kernel(const T* __restrict__ w0, const T* __restrict__ w1,
                                                   const T* __restrict__ w2, const T* __restrict__ invw, int n0,
                                                   int n1, int n2, int two_p, T* __restrict__ gp) {
  const int n1f = n1 + two_p, n2f = n2 + two_p;
  const long total = (long)n0 * n1f * n2f;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long per = (long)n1f * n2f;
  const int m = (int)(idx / per);
  const int r = (int)(idx - (long)m * per);
  const int y = r / n2f, z = r - (r / n2f) * n2f;
  const long plane = (long)n1 * n2;
  const bool own = y < n1 && z < n2;
  T* o = gp + 4 * idx;
  o[0] = own ? w0[(long)(m + two_p - 1) * plane + (long)y * n2 + z] : T(0);
  o[1] = z < n2 ? w1[(long)m * n1f * n2 + (long)y * n2 + z] : T(0);
  o[2] = y < n1 ? w2[(long)m * n1 * n2f + (long)y * n2f + z] : T(0);
  o[3] = own ? invw[(long)m * plane + (long)y * n2 + z] : T(0);
}

template <typename T>
static int pack_impl(const T* w0, const T* w1, const T* w2, const T* invw, int n0, int n1, int n2, int order, T* gp,
                     hipStream_t s) {
  if (!w0 || !w1 || !w2 || !invw || !gp || n0 < 1 || n1 < 1 || n2 < 1)
    return fail(CDFT_E_ARG, "hdiv_pack: bad arguments");
  if (order != 2 && order != 4 && order != 6 && order != 8 && order != 12)
    return fail(CDFT_E_ARG, "hdiv_pack: bad order %d", order);
  const long total = (long)n0 * (n1 + order) * (n2 + order);
  hipLaunchKernelGGL(pack_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, w0, w1, w2, invw, n0, n1,
                     n2, order, gp);
  return check_launch("hdiv_pack");
}

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Segment length override (0 = automatic); set by cdft_hdiv_set_segment.
static int g_segment = 0;
// Tile extent overrides (0 = automatic): the largest tile extent along y (1..TMAX; the tiles then
// split n1 evenly) and the tile width along z (13 or 16); set by cdft_hdiv_set_tile. No tiling
// changes an output bit: every point's sums take the same operands in the same order.
static int g_tile_y = 0, g_tile_z = 0;
// Compute units of the current device, read once per device: the unit of the fill rule below.
// 256 (MI355X) when no device is visible, as for the host-only launch-shape query.
static int device_cus() {
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();  // leave no error behind for the next launch check
    return 256;
  }
  int cus = cached[dev].load(std::memory_order_relaxed);
  if (cus > 0) return cus;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
    (void)hipGetLastError();
    cus = 256;
  }
  cached[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

// Segment count of the fill rule. The dispatcher deals a launch's blocks round-robin over the CUs
// and, up to 16 waves per CU (four per SIMD, the order-8 float64 kernels' occupancy), runs them all
// at once; the launch ends when the most loaded CU ends. Per-block timestamps show a block's life
// growing with the blocks beside it on its CU (65^3, k = 5: 14.1 us beside two others, 16.4 us
// beside three; 81^3: 28.5 us at four per CU, 32.1 us at five), so a block count just above a
// multiple of the CU count is the worst case: at 81^3, k = 5, 1,050 blocks (26 CUs with five)
// take 39.7 us and 1,260 blocks take 35.6 us. For each fill b = 1..5 blocks per CU the rule takes
// the most segments that fit (segments of at least 4 planes) and scores them by (4 + planes per
// segment) x g(b): 4 planes stand for the segment's x warm-up, and g is the relative block life
// at b per CU fitted to launch sweeps at 65^3 (k = 5, 8), 81^3 and 101 x 81 x 81 (float64 and
// float32, tilings 8-16 x 13-16, 1-16 segments). The choice is the best or within 3% of the best
// segment count measured on each of those shapes and on 57^3.
static int fill_segments(int n0, long columns, int threads, int waves_per_cu = 16) {
  static const double g[6] = {0.0, 0.57, 0.68, 0.90, 1.0, 1.29};
  const long cus = device_cus();
  const int cap = std::min(5, std::max(1, waves_per_cu / std::max(1, threads / 64)));
  const int most = ceil_div(n0, 4);
  int best = 1;
  double best_cost = 1e300;
  for (int b = 1; b <= cap; ++b) {
    const int nseg = (int)std::max(1L, std::min<long>((long)b * cus / columns, most));
    const long per_cu = (columns * nseg + cus - 1) / cus;  // the most loaded CU
    const double cost = (4.0 + ceil_div(n0, nseg)) * g[std::min<long>(std::max(per_cu, 1L), 5)];
    if (cost < best_cost) {
      best_cost = cost;
      best = nseg;
    }
  }
  return best;
}

// The column kernel's z tiles are 13 or 16 wide (its LDS pitches are built for those): the one that
// leaves fewer idle columns (65 = 5 x 13, 64 = 4 x 16).
static int column_tile_width(int n2) {
  const int idle13 = ceil_div(n2, 13) * 13 - n2, idle16 = ceil_div(n2, 16) * 16 - n2;
  return idle13 < idle16 ? 13 : 16;
}

// tz: the tile width (0 = the override of cdft_hdiv_set_tile, else column_tile_width).
static March march_dims(int n0, int n1, int n2, int k, int reach, int tz = 0) {
  March d;
  // (a vector-only width in the override is for batch plans: the column kernel takes its own choice)
  d.tz = tz ? tz : (g_tile_z == 13 || g_tile_z == 16) ? g_tile_z : column_tile_width(n2);
  d.tiles_z = ceil_div(n2, d.tz);
  d.tiles_y = ceil_div(n1, g_tile_y ? g_tile_y : TMAX);
  d.ty = ceil_div(n1, d.tiles_y);
  const int cover = d.ty * d.tz;
  const int halo = ceil_div(2 * reach * (d.ty + d.tz), halo_slots(reach));  // every halo value has a slot
  int th = cover > halo ? cover : halo;
  d.threads = ceil_div(th, 64) * 64;
  const int per_plane = d.tiles_y * d.tiles_z;
  d.nseg = g_segment > 0 ? ceil_div(n0, g_segment) : fill_segments(n0, (long)per_plane * k, d.threads);
  if (d.nseg < 1) d.nseg = 1;
  if (d.nseg > n0) d.nseg = n0;
  d.seg = ceil_div(n0, d.nseg);
  d.tiles = (unsigned)per_plane * d.nseg;
  return d;
}

// ---- launch shape of the vector kernel ----

// Which kernel the order-8 apply and Chebyshev steps use: 0 = the vector kernel (default), 1 = the
// column kernel (set by cdft_hdiv_set_kernel; both give bitwise the same outputs).
static std::atomic<int> g_kernel{0};

// Tile widths the vector kernel is built for, by element size.
static const int VEC_TZ_F64[] = {16, 20, 22, 24, 28};
static const int VEC_TZ_F32[] = {16, 20, 24, 28};

static bool vec_width_ok(int tz, int elem) {
  if (elem == 8) return std::find(std::begin(VEC_TZ_F64), std::end(VEC_TZ_F64), tz) != std::end(VEC_TZ_F64);
  return std::find(std::begin(VEC_TZ_F32), std::end(VEC_TZ_F32), tz) != std::end(VEC_TZ_F32);
}

// Per-shape launch settings (tile rows, tile width, planes per segment) put in force by
// cdft_hdiv_vec_set, typically from a timing sweep on the device: a launch whose grid, block width and
// element size match an entry uses it. A small table under a lock; the oldest entry makes room.
struct VecSetting {
  int n0, n1, n2, k, elem, ty, tz, seg;
};
static std::mutex g_vec_mu;
static VecSetting g_vec_table[64];
static int g_vec_count = 0, g_vec_next = 0;

static bool vec_lookup(int n0, int n1, int n2, int k, int elem, VecSetting& out) {
  std::lock_guard<std::mutex> lock(g_vec_mu);
  for (int i = 0; i < g_vec_count; ++i) {
    const VecSetting& e = g_vec_table[i];
    if (e.n0 == n0 && e.n1 == n1 && e.n2 == n2 && e.k == k && e.elem == elem) {
      out = e;
      return true;
    }
  }
  return false;
}

// The vector kernel's launch: by default the tile width with the fewest idle columns (ties: the
// wider), the largest tile rows up to VEC_TMAX_Y that split n1 evenly, and the segment count of the
// CU-fill rule at the kernel's occupancy (three waves per SIMD in float64, two in float32).
// The vector kernel's launch for tile width tz, tiles of at most ty_max rows and segments of seg
// planes (0: the CU-fill rule).
static March vec_dims_with(int n0, int n1, int n2, int k, int elem, int tz, int ty_max, int seg) {
  const int cz = 16 / elem;
  March d;
  d.tz = tz;
  d.tiles_z = ceil_div(n2, d.tz);
  d.tiles_y = ceil_div(n1, ty_max);
  d.ty = ceil_div(n1, d.tiles_y);
  const int tzv = d.tz / cz;
  const int reach = 7;
  const int halo = ceil_div(2 * reach * tzv + (8 / cz + ceil_div(reach, cz)) * d.ty, MAXH);
  const int cover = d.ty * tzv;
  d.threads = ceil_div(std::max(cover, halo), 64) * 64;
  const int per_plane = d.tiles_y * d.tiles_z;
  if (seg > 0)
    d.nseg = ceil_div(n0, seg);
  else
    d.nseg = fill_segments(n0, (long)per_plane * k, d.threads, elem == 8 ? 12 : 8);
  d.nseg = std::max(1, std::min(d.nseg, n0));
  d.seg = ceil_div(n0, d.nseg);
  d.tiles = (unsigned)per_plane * d.nseg;
  return d;
}

// The vector width with the fewest idle columns over n2 (ties: the wider).
static int vec_best_width(int n2, int elem) {
  int best = 0, idle = 1 << 30;
  const int* w = elem == 8 ? VEC_TZ_F64 : VEC_TZ_F32;
  const int nw = elem == 8 ? (int)std::size(VEC_TZ_F64) : (int)std::size(VEC_TZ_F32);
  for (int i = 0; i < nw; ++i) {
    const int id = ceil_div(n2, w[i]) * w[i] - n2;
    if (id <= idle) {
      idle = id;
      best = w[i];
    }
  }
  return best;
}

static March vec_dims(int n0, int n1, int n2, int k, int elem) {
  VecSetting set{};
  if (vec_lookup(n0, n1, n2, k, elem, set)) return vec_dims_with(n0, n1, n2, k, elem, set.tz, set.ty, set.seg);
  return vec_dims_with(n0, n1, n2, k, elem, vec_best_width(n2, elem), VEC_TMAX_Y, 0);
}

template <typename T, int MODE, int TZ>
static int launch_vec_tz(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2, const T* invw,
                         const T* pot, int k, int n0, int n1, int n2, const Taps<4>& taps, const double* table,
                         int step, const March& d, hipStream_t stream) {
  const unsigned long blocks = (unsigned long)d.tiles * (unsigned long)k;
  if (blocks > 0x7fffffffUL) return fail(CDFT_E_ARG, "hdiv: too many blocks");
  hipLaunchKernelGGL((hdiv_vec_kernel<T, MODE, TZ>), dim3((unsigned)blocks), dim3(d.threads), 0, stream, phi, prev,
                     out, w0, w1, w2, invw, pot, n0, n1, n2, k, taps, table, step, d);
  return check_launch("hdiv_vec_kernel");
}

template <typename T, int MODE>
static int launch_vec(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2, const T* invw,
                      const T* pot, int k, int n0, int n1, int n2, const Taps<4>& taps, const double* table, int step,
                      hipStream_t stream) {
  const March d = vec_dims(n0, n1, n2, k, (int)sizeof(T));
  if (d.ty > VEC_TMAX_Y || d.threads > NTMAX) return fail(CDFT_E_ARG, "hdiv: vector tile needs %d threads", d.threads);
#define CDFT_VEC_CASE(W) \
  case W: return launch_vec_tz<T, MODE, W>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, taps, table, step, d, stream);
  if constexpr (sizeof(T) == 8) {
    switch (d.tz) {
      CDFT_VEC_CASE(16)
      CDFT_VEC_CASE(20)
      CDFT_VEC_CASE(22)
      CDFT_VEC_CASE(24)
      CDFT_VEC_CASE(28)
    }
  } else {
    switch (d.tz) {
      CDFT_VEC_CASE(16)
      CDFT_VEC_CASE(20)
      CDFT_VEC_CASE(24)
      CDFT_VEC_CASE(28)
    }
  }
#undef CDFT_VEC_CASE
  return fail(CDFT_E_ARG, "hdiv: no vector kernel of tile width %d", d.tz);
}

template <typename T, int P, int MODE>
static int launch_p(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2,
                    const T* invw, const T* pot, const T* gp, int k, int n0, int n1, int n2, const double* coef,
                    const double* table, int step, hipStream_t stream) {
  Taps<P> taps;
  for (int i = 0; i < 2 * P; ++i) taps.c[i] = coef[i];
  if constexpr (P == 4 && MODE <= 2) {
    if (g_kernel.load(std::memory_order_relaxed) == 0) {
      if ((long)(n0 + 2 * P) * (n1 + 2 * P) * (n2 + 2 * P) * (long)sizeof(T) >= 0x7fffffffL)
        return fail(CDFT_E_ARG, "hdiv: grid %d x %d x %d too large for one launch", n0, n1, n2);
      return launch_vec<T, MODE>(phi, prev, out, w0, w1, w2, invw, pot, k, n0, n1, n2, taps, table, step, stream);
    }
  }
  const March d = march_dims(n0, n1, n2, k, 2 * P - 1);
  if (d.threads > NTMAX) return fail(CDFT_E_ARG, "hdiv: tile needs %d threads", d.threads);
  // Offsets inside one vector and inside the extended face-weight arrays are 32-bit.
  if ((long)(n0 + 2 * P) * (n1 + 2 * P) * (n2 + 2 * P) * 4 * (long)sizeof(T) >= 0x7fffffffL)
    return fail(CDFT_E_ARG, "hdiv: grid %d x %d x %d too large for one launch", n0, n1, n2);
  const unsigned long blocks = (unsigned long)d.tiles * (unsigned long)k;
  if (blocks > 0x7fffffffUL) return fail(CDFT_E_ARG, "hdiv: too many blocks");
  const dim3 grid((unsigned)blocks), block(d.threads);
  if (d.tz == 13)
    hipLaunchKernelGGL((hdiv_kernel<T, P, MODE, 13>), grid, block, 0, stream, phi, prev, out, w0, w1, w2,
                       invw, pot, gp, n0, n1, n2, k, taps, table, step, d);
  else
    hipLaunchKernelGGL((hdiv_kernel<T, P, MODE, 16>), grid, block, 0, stream, phi, prev, out, w0, w1, w2,
                       invw, pot, gp, n0, n1, n2, k, taps, table, step, d);
  return check_launch("hdiv_kernel");
}

template <typename T, int MODE>
static int launch(const T* phi, const T* prev, T* out, const T* w0, const T* w1, const T* w2,
                  const T* invw, const T* pot, const T* gp, int k, int n0, int n1, int n2, const double* coef,
                  int order, const double* table, int step, hipStream_t stream) {
  switch (order) {
    case 2: return launch_p<T, 1, MODE>(phi, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, table, step, stream);
    case 4: return launch_p<T, 2, MODE>(phi, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, table, step, stream);
    case 6: return launch_p<T, 3, MODE>(phi, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, table, step, stream);
    case 8: return launch_p<T, 4, MODE>(phi, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, table, step, stream);
    case 12:
      if constexpr (std::is_same_v<T, double>)
        return launch_p<T, 6, MODE>(phi, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, table, step, stream);
      else
        return fail(CDFT_E_ARG, "stencil order 12 runs in float64 only");
    default: return fail(CDFT_E_ARG, "unsupported stencil order %d (2, 4, 6, 8 or 12)", order);
  }
}

static int validate(const void* phi, const void* out, const void* w0, const void* w1,
                    const void* w2, const void* invw, const void* pot, const void* gp, int k, int n0, int n1,
                    int n2, const double* coef, int order) {
  if (!phi || !out || !w0 || !w1 || !w2 || !invw || !pot || !gp || !coef)
    return fail(CDFT_E_ARG, "hdiv: null pointer argument");
  if (k < 1 || n0 < 1 || n1 < 1 || n2 < 1)
    return fail(CDFT_E_ARG, "hdiv: sizes must be positive (k=%d, n=%d,%d,%d)", k, n0, n1, n2);
  if ((long double)n0 * n1 * n2 * k > 9.0e18L || k > 65535) return fail(CDFT_E_ARG, "hdiv: block too large");
  if (order != 2 && order != 4 && order != 6 && order != 8 && order != 12)
    return fail(CDFT_E_ARG, "hdiv: unsupported stencil order %d (2, 4, 6, 8 or 12)", order);
  if (phi == out) return fail(CDFT_E_ARG, "hdiv: out must not alias phi");
  return CDFT_OK;
}

template <typename T>
static int filter_impl(const T* x, T* out, T* work1, T* work2, const T* w0, const T* w1,
                       const T* w2, const T* invw, const T* pot, const T* gp, int k, int n0, int n1, int n2,
                       const double* coef, int order, const double* table, int degree,
                       hipStream_t stream) {
  int rc = validate(x, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
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
      rc = launch<T, 1>(y, nullptr, dst, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, s, stream);
    } else {
      rc = launch<T, 2>(y, prev, dst, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, s, stream);
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
                                   const double* w2, const double* invw, const double* pot, const double* gp,
                                   double* out, int k, int n0, int n1, int n2, const double* coef,
                                   int order, void* stream) {
  int rc = validate(phi, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  return launch<double, 0>(phi, nullptr, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order,
                           nullptr, 0, (hipStream_t)stream);
}

extern "C" int cdft_hdiv_apply_f32(const float* phi, const float* w0, const float* w1,
                                   const float* w2, const float* invw, const float* pot, const float* gp, float* out,
                                   int k, int n0, int n1, int n2, const double* coef, int order,
                                   void* stream) {
  int rc = validate(phi, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  return launch<float, 0>(phi, nullptr, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order,
                          nullptr, 0, (hipStream_t)stream);
}

extern "C" int cdft_cheb_step_f64(const double* y, const double* prev, const double* w0,
                                  const double* w1, const double* w2, const double* invw,
                                  const double* pot, const double* gp, double* out, int k, int n0, int n1, int n2,
                                  const double* coef, int order, const double* table, int step,
                                  void* stream) {
  int rc = validate(y, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!table || step < 0) return fail(CDFT_E_ARG, "cheb_step: null table or negative step");
  if (step > 0 && !prev) return fail(CDFT_E_ARG, "cheb_step: step > 0 needs the previous block");
  if (step == 0)
    return launch<double, 1>(y, nullptr, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, 0, (hipStream_t)stream);
  return launch<double, 2>(y, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, step, (hipStream_t)stream);
}

extern "C" int cdft_cheb_step_f32(const float* y, const float* prev, const float* w0,
                                  const float* w1, const float* w2, const float* invw,
                                  const float* pot, const float* gp, float* out, int k, int n0, int n1, int n2,
                                  const double* coef, int order, const double* table, int step,
                                  void* stream) {
  int rc = validate(y, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!table || step < 0) return fail(CDFT_E_ARG, "cheb_step: null table or negative step");
  if (step > 0 && !prev) return fail(CDFT_E_ARG, "cheb_step: step > 0 needs the previous block");
  if (step == 0)
    return launch<float, 1>(y, nullptr, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, 0, (hipStream_t)stream);
  return launch<float, 2>(y, prev, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, table, step, (hipStream_t)stream);
}

extern "C" int cdft_cheb_filter_f64(const double* x, double* out, double* work1, double* work2,
                                    const double* w0, const double* w1, const double* w2,
                                    const double* invw, const double* pot, const double* gp, int k, int n0, int n1,
                                    int n2, const double* coef, int order, const double* table,
                                    int degree, void* stream) {
  return filter_impl<double>(x, out, work1, work2, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef,
                             order, table, degree, (hipStream_t)stream);
}

extern "C" int cdft_cheb_filter_f32(const float* x, float* out, float* work1, float* work2,
                                    const float* w0, const float* w1, const float* w2,
                                    const float* invw, const float* pot, const float* gp, int k, int n0, int n1,
                                    int n2, const double* coef, int order, const double* table,
                                    int degree, void* stream) {
  return filter_impl<float>(x, out, work1, work2, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef,
                            order, table, degree, (hipStream_t)stream);
}

namespace cdft {
__global__ void scale_rows_kernel(const double* __restrict__ x, const double* __restrict__ w, long n, long total,
                                  double* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < total) out[p] = w[p % n] * x[p];
}
__global__ void scale_rows_kernel_f32(const float* __restrict__ x, const float* __restrict__ w, long n, long total,
                                      float* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < total) out[p] = w[p % n] * x[p];
}
}  // namespace cdft

template <typename T>
static int adjoint_impl(const T* x, const T* w0, const T* w1, const T* w2, const T* invw, const T* pot, const T* gp, T* out,
                        T* work, int k, int n0, int n1, int n2, const double* coef, int order, hipStream_t s) {
  int rc = validate(x, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!work || work == x || work == out) return fail(CDFT_E_ARG, "hdiv_apply_adjoint: work must be a distinct (k, N) block");
  const long n = (long)n0 * n1 * n2, total = n * k;
  const unsigned nb = (unsigned)((total + 255) / 256);
  if (sizeof(T) == 8)
    hipLaunchKernelGGL(scale_rows_kernel, dim3(nb), dim3(256), 0, s, (const double*)x, (const double*)invw, n, total,
                       (double*)work);
  else
    hipLaunchKernelGGL(scale_rows_kernel_f32, dim3(nb), dim3(256), 0, s, (const float*)x, (const float*)invw, n,
                       total, (float*)work);
  rc = check_launch("scale_rows_kernel");
  if (rc) return rc;
  return launch<T, 3>(work, x, out, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, nullptr, 0, s);
}

extern "C" int cdft_hdiv_apply_adjoint_f64(const double* x, const double* w0, const double* w1, const double* w2,
                                           const double* invw, const double* pot, const double* gp, double* out, double* work, int k,
                                           int n0, int n1, int n2, const double* coef, int order, void* stream) {
  return adjoint_impl<double>(x, w0, w1, w2, invw, pot, gp, out, work, k, n0, n1, n2, coef, order, (hipStream_t)stream);
}

extern "C" int cdft_hdiv_apply_adjoint_f32(const float* x, const float* w0, const float* w1, const float* w2,
                                           const float* invw, const float* pot, const float* gp, float* out, float* work, int k,
                                           int n0, int n1, int n2, const double* coef, int order, void* stream) {
  return adjoint_impl<float>(x, w0, w1, w2, invw, pot, gp, out, work, k, n0, n1, n2, coef, order, (hipStream_t)stream);
}

namespace cdft {
// ||A phi_v - e_v phi_v||_q from the stencil blocks' partials: one wavefront per vector, lane l sums the
// partials of tiles l, l+64, ... in order, then a fixed butterfly.
__device__ __forceinline__ void resnorm_combine(const double* __restrict__ partials, int tiles, int k, int v,
                                                double* __restrict__ norms) {
  double s = 0.0;
  for (int t = threadIdx.x; t < tiles; t += 64) s += partials[(long)t * k + v];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) norms[v] = sqrt(s > 0.0 ? s : 0.0);
}

__global__ __launch_bounds__(64) void resnorm_tiles_kernel(const double* __restrict__ partials, int tiles, int k,
                                                           double* __restrict__ norms) {
  resnorm_combine(partials, tiles, k, blockIdx.x, norms);
}

// The end of a Rayleigh-Ritz pass in one launch of k + 2 blocks: blocks 0..k-1 combine the residual
// norms, block k builds the next filter's Chebyshev table, block k + 1 writes max |G - I|. The three
// are independent (the table reads the eigenvalues and bounds, the error reads the Gram matrix).
__global__ __launch_bounds__(64) void resnorm_tail_kernel(const double* __restrict__ partials, int tiles, int k,
                                                          double* __restrict__ norms, const double* __restrict__ evals,
                                                          int index, const double* __restrict__ bounds, int degree,
                                                          double* __restrict__ table, const double* __restrict__ gram,
                                                          double* __restrict__ orth) {
  const int b = blockIdx.x;
  if (b < k) {
    resnorm_combine(partials, tiles, k, b, norms);
    return;
  }
  if (threadIdx.x != 0) return;
  if (b == k)
    cheb_table_rows(evals, index, bounds, degree, table);
  else
    max_abs_minus_identity(gram, k, orth);
}
}  // namespace cdft

// The stencil launch of the residual norms (MODE 4); returns the number of tile partials per vector.
static int resnorm_apply(const double* phi, const double* w0, const double* w1, const double* w2,
                         const double* invw, const double* pot, const double* gp, const double* evals, const double* q,
                         int k, int n0, int n1, int n2, const double* coef, int order, double* partials,
                         long partials_len, const double* norms, hipStream_t s, int* tiles) {
  int rc = validate(phi, partials, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order);
  if (rc) return rc;
  if (!evals || !q || !norms) return fail(CDFT_E_ARG, "hdiv_resnorm: null evals, q or norms");
  const March d = march_dims(n0, n1, n2, k, order - 1);
  const long blocks = (long)d.tiles * k;
  if (partials_len < blocks)
    return fail(CDFT_E_CAPACITY, "hdiv_resnorm: partials hold %ld doubles, the launch needs %ld", partials_len, blocks);
  *tiles = (int)d.tiles;
  return launch<double, 4>(phi, q, partials, w0, w1, w2, invw, pot, gp, k, n0, n1, n2, coef, order, evals, 0, s);
}

extern "C" int cdft_hdiv_resnorm_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                                     const double* invw, const double* pot, const double* gp, const double* evals, const double* q, int k,
                                     int n0, int n1, int n2, const double* coef, int order, double* partials,
                                     long partials_len, double* norms, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int tiles = 0;
  int rc = resnorm_apply(phi, w0, w1, w2, invw, pot, gp, evals, q, k, n0, n1, n2, coef, order, partials, partials_len,
                         norms, s, &tiles);
  if (rc) return rc;
  hipLaunchKernelGGL(resnorm_tiles_kernel, dim3(k), dim3(64), 0, s, partials, tiles, k, norms);
  return check_launch("resnorm_tiles");
}

extern "C" int cdft_hdiv_resnorm_tail_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                                          const double* invw, const double* pot, const double* gp, const double* evals,
                                          const double* q, int k, int n0, int n1, int n2, const double* coef, int order,
                                          double* partials, long partials_len, double* norms, int index,
                                          const double* bounds, int degree, double* table, const double* gram,
                                          double* orth, void* stream) {
  if (!bounds || !table || !gram || !orth || index < 0 || index >= k || degree < 1)
    return fail(CDFT_E_ARG, "hdiv_resnorm_tail: null bounds, table, gram or orth, or index/degree out of range");
  hipStream_t s = (hipStream_t)stream;
  int tiles = 0;
  int rc = resnorm_apply(phi, w0, w1, w2, invw, pot, gp, evals, q, k, n0, n1, n2, coef, order, partials, partials_len,
                         norms, s, &tiles);
  if (rc) return rc;
  hipLaunchKernelGGL(resnorm_tail_kernel, dim3(k + 2), dim3(64), 0, s, partials, tiles, k, norms, evals, index, bounds,
                     degree, table, gram, orth);
  return check_launch("resnorm_tail");
}

extern "C" long cdft_hdiv_pack_len(int n0, int n1, int n2, int order) {
  return (n0 < 1 || n1 < 1 || n2 < 1 || order < 1) ? 0L : 4L * n0 * (n1 + order) * (n2 + order);
}

extern "C" int cdft_hdiv_pack_f64(const double* w0, const double* w1, const double* w2, const double* invw, int n0,
                                  int n1, int n2, int order, double* gp, void* stream) {
  return pack_impl<double>(w0, w1, w2, invw, n0, n1, n2, order, gp, (hipStream_t)stream);
}

extern "C" int cdft_hdiv_pack_f32(const float* w0, const float* w1, const float* w2, const float* invw, int n0, int n1,
                                  int n2, int order, float* gp, void* stream) {
  return pack_impl<float>(w0, w1, w2, invw, n0, n1, n2, order, gp, (hipStream_t)stream);
}

extern "C" int cdft_hdiv_set_segment(int planes) {
  if (planes < 0) return fail(CDFT_E_ARG, "hdiv_set_segment: planes must be >= 0 (0 = automatic)");
  g_segment = planes;
  return CDFT_OK;
}

extern "C" int cdft_hdiv_set_tile(int tile_y, int tile_z) {
  if (tile_y < 0 || tile_y > TMAX) return fail(CDFT_E_ARG, "hdiv_set_tile: tile_y must be 0 (automatic) or 1..%d", TMAX);
  if (tile_z != 0 && tile_z != 13 && tile_z != 16 && !vec_width_ok(tile_z, 8))
    return fail(CDFT_E_ARG, "hdiv_set_tile: tile_z must be 0 (automatic), 13, 16, 20, 22, 24 or 28");
  g_tile_y = tile_y;
  g_tile_z = tile_z;
  return CDFT_OK;
}

extern "C" int cdft_hdiv_launch_shape(int n0, int n1, int n2, int k, int order, int* shape6) {
  if (!shape6 || n0 < 1 || n1 < 1 || n2 < 1 || k < 1) return fail(CDFT_E_ARG, "hdiv_launch_shape: bad arguments");
  if (order != 2 && order != 4 && order != 6 && order != 8 && order != 12)
    return fail(CDFT_E_ARG, "hdiv_launch_shape: unsupported stencil order %d", order);
  const March d = march_dims(n0, n1, n2, k, order - 1);
  shape6[0] = d.ty;
  shape6[1] = d.tz;
  shape6[2] = d.seg;
  shape6[3] = d.nseg;
  shape6[4] = d.threads;
  shape6[5] = (int)(d.tiles * (unsigned)k);
  return CDFT_OK;
}

extern "C" int cdft_hdiv_set_kernel(int which) {
  if (which != 0 && which != 1) return fail(CDFT_E_ARG, "hdiv_set_kernel: 0 (vector kernel) or 1 (column kernel)");
  g_kernel.store(which, std::memory_order_relaxed);
  return CDFT_OK;
}

extern "C" int cdft_hdiv_vec_set(int n0, int n1, int n2, int k, int elem_bytes, int tile_y, int tile_z, int segment) {
  if (n0 < 1 || n1 < 1 || n2 < 1 || k < 1 || (elem_bytes != 4 && elem_bytes != 8))
    return fail(CDFT_E_ARG, "hdiv_vec_set: bad shape, block width or element size");
  if (tile_y == 0) {  // remove the entry
    std::lock_guard<std::mutex> lock(g_vec_mu);
    for (int i = 0; i < g_vec_count; ++i) {
      VecSetting& e = g_vec_table[i];
      if (e.n0 == n0 && e.n1 == n1 && e.n2 == n2 && e.k == k && e.elem == elem_bytes) {
        e = g_vec_table[--g_vec_count];
        g_vec_next = g_vec_count % 64;
        break;
      }
    }
    return CDFT_OK;
  }
  if (tile_y < 1 || tile_y > VEC_TMAX_Y) return fail(CDFT_E_ARG, "hdiv_vec_set: tile_y must be 1..%d", VEC_TMAX_Y);
  if (!vec_width_ok(tile_z, elem_bytes)) return fail(CDFT_E_ARG, "hdiv_vec_set: no vector kernel of tile width %d", tile_z);
  if (segment < 0) return fail(CDFT_E_ARG, "hdiv_vec_set: segment must be >= 0 (0 = automatic)");
  const VecSetting entry{n0, n1, n2, k, elem_bytes, tile_y, tile_z, segment};
  std::lock_guard<std::mutex> lock(g_vec_mu);
  for (int i = 0; i < g_vec_count; ++i) {
    VecSetting& e = g_vec_table[i];
    if (e.n0 == n0 && e.n1 == n1 && e.n2 == n2 && e.k == k && e.elem == elem_bytes) {
      e = entry;
      return CDFT_OK;
    }
  }
  if (g_vec_count < 64) {
    g_vec_table[g_vec_count++] = entry;
  } else {
    g_vec_table[g_vec_next] = entry;
    g_vec_next = (g_vec_next + 1) % 64;
  }
  return CDFT_OK;
}

extern "C" int cdft_hdiv_vec_shape(int n0, int n1, int n2, int k, int elem_bytes, int* shape6) {
  if (!shape6 || n0 < 1 || n1 < 1 || n2 < 1 || k < 1 || (elem_bytes != 4 && elem_bytes != 8))
    return fail(CDFT_E_ARG, "hdiv_vec_shape: bad arguments");
  const March d = vec_dims(n0, n1, n2, k, elem_bytes);
  shape6[0] = d.ty;
  shape6[1] = d.tz;
  shape6[2] = d.seg;
  shape6[3] = d.nseg;
  shape6[4] = d.threads;
  shape6[5] = (int)(d.tiles * (unsigned)k);
  return CDFT_OK;
}

// ---- batched grids ----

namespace cdft {

template <typename T, int P, int MODE>
static int launch_batched_p(const void* desc, int n_grids, const int* launch, int step, int degree, hipStream_t s) {
  const dim3 grid((unsigned)launch[0]), block((unsigned)launch[1]);
  const BatchGrid<T>* g = (const BatchGrid<T>*)desc;
  if constexpr (P == 4 && MODE <= 2) {
    if (launch[2] < 0) {  // the vector kernel, tile width -launch[2]
#define CDFT_BVEC(W)                                                                                               \
  case W:                                                                                                          \
    hipLaunchKernelGGL((hdiv_batched_vec_kernel<T, MODE, W>), grid, block, 0, s, g, n_grids, step, degree); \
    return check_launch("hdiv_batched_vec_kernel");
      switch (-launch[2]) {
        CDFT_BVEC(16)
        CDFT_BVEC(20)
        CDFT_BVEC(24)
        CDFT_BVEC(28)
        case 22:
          if constexpr (sizeof(T) == 8) {
            hipLaunchKernelGGL((hdiv_batched_vec_kernel<T, MODE, 22>), grid, block, 0, s, g, n_grids, step, degree);
            return check_launch("hdiv_batched_vec_kernel");
          }
          break;
      }
#undef CDFT_BVEC
      return fail(CDFT_E_ARG, "batched: no vector kernel of tile width %d", -launch[2]);
    }
  }
  if (launch[2] == 13)
    hipLaunchKernelGGL((hdiv_batched_kernel<T, P, MODE, 13>), grid, block, 0, s, g, n_grids, step, degree);
  else
    hipLaunchKernelGGL((hdiv_batched_kernel<T, P, MODE, 16>), grid, block, 0, s, g, n_grids, step, degree);
  return check_launch("hdiv_batched_kernel");
}

template <typename T, int MODE>
static int launch_batched(const void* desc, int n_grids, const int* launch, int order, int step, int degree,
                          hipStream_t s) {
  switch (order) {
    case 2: return launch_batched_p<T, 1, MODE>(desc, n_grids, launch, step, degree, s);
    case 4: return launch_batched_p<T, 2, MODE>(desc, n_grids, launch, step, degree, s);
    case 6: return launch_batched_p<T, 3, MODE>(desc, n_grids, launch, step, degree, s);
    case 8: return launch_batched_p<T, 4, MODE>(desc, n_grids, launch, step, degree, s);
    default: return fail(CDFT_E_ARG, "batched: unsupported stencil order %d", order);
  }
}

static int check_batch_launch(const void* desc, int n_grids, const int* launch, int order) {
  if (!desc || !launch || n_grids < 1) return fail(CDFT_E_ARG, "batched: null descriptor or no grids");
  const bool vec = launch[2] < 0 && order == 8 && vec_width_ok(-launch[2], 8);  // float32: 16, 20, 24, 28 too
  if (launch[0] < 1 || launch[1] < 64 || launch[1] > NTMAX || (launch[2] != 13 && launch[2] != 16 && !vec))
    return fail(CDFT_E_ARG, "batched: launch {%d, %d, %d} is not one cdft_batch_plan wrote", launch[0], launch[1],
                launch[2]);
  if (order != 2 && order != 4 && order != 6 && order != 8) return fail(CDFT_E_ARG, "batched: bad order %d", order);
  return CDFT_OK;
}

}  // namespace cdft

static_assert(sizeof(BatchGrid<double>) == sizeof(BatchGrid<float>), "one descriptor size for both precisions");

extern "C" long cdft_batch_desc_bytes(int n_grids) {
  return n_grids < 1 ? 0L : (long)n_grids * (long)sizeof(BatchGrid<double>);
}

namespace cdft {

// The descriptor table of a batch (the element type only changes the pointer types; the layout is
// the same for float and double).
template <typename T>
static int batch_plan(int n_grids, const int* shapes, int k, int order, const double* coefs,
                      const void* const* ptrs, void* desc, long desc_bytes, int* launch) {
  if (n_grids < 1 || !shapes || !coefs || !ptrs || !desc || !launch) return fail(CDFT_E_ARG, "batch_plan: null argument");
  if (desc_bytes < cdft_batch_desc_bytes(n_grids)) return fail(CDFT_E_ARG, "batch_plan: descriptor buffer too small");
  if (order != 2 && order != 4 && order != 6 && order != 8) return fail(CDFT_E_ARG, "batch_plan: bad order %d", order);
  if (k < 1 || k > 65535) return fail(CDFT_E_ARG, "batch_plan: bad block width %d", k);
  BatchGrid<T>* g = (BatchGrid<T>*)desc;
  unsigned long start = 0;
  int threads = 64;
  // Order 8 runs on the vector kernel (unless cdft_hdiv_set_kernel(1) or a column-only tile width is
  // in force): one vector width for the batch, the override of cdft_hdiv_set_tile or the width that
  // leaves the fewest idle points over the whole batch; tile rows and segment lengths likewise.
  const int elem = (int)sizeof(T);
  const bool vec = order == 8 && g_kernel.load(std::memory_order_relaxed) == 0 &&
                   (g_tile_z == 0 || vec_width_ok(g_tile_z, elem));
  int vtz = g_tile_z;
  if (vec && vtz == 0) {
    const int* w = elem == 8 ? VEC_TZ_F64 : VEC_TZ_F32;
    const int nw = elem == 8 ? (int)std::size(VEC_TZ_F64) : (int)std::size(VEC_TZ_F32);
    long best_idle = -1;
    for (int c = 0; c < nw; ++c) {
      long idle = 0;
      for (int i = 0; i < n_grids; ++i) {
        const long n0 = shapes[3 * i], n1 = shapes[3 * i + 1], n2 = shapes[3 * i + 2];
        idle += n0 * n1 * (ceil_div((int)n2, w[c]) * w[c] - n2);
      }
      if (best_idle < 0 || idle <= best_idle) {
        best_idle = idle;
        vtz = w[c];
      }
    }
  }
  // The column kernel: one launch runs one tile width, the override of cdft_hdiv_set_tile, else the
  // width that leaves the fewest idle points over the whole batch (grids of different z extents may
  // each prefer another; the outputs do not depend on the tiling).
  int tz = (g_tile_z == 13 || g_tile_z == 16) ? g_tile_z : 0;
  if (!vec && tz == 0) {
    long idle[2] = {0, 0};
    for (int i = 0; i < n_grids; ++i) {
      const long n0 = shapes[3 * i], n1 = shapes[3 * i + 1], n2 = shapes[3 * i + 2];
      idle[0] += n0 * n1 * (ceil_div((int)n2, 13) * 13 - n2);
      idle[1] += n0 * n1 * (ceil_div((int)n2, 16) * 16 - n2);
    }
    tz = idle[0] < idle[1] ? 13 : 16;
  }
  for (int i = 0; i < n_grids; ++i) {
    const int n0 = shapes[3 * i], n1 = shapes[3 * i + 1], n2 = shapes[3 * i + 2];
    if (n0 < 1 || n1 < 1 || n2 < 1) return fail(CDFT_E_ARG, "batch_plan: grid %d has a bad shape", i);
    if ((long)(n0 + order) * (n1 + order) * (n2 + order) * 4 * (long)sizeof(T) >= 0x7fffffffL)
      return fail(CDFT_E_ARG, "batch_plan: grid %d too large for 32-bit offsets", i);
    const March d = vec ? vec_dims_with(n0, n1, n2, k, elem, vtz, g_tile_y ? g_tile_y : VEC_TMAX_Y, g_segment)
                        : march_dims(n0, n1, n2, k, order - 1, tz);
    if (d.threads > NTMAX || (vec && d.ty > VEC_TMAX_Y))
      return fail(CDFT_E_ARG, "batch_plan: grid %d needs %d threads", i, d.threads);
    threads = d.threads > threads ? d.threads : threads;
    const void* const* p = ptrs + 11 * i;
    for (int j = 0; j < 11; ++j)
      if (!p[j]) return fail(CDFT_E_ARG, "batch_plan: grid %d pointer %d is null", i, j);
    g[i].x = (const T*)p[0];
    g[i].out = (T*)p[1];
    g[i].work1 = (T*)p[2];
    g[i].work2 = (T*)p[3];
    g[i].w0 = (const T*)p[4];
    g[i].w1 = (const T*)p[5];
    g[i].w2 = (const T*)p[6];
    g[i].invw = (const T*)p[7];
    g[i].pot = (const T*)p[8];
    g[i].gp = (const T*)p[9];
    g[i].table = (const double*)p[10];
    for (int j = 0; j < 8; ++j) g[i].taps[j] = j < order ? coefs[order * i + j] : 0.0;
    g[i].n0 = n0;
    g[i].n1 = n1;
    g[i].n2 = n2;
    g[i].k = k;
    g[i].md = d;
    g[i].block_start = (unsigned)start;
    g[i].blocks = d.tiles * (unsigned)k;
    start += (unsigned long)d.tiles * (unsigned long)k;
    if (start > 0x7fffffffUL) return fail(CDFT_E_ARG, "batch_plan: too many blocks");
  }
  launch[0] = (int)start;
  launch[1] = threads;
  launch[2] = vec ? -vtz : tz;
  return CDFT_OK;
}

}  // namespace cdft

extern "C" int cdft_batch_plan_f64(int n_grids, const int* shapes, int k, int order, const double* coefs,
                                   const void* const* ptrs, void* desc, long desc_bytes, int* launch) {
  return batch_plan<double>(n_grids, shapes, k, order, coefs, ptrs, desc, desc_bytes, launch);
}

extern "C" int cdft_batch_plan_f32(int n_grids, const int* shapes, int k, int order, const double* coefs,
                                   const void* const* ptrs, void* desc, long desc_bytes, int* launch) {
  return batch_plan<float>(n_grids, shapes, k, order, coefs, ptrs, desc, desc_bytes, launch);
}

extern "C" int cdft_cheb_filter_batched_f64(const void* desc, int n_grids, const int* launch, int order, int degree,
                                            void* stream) {
  int rc = check_batch_launch(desc, n_grids, launch, order);
  if (rc) return rc;
  if (degree < 1) return fail(CDFT_E_ARG, "cheb_filter_batched: degree must be >= 1");
  for (int s = 0; s < degree; ++s) {
    rc = s == 0 ? launch_batched<double, 1>(desc, n_grids, launch, order, s, degree, (hipStream_t)stream)
                : launch_batched<double, 2>(desc, n_grids, launch, order, s, degree, (hipStream_t)stream);
    if (rc) return rc;
  }
  return CDFT_OK;
}

extern "C" int cdft_hdiv_apply_batched_f64(const void* desc, int n_grids, const int* launch, int order, void* stream) {
  int rc = check_batch_launch(desc, n_grids, launch, order);
  if (rc) return rc;
  return launch_batched<double, 0>(desc, n_grids, launch, order, 0, 1, (hipStream_t)stream);
}

extern "C" int cdft_cheb_filter_batched_f32(const void* desc, int n_grids, const int* launch, int order, int degree,
                                            void* stream) {
  int rc = check_batch_launch(desc, n_grids, launch, order);
  if (rc) return rc;
  if (degree < 1) return fail(CDFT_E_ARG, "cheb_filter_batched: degree must be >= 1");
  for (int s = 0; s < degree; ++s) {
    rc = s == 0 ? launch_batched<float, 1>(desc, n_grids, launch, order, s, degree, (hipStream_t)stream)
                : launch_batched<float, 2>(desc, n_grids, launch, order, s, degree, (hipStream_t)stream);
    if (rc) return rc;
  }
  return CDFT_OK;
}

extern "C" int cdft_hdiv_apply_batched_f32(const void* desc, int n_grids, const int* launch, int order, void* stream) {
  int rc = check_batch_launch(desc, n_grids, launch, order);
  if (rc) return rc;
  return launch_batched<float, 0>(desc, n_grids, launch, order, 0, 1, (hipStream_t)stream);
}
