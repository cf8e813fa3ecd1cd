// Geometry and cusp-aware quadrature kernels (float64), built once per geometry and reused.
//
//   cusp_grid_fields : u, grad u, f^2 = exp(-2u) and W = -|grad u|^2/2 on the grid, with W and
//                      grad u replaced by their 4^3 Gauss-Legendre cell averages within 2h of a
//                      nucleus (reference src/cdft/operators/cusp.py:46-145)
//   face_weights     : f^2 at the staggered faces of each axis (cusp.py:168-214), max u reported
//   node_fields      : f^2, grad u, W and the partition weight P at the sphere nodes
//                      (quadrature.py:226-303, 340-419)
//   far_factor       : prod_a (1 - p_a) on the grid (quadrature.py:499-504)
//   node_stencils    : base cell and the three 1-D degree-7 Lagrange weight rows of each node
//                      (quadrature.py:56-93, 506-560)
//   interpolate      : out[f][node] = sum_abc wx wy wz field[f][base + abc] -- matrix-free L
//   lift             : out[f][g] = sum over nodes whose stencil covers g of x[f][node] * weight --
//                      matrix-free L^T in two separable stages (per z-column of cells, then a
//                      64-term combine per grid point), deterministic (no atomics)
//   gradient         : order-2P central first derivatives with a zero halo (laplacian.py:111-144)
#include "common.h"

namespace cdft {

constexpr int MAXAT = 8;

struct Nuclei {
  int n;
  double z[MAXAT];
  double px[MAXAT], py[MAXAT], pz[MAXAT];
};

struct GridDesc {
  int n0, n1, n2;
  double h, ox, oy, oz;
};

struct CellRule {  // the 4^3 Gauss-Legendre cell average
  double off[4];   // node offsets (GL node / 2 * h)
  double w[64];    // products wx*wy*wz in loop order (x outer, z inner)
  double total;    // their sum, accumulated in loop order
};

struct Spheres {
  int n;
  double cx[MAXAT], cy[MAXAT], cz[MAXAT];
  double radius[MAXAT], tcentre[MAXAT], twidth[MAXAT];
  double cell_width;  // Becke erf width, already times sqrt(2)
};

// Distances and lattice coordinates are formed with separately rounded products and sums (no fused
// multiply-add), as the reference's torch CPU expressions round them: a point two spacings from a
// nucleus sits on the cell-average threshold, and a coordinate one ulp off moves it across.
__device__ __forceinline__ double dist3(double dx, double dy, double dz) {
  return sqrt(rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz)));
}

// origin + index * h, the reference's grid point coordinate (grid.py:667-670, cusp.py:190-196).
__device__ __forceinline__ double lattice(double origin, double index, double h) {
  return rn_add(origin, rn_mul(index, h));
}

__global__ void cusp_grid_fields_kernel(GridDesc g, Nuclei nu, double reach, CellRule rule,
                                        double* __restrict__ u_out, double* __restrict__ gu,
                                        double* __restrict__ weight, double* __restrict__ wpot) {
  const long n = (long)g.n0 * g.n1 * g.n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % g.n2);
  const int i1 = (int)((p / g.n2) % g.n1);
  const int i0 = (int)(p / ((long)g.n1 * g.n2));
  const double x = lattice(g.ox, (double)i0, g.h), y = lattice(g.oy, (double)i1, g.h), z = lattice(g.oz, (double)i2, g.h);
  double u = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  bool near = false;
  for (int a = 0; a < nu.n; ++a) {
    const double dx = x - nu.px[a], dy = y - nu.py[a], dz = z - nu.pz[a];
    const double d = dist3(dx, dy, dz);
    u += nu.z[a] * d;
    if (d > 1.0e-12) {
      const double safe = d > 1.0e-30 ? d : 1.0e-30;
      gx += nu.z[a] * (dx / safe);
      gy += nu.z[a] * (dy / safe);
      gz += nu.z[a] * (dz / safe);
    }
    if (d <= reach) near = true;
  }
  double w = -0.5 * ((gx * gx + gy * gy) + gz * gz);
  if (near && nu.n > 0) {
    double ms = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    int q = 0;
    for (int a0 = 0; a0 < 4; ++a0)
      for (int a1 = 0; a1 < 4; ++a1)
        for (int a2 = 0; a2 < 4; ++a2, ++q) {
          const double sx = x + rule.off[a0], sy = y + rule.off[a1], sz = z + rule.off[a2];
          double hx = 0.0, hy = 0.0, hz = 0.0;
          for (int a = 0; a < nu.n; ++a) {
            const double dx = sx - nu.px[a], dy = sy - nu.py[a], dz = sz - nu.pz[a];
            double d = dist3(dx, dy, dz);
            d = d > 1.0e-300 ? d : 1.0e-300;
            hx = hx + nu.z[a] * dx / d;
            hy = hy + nu.z[a] * dy / d;
            hz = hz + nu.z[a] * dz / d;
          }
          ms = ms + rule.w[q] * ((hx * hx + hy * hy) + hz * hz);
          mx = mx + rule.w[q] * hx;
          my = my + rule.w[q] * hy;
          mz = mz + rule.w[q] * hz;
        }
    w = -0.5 * ms / rule.total;
    gx = mx / rule.total;
    gy = my / rule.total;
    gz = mz / rule.total;
  }
  u_out[p] = u;
  gu[p] = gx;
  gu[n + p] = gy;
  gu[2 * n + p] = gz;
  weight[p] = exp(-2.0 * u);
  wpot[p] = w;
}

__global__ void face_weights_kernel(GridDesc g, Nuclei nu, int axis, int order,
                                    double* __restrict__ out, unsigned long long* __restrict__ umax) {
  int s0 = g.n0, s1 = g.n1, s2 = g.n2;
  if (axis == 0) s0 += order;
  if (axis == 1) s1 += order;
  if (axis == 2) s2 += order;
  const long n = (long)s0 * s1 * s2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % s2);
  const int i1 = (int)((p / s2) % s1);
  const int i0 = (int)(p / ((long)s1 * s2));
  const int half = order / 2;
  double c0 = (double)i0, c1 = (double)i1, c2 = (double)i2;
  if (axis == 0) c0 = c0 - half + 0.5;
  if (axis == 1) c1 = c1 - half + 0.5;
  if (axis == 2) c2 = c2 - half + 0.5;
  const double x = lattice(g.ox, c0, g.h), y = lattice(g.oy, c1, g.h), z = lattice(g.oz, c2, g.h);
  double u = 0.0;
  for (int a = 0; a < nu.n; ++a) {
    const double dx = x - nu.px[a], dy = y - nu.py[a], dz = z - nu.pz[a];
    u = u + nu.z[a] * dist3(dx, dy, dz);
  }
  out[p] = exp(-2.0 * u);
  atomicMax(umax, (unsigned long long)__double_as_longlong(u));  // u >= 0: integer order = float order
}

__device__ __forceinline__ double bump(const Spheres& s, int a, double x, double y, double z, double* dist) {
  const double d = dist3(x - s.cx[a], y - s.cy[a], z - s.cz[a]);
  *dist = d;
  if (!(d < s.radius[a])) return 0.0;
  return 0.5 * erfc((d - s.tcentre[a]) / s.twidth[a]);
}

__device__ double far_factor_at(const Spheres& s, double x, double y, double z) {
  double far = 1.0, d;
  for (int a = 0; a < s.n; ++a) far = far * (1.0 - bump(s, a, x, y, z, &d));
  return far;
}

__global__ void far_factor_kernel(GridDesc g, Spheres s, double* __restrict__ far) {
  const long n = (long)g.n0 * g.n1 * g.n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % g.n2);
  const int i1 = (int)((p / g.n2) % g.n1);
  const int i0 = (int)(p / ((long)g.n1 * g.n2));
  far[p] = far_factor_at(s, lattice(g.ox, (double)i0, g.h), lattice(g.oy, (double)i1, g.h), lattice(g.oz, (double)i2, g.h));
}

// Partition P_a = Q (beta_a + sum_b beta_b [s_b >= R_b]) and the analytic fields at each node.
__global__ void node_fields_kernel(const double* __restrict__ pts, const int* __restrict__ owner,
                                   const double* __restrict__ base_w, long m, Nuclei nu, Spheres s,
                                   double* __restrict__ weights, double* __restrict__ f2,
                                   double* __restrict__ gu, double* __restrict__ wpot,
                                   double* __restrict__ partition) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
  double near = 1.0 - far_factor_at(s, x, y, z);
  double share;
  const int own = owner[i];
  if (s.n == 1) {
    share = 1.0;
  } else {
    double dist[MAXAT], cells[MAXAT];
    for (int a = 0; a < s.n; ++a) dist[a] = dist3(x - s.cx[a], y - s.cy[a], z - s.cz[a]);
    double total = 0.0;
    for (int a = 0; a < s.n; ++a) {
      double c = 1.0;
      for (int b = 0; b < s.n; ++b)
        if (b != a) c = c * 0.5 * erfc((dist[a] - dist[b]) / s.cell_width);
      cells[a] = c;
    }
    for (int a = 0; a < s.n; ++a) total += cells[a];
    total = total > 1.0e-300 ? total : 1.0e-300;
    for (int a = 0; a < s.n; ++a) cells[a] = cells[a] / total;
    share = cells[own];
    for (int b = 0; b < s.n; ++b) {
      if (b == own) continue;
      share = share + cells[b] * ((dist[b] >= s.radius[b]) ? 1.0 : 0.0);
    }
  }
  const double part = share * near;
  partition[i] = part;
  weights[i] = base_w[i] * part;
  double u = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (int a = 0; a < nu.n; ++a) {
    const double dx = x - nu.px[a], dy = y - nu.py[a], dz = z - nu.pz[a];
    double d = dist3(dx, dy, dz);
    d = d > 1.0e-300 ? d : 1.0e-300;
    u = u + nu.z[a] * d;
    gx = gx + nu.z[a] * (dx / d);
    gy = gy + nu.z[a] * (dy / d);
    gz = gz + nu.z[a] * (dz / d);
  }
  f2[i] = exp(-2.0 * u);
  gu[i] = gx;
  gu[m + i] = gy;
  gu[2 * m + i] = gz;
  wpot[i] = -0.5 * ((gx * gx + gy * gy) + gz * gz);
}

// Lagrange rows, degree 7 (8 points at offsets -3..4 around the base cell).
__global__ void node_stencils_kernel(const double* __restrict__ pts, long m, GridDesc g,
                                     int* __restrict__ base, double* __restrict__ lw,
                                     int* __restrict__ bad) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double origin[3] = {g.ox, g.oy, g.oz};
  const int dims[3] = {g.n0, g.n1, g.n2};
  for (int ax = 0; ax < 3; ++ax) {
    const double frac = (pts[3 * i + ax] - origin[ax]) / g.h;
    const double b = floor(frac);
    const double t = frac - b;
    const int bi = (int)b;
    base[3 * i + ax] = bi;
    if (bi - 3 < 0 || bi + 4 >= dims[ax]) atomicOr(bad, 1);
    double x[8];
    for (int j = 0; j < 8; ++j) x[j] = t - (double)(j - 3);
    for (int j = 0; j < 8; ++j) {
      double w = 1.0;
      for (int mm = 0; mm < 8; ++mm) {
        if (mm == j) continue;
        w = w * x[mm] / (double)(j - mm);
      }
      lw[(long)i * 24 + ax * 8 + j] = w;
    }
  }
}

// One wavefront per node: lane 8 j1 + j2 takes the (y, z) offsets (j1, j2) of the 8^3 stencil and
// walks the 8 x offsets, so every load instruction reads 8 rows of 8 consecutive values; the 64
// partial sums are combined by a fixed butterfly (deterministic). nf <= 4 fields.
__global__ __launch_bounds__(256) void interpolate_kernel(const double* __restrict__ field, int nf,
                                                          long nfield, const int* __restrict__ base,
                                                          const double* __restrict__ lw, long m,
                                                          GridDesc g, double* __restrict__ out) {
  const long node = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // wave-uniform
  if (node >= m) return;
  const int lane = (int)(threadIdx.x & 63);
  const int j1 = lane >> 3, j2 = lane & 7;
  const double* w = lw + node * 24;
  const double wyz = w[8 + j1] * w[16 + j2];
  const long plane = (long)g.n1 * g.n2;
  const long b0 = base[3 * node] - 3;
  const long col = (long)(base[3 * node + 1] - 3 + j1) * g.n2 + (base[3 * node + 2] - 3 + j2);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const double wa = w[a] * wyz;
    const long idx = (b0 + a) * plane + col;
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (f < nf) acc[f] += wa * field[f * nfield + idx];
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (f >= nf) break;
    double v = acc[f];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) out[f * m + node] = v;
  }
}

// The same interpolant with the stencil block shared: one wavefront per chunk of at most
// INTERP_CHUNK consecutive nodes that share one base cell (nodes are numbered in base-cell order, so a
// cell's nodes are consecutive). Lane 8 j1 + j2 loads its (y, z) column of the cell's 8^3 block for
// every field once; then, for each node of the chunk, it forms exactly interpolate_kernel's sum over
// the 8 x offsets and the same butterfly, so every output is bitwise interpolate_kernel's. A node's
// 8^3 loads (8 rows of 64 bytes per field) are made once per chunk instead of once per node.
constexpr int INTERP_CHUNK = 16;

template <int NF>
__global__ __launch_bounds__(256) void interpolate_chunks_kernel(const double* __restrict__ field, long nfield,
                                                                 const int* __restrict__ chunk_node, int n_chunks,
                                                                 const int* __restrict__ base,
                                                                 const double* __restrict__ lw, long m, GridDesc g,
                                                                 double* __restrict__ out) {
  const int chunk = (int)(((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);  // wave-uniform
  if (chunk >= n_chunks) return;
  const int lane = (int)(threadIdx.x & 63);
  const int j1 = lane >> 3, j2 = lane & 7;
  const int first = chunk_node[chunk], last = chunk_node[chunk + 1];
  const long plane = (long)g.n1 * g.n2;
  const long b0 = base[3 * first] - 3;
  const long col = (long)(base[3 * first + 1] - 3 + j1) * g.n2 + (base[3 * first + 2] - 3 + j2);
  double v[NF][8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long idx = (b0 + a) * plane + col;
#pragma unroll
    for (int f = 0; f < NF; ++f) v[f][a] = field[f * nfield + idx];
  }
  for (int node = first; node < last; ++node) {
    const double* w = lw + (long)node * 24;
    const double wyz = w[8 + j1] * w[16 + j2];
    double acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double wa = w[a] * wyz;
#pragma unroll
      for (int f = 0; f < NF; ++f) acc[f] += wa * v[f][a];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double x = acc[f];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
      if (lane == 0) out[f * m + node] = x;
    }
  }
}

struct CellBox {
  int lo0, lo1, lo2;  // smallest base cell
  int c0, c1, c2;     // cell-box extents
};

// Stage 1 of the transpose, one block of 64 threads per chunk of at most 64 consecutive nodes of
// one z-column (b0, b1) of the cell box. Nodes are numbered in base-cell order, so a chunk's nodes
// are sorted by b2 and reach z-cells [zlo, zlo + hz) (relative to the box's lo2, 8 cells per node).
// Thread t = 8 j0 + j1 accumulates Z[t][iz - zlo] = sum over the chunk's nodes of
// x[n] (wx_n[j0] wy_n[j1]) wz_n[iz - b2_n] in a sliding window of 8 registers, retiring the lowest
// z-cell whenever b2 advances: no dynamic register indexing, no atomics. Splitting columns into
// chunks keeps the serial walk short (the columns through a nucleus hold ~1,100 nodes) and gives
// ~1,000 blocks instead of one per column.
template <int NF>
__global__ __launch_bounds__(64) void lift_chunks_kernel(const double* __restrict__ x, long m,
                                                         const int* __restrict__ base,
                                                         const double* __restrict__ lw, int lo2,
                                                         const int* __restrict__ chunk_node,
                                                         const int* __restrict__ chunk_zlo,
                                                         const int* __restrict__ chunk_hz,
                                                         const int* __restrict__ chunk_off, long total,
                                                         double* __restrict__ z) {
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int j0 = t >> 3, j1 = t & 7;
  const int s = chunk_node[c], cnt = chunk_node[c + 1] - s;
  const int zlo = chunk_zlo[c], hz = chunk_hz[c];
  double* zc = z + chunk_off[c] + (long)t * hz;
  __shared__ double s_lw[64][25];
  __shared__ double s_x[NF][64];
  __shared__ int s_b2[64];
  const double* lc = lw + (long)s * 24;
#pragma unroll
  for (int q = 0; q < 24; ++q) {
    const int idx = q * 64 + t;  // coalesced over the chunk's contiguous 24-double records
    const int r = idx / 24;
    if (r < cnt) s_lw[r][idx - 24 * r] = lc[idx];
  }
  if (t < cnt) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_x[f][t] = x[f * m + s + t];
    s_b2[t] = base[3 * (s + t) + 2] - lo2;
  }
  __syncthreads();
  double acc[NF][8];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[f][k] = 0.0;
  int cur = zlo;
  for (int u = 0; u < cnt; ++u) {
    const int b2 = s_b2[u];
    while (cur < b2) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        zc[f * total + (cur - zlo)] = acc[f][0];
#pragma unroll
        for (int k = 0; k < 7; ++k) acc[f][k] = acc[f][k + 1];
        acc[f][7] = 0.0;
      }
      ++cur;
    }
    double wz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) wz[k] = s_lw[u][16 + k];
    const double wxy = s_lw[u][j0] * s_lw[u][8 + j1];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double w = wxy * wz[k];
#pragma unroll
      for (int f = 0; f < NF; ++f) acc[f][k] += s_x[f][u] * w;
    }
  }
  while (cur < zlo + hz) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      zc[f * total + (cur - zlo)] = acc[f][0];
#pragma unroll
      for (int k = 0; k < 7; ++k) acc[f][k] = acc[f][k + 1];
      acc[f][7] = 0.0;
    }
    ++cur;
  }
}

// Stage 2: each column's chunk slabs summed (chunks in node order) into one column slab
// Zcol[f][col][t][iz], iz in [0, c2 + 7): one thread per (column, t, iz) entry, blocks of 256 entries
// of one column (blockIdx.y). A column's chunks start at increasing z-cells, so the chunks reaching
// iz are a prefix of the column's chunk list (less those already past iz); they are visited four at
// a time with all loads issued before the in-order sums.
template <int NF>
__global__ __launch_bounds__(256) void lift_columns_kernel(const double* __restrict__ z, CellBox cb,
                                                           const int* __restrict__ chunk_zlo,
                                                           const int* __restrict__ chunk_hz,
                                                           const int* __restrict__ chunk_off,
                                                           const int* __restrict__ col_chunk, long total,
                                                           double* __restrict__ zcol) {
  const int col = blockIdx.y;
  const int ez = cb.c2 + 7;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= 64 * ez) return;
  const long ncols = (long)cb.c0 * cb.c1;
  const int c0 = col_chunk[col], c1 = col_chunk[col + 1];
  const int t = e / ez, iz = e - (e / ez) * ez;
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = 0.0;
  for (int c = c0; c < c1; c += 4) {
    double v[4][NF];
    bool in[4];
    bool more = true;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int cc = c + u;
      in[u] = false;
      if (cc < c1) {
        const int zlo = chunk_zlo[cc];
        if (zlo > iz) more = false;
        const int r = iz - zlo, hz = chunk_hz[cc];
        in[u] = zlo <= iz && r < hz;
        const long idx = in[u] ? (long)chunk_off[cc] + (long)t * hz + r : 0L;  // in range either way
#pragma unroll
        for (int f = 0; f < NF; ++f) v[u][f] = in[u] ? z[f * total + idx] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (in[u])
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] += v[u][f];
    if (!more) break;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) zcol[((f * ncols + col) * 64 + t) * ez + iz] = acc[f];
}

// Stage 3: out[f][p] = add[f][p] + sum_{j0, j1} Zcol[f][col(i0 - j0 + 3, i1 - j1 + 3)][8 j0 + j1][iz]
// (j0 outer, j1 inner), iz = i2 - lo2 + 3; add (or zero) where no node reaches.
template <int NF>
__global__ void lift_combine_kernel(const double* __restrict__ zcol, CellBox cb, GridDesc g,
                                    const double* __restrict__ add, double* __restrict__ out) {
  const long n = (long)g.n0 * g.n1 * g.n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % g.n2);
  const int i1 = (int)((p / g.n2) % g.n1);
  const int i0 = (int)(p / ((long)g.n1 * g.n2));
  const int ez = cb.c2 + 7;
  const int iz = i2 - (cb.lo2 - 3);
  const long ncols = (long)cb.c0 * cb.c1;
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = 0.0;
  if (iz >= 0 && iz < ez) {
    for (int j0 = 0; j0 < 8; ++j0) {
      const int b0 = i0 - j0 + 3 - cb.lo0;
      if (b0 < 0 || b0 >= cb.c0) continue;
      // the row's eight loads issued together, then summed in j1 order; a column outside the box
      // adds +0.0, which leaves the sum's bits as skipping it would (the sum starts at +0.0, so it
      // is never -0.0)
      double v[8][NF];
#pragma unroll
      for (int j1 = 0; j1 < 8; ++j1) {
        const int b1 = i1 - j1 + 3 - cb.lo1;
        const bool in = b1 >= 0 && b1 < cb.c1;
        const long idx = in ? (((long)b0 * cb.c1 + b1) * 64 + j0 * 8 + j1) * ez + iz : 0L;
#pragma unroll
        for (int f = 0; f < NF; ++f) v[j1][f] = in ? zcol[f * ncols * 64 * ez + idx] : 0.0;
      }
#pragma unroll
      for (int j1 = 0; j1 < 8; ++j1)
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] += v[j1][f];
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) out[f * n + p] = add ? acc[f] + add[f * n + p] : acc[f];
}

template <int P>
__global__ void gradient_kernel(const double* __restrict__ field, int nf, GridDesc g,
                                const double* __restrict__ coef /* 2P+1 central taps / h */,
                                double* __restrict__ out) {
  const long n = (long)g.n0 * g.n1 * g.n2;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i2 = (int)(p % g.n2);
  const int i1 = (int)((p / g.n2) % g.n1);
  const int i0 = (int)(p / ((long)g.n1 * g.n2));
  const long plane = (long)g.n1 * g.n2;
  const int idx[3] = {i0, i1, i2};
  const int dims[3] = {g.n0, g.n1, g.n2};
  const long stride[3] = {plane, (long)g.n2, 1};
  for (int f = 0; f < nf; ++f) {
    const double* src = field + (long)f * n;
    for (int ax = 0; ax < 3; ++ax) {
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
      out[((long)f * 3 + ax) * n + p] = acc;
    }
  }
}

// Unit cusp density exp(-2 Z s) and its closed-form potential (pi/Z^3)[1 - e^{-2Zs}(1 + Zs)]/s with
// the small-s series below Zs = 1e-3 (reference quadrature.py:805-820, scf/step.py:128-165), per
// nucleus, at the grid points (pts == NULL) or at a list of points.
__global__ void hydrogenic_fields_kernel(const double* __restrict__ pts, long m, GridDesc g, Nuclei nu,
                                         double* __restrict__ dens, double* __restrict__ v1s) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  double x, y, z;
  if (pts) {
    x = pts[3 * p];
    y = pts[3 * p + 1];
    z = pts[3 * p + 2];
  } else {
    const int i2 = (int)(p % g.n2);
    const int i1 = (int)((p / g.n2) % g.n1);
    const int i0 = (int)(p / ((long)g.n1 * g.n2));
    x = lattice(g.ox, (double)i0, g.h);
    y = lattice(g.oy, (double)i1, g.h);
    z = lattice(g.oz, (double)i2, g.h);
  }
  for (int a = 0; a < nu.n; ++a) {
    const double s = dist3(x - nu.px[a], y - nu.py[a], z - nu.pz[a]);
    const double zz = nu.z[a];
    if (dens) dens[(long)a * m + p] = exp(-2.0 * zz * s);
    if (v1s) {
      const double t = zz * s;
      double v;
      if (t < 1.0e-3) {
        v = (M_PI / (zz * zz)) * (1.0 - (2.0 / 3.0) * (t * t) + (2.0 / 3.0) * (t * t * t) - 0.4 * (t * t * t * t) +
                                  (8.0 / 45.0) * (t * t * t * t * t));
      } else {
        const double safe = s > 1.0e-300 ? s : 1.0e-300;
        v = (M_PI / (zz * zz * zz)) * (1.0 - exp(-2.0 * t) * (1.0 + t)) / safe;
      }
      v1s[(long)a * m + p] = v;
    }
  }
}

static unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace cdft

using namespace cdft;

static int check_nuclei(int n_atoms) {
  if (n_atoms < 0 || n_atoms > MAXAT) return fail(CDFT_E_ARG, "at most %d nuclei (got %d)", MAXAT, n_atoms);
  return CDFT_OK;
}

static Nuclei make_nuclei(int n_atoms, const double* charges, const double* positions) {
  Nuclei nu;
  nu.n = n_atoms;
  for (int a = 0; a < n_atoms; ++a) {
    nu.z[a] = charges[a];
    nu.px[a] = positions[3 * a];
    nu.py[a] = positions[3 * a + 1];
    nu.pz[a] = positions[3 * a + 2];
  }
  return nu;
}

static GridDesc make_grid(const int* shape, double h, const double* origin) {
  GridDesc g;
  g.n0 = shape[0];
  g.n1 = shape[1];
  g.n2 = shape[2];
  g.h = h;
  g.ox = origin[0];
  g.oy = origin[1];
  g.oz = origin[2];
  return g;
}

// charges, positions (n_atoms x 3), shape, origin, gl_offsets (4), gl_weights (64): host arrays.
extern "C" int cdft_cusp_grid_fields_f64(int n_atoms, const double* charges, const double* positions,
                                         const int* shape, double h, const double* origin,
                                         double reach, const double* gl_offsets,
                                         const double* gl_weights, double gl_total, double* u,
                                         double* grad_u, double* weight, double* wpot, void* stream) {
  int rc = check_nuclei(n_atoms);
  if (rc) return rc;
  if (!u || !grad_u || !weight || !wpot || !shape || !origin) return fail(CDFT_E_ARG, "cusp_grid_fields: null pointer");
  GridDesc g = make_grid(shape, h, origin);
  CellRule rule;
  for (int i = 0; i < 4; ++i) rule.off[i] = gl_offsets[i];
  for (int i = 0; i < 64; ++i) rule.w[i] = gl_weights[i];
  rule.total = gl_total;
  const long n = (long)g.n0 * g.n1 * g.n2;
  hipLaunchKernelGGL(cusp_grid_fields_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     make_nuclei(n_atoms, charges, positions), reach, rule, u, grad_u, weight, wpot);
  return check_launch("cusp_grid_fields");
}

// umax: device unsigned long long, zeroed by the caller; receives the bits of max u (>= 0).
extern "C" int cdft_face_weights_f64(int n_atoms, const double* charges, const double* positions,
                                     const int* shape, double h, const double* origin, int order,
                                     double* w0, double* w1, double* w2, unsigned long long* umax,
                                     void* stream) {
  int rc = check_nuclei(n_atoms);
  if (rc) return rc;
  if (!w0 || !w1 || !w2 || !umax) return fail(CDFT_E_ARG, "face_weights: null pointer");
  GridDesc g = make_grid(shape, h, origin);
  Nuclei nu = make_nuclei(n_atoms, charges, positions);
  double* outs[3] = {w0, w1, w2};
  for (int axis = 0; axis < 3; ++axis) {
    long n = (long)(g.n0 + (axis == 0 ? order : 0)) * (g.n1 + (axis == 1 ? order : 0)) * (g.n2 + (axis == 2 ? order : 0));
    hipLaunchKernelGGL(face_weights_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g, nu,
                       axis, order, outs[axis], umax);
    rc = check_launch("face_weights");
    if (rc) return rc;
  }
  return CDFT_OK;
}

static Spheres make_spheres(int n, const double* sph, double cell_width) {
  // sph: n rows of (cx, cy, cz, radius, taper_centre, taper_width)
  Spheres s;
  s.n = n;
  for (int a = 0; a < n; ++a) {
    s.cx[a] = sph[6 * a];
    s.cy[a] = sph[6 * a + 1];
    s.cz[a] = sph[6 * a + 2];
    s.radius[a] = sph[6 * a + 3];
    s.tcentre[a] = sph[6 * a + 4];
    s.twidth[a] = sph[6 * a + 5];
  }
  s.cell_width = cell_width;
  return s;
}

extern "C" int cdft_node_fields_f64(const double* pts, const int* owner, const double* base_w, long m,
                                    int n_atoms, const double* charges, const double* positions,
                                    const double* spheres, double cell_width, double* weights,
                                    double* f2, double* grad_u, double* wpot, double* partition,
                                    void* stream) {
  int rc = check_nuclei(n_atoms);
  if (rc) return rc;
  if (!pts || !owner || !base_w || !weights || !f2 || !grad_u || !wpot || !partition || m < 1)
    return fail(CDFT_E_ARG, "node_fields: bad arguments");
  hipLaunchKernelGGL(node_fields_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, (hipStream_t)stream, pts, owner,
                     base_w, m, make_nuclei(n_atoms, charges, positions), make_spheres(n_atoms, spheres, cell_width),
                     weights, f2, grad_u, wpot, partition);
  return check_launch("node_fields");
}

extern "C" int cdft_far_factor_f64(const int* shape, double h, const double* origin, int n_spheres,
                                   const double* spheres, double* far, void* stream) {
  int rc = check_nuclei(n_spheres);
  if (rc) return rc;
  if (!far) return fail(CDFT_E_ARG, "far_factor: null pointer");
  GridDesc g = make_grid(shape, h, origin);
  const long n = (long)g.n0 * g.n1 * g.n2;
  hipLaunchKernelGGL(far_factor_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, g,
                     make_spheres(n_spheres, spheres, 0.0), far);
  return check_launch("far_factor");
}

extern "C" int cdft_node_stencils_f64(const double* pts, long m, const int* shape, double h,
                                      const double* origin, int* base, double* lw, int* bad,
                                      void* stream) {
  if (!pts || !base || !lw || !bad || m < 1) return fail(CDFT_E_ARG, "node_stencils: bad arguments");
  GridDesc g = make_grid(shape, h, origin);
  hipLaunchKernelGGL(node_stencils_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, (hipStream_t)stream, pts, m, g,
                     base, lw, bad);
  return check_launch("node_stencils");
}

extern "C" int cdft_interpolate_f64(const double* field, int nf, const int* base, const double* lw,
                                    long m, const int* shape, double h, const double* origin,
                                    double* out, void* stream) {
  if (!field || !base || !lw || !out || nf < 1 || nf > 4 || m < 1)
    return fail(CDFT_E_ARG, "interpolate: bad arguments (nf=%d)", nf);
  GridDesc g = make_grid(shape, h, origin);
  const long n = (long)g.n0 * g.n1 * g.n2;
  hipLaunchKernelGGL(interpolate_kernel, dim3(blocks_for(m * 64, 256)), dim3(256), 0, (hipStream_t)stream, field,
                     nf, n, base, lw, m, g, out);
  return check_launch("interpolate");
}

extern "C" int cdft_interpolate_chunks_f64(const double* field, int nf, const int* chunk_node, int n_chunks,
                                           const int* base, const double* lw, long m, const int* shape, double h,
                                           const double* origin, double* out, void* stream) {
  if (!field || !chunk_node || !base || !lw || !out || nf < 1 || nf > 4 || m < 1 || n_chunks < 1)
    return fail(CDFT_E_ARG, "interpolate_chunks: bad arguments (nf=%d, chunks=%d)", nf, n_chunks);
  GridDesc g = make_grid(shape, h, origin);
  const long n = (long)g.n0 * g.n1 * g.n2;
  const dim3 grid(blocks_for(64L * n_chunks, 256)), block(256);
  hipStream_t s = (hipStream_t)stream;
  switch (nf) {
    case 1: hipLaunchKernelGGL(interpolate_chunks_kernel<1>, grid, block, 0, s, field, n, chunk_node, n_chunks, base, lw, m, g, out); break;
    case 2: hipLaunchKernelGGL(interpolate_chunks_kernel<2>, grid, block, 0, s, field, n, chunk_node, n_chunks, base, lw, m, g, out); break;
    case 3: hipLaunchKernelGGL(interpolate_chunks_kernel<3>, grid, block, 0, s, field, n, chunk_node, n_chunks, base, lw, m, g, out); break;
    default: hipLaunchKernelGGL(interpolate_chunks_kernel<4>, grid, block, 0, s, field, n, chunk_node, n_chunks, base, lw, m, g, out); break;
  }
  return check_launch("interpolate_chunks");
}

// cell_box (host): lo0, lo1, lo2, c0, c1, c2; the chunk tables (device, int32): chunk_node
// (n_chunks + 1 node offsets), chunk_zlo and chunk_hz (first z-cell relative to lo2 and height of
// each chunk's slab), chunk_off (n_chunks + 1 offsets of the slabs, in doubles per field, each slab
// 64 * hz), col_chunk (c0 * c1 + 1 chunk offsets per column). work: nf * (work_per_field +
// c0 c1 64 (c2 + 7)) doubles, work_per_field = chunk_off[n_chunks]. add may be NULL.
extern "C" int cdft_lift_f64(const double* x, int nf, long m, const int* base, const double* lw,
                             const int* shape, double h, const double* origin, const int* cell_box,
                             int n_chunks, const int* chunk_node, const int* chunk_zlo, const int* chunk_hz,
                             const int* chunk_off, const int* col_chunk, long work_per_field,
                             const double* add, double* work, long work_len, double* out, void* stream) {
  if (!x || !base || !lw || !out || !cell_box || !work || !chunk_node || !chunk_zlo || !chunk_hz || !chunk_off ||
      !col_chunk || nf < 1 || nf > 4 || n_chunks < 1)
    return fail(CDFT_E_ARG, "lift: bad arguments (nf=%d, chunks=%d)", nf, n_chunks);
  const long col_doubles = (long)cell_box[3] * cell_box[4] * 64 * (cell_box[5] + 7);
  if ((long)cell_box[3] * cell_box[4] > 65535)
    return fail(CDFT_E_ARG, "lift: %d x %d node columns exceed one launch's second grid dimension", cell_box[3],
                cell_box[4]);
  if (work_len < (work_per_field + col_doubles) * nf)
    return fail(CDFT_E_CAPACITY, "lift: workspace holds %ld doubles, needs %ld", work_len,
                (work_per_field + col_doubles) * nf);
  GridDesc g = make_grid(shape, h, origin);
  CellBox cb;
  cb.lo0 = cell_box[0];
  cb.lo1 = cell_box[1];
  cb.lo2 = cell_box[2];
  cb.c0 = cell_box[3];
  cb.c1 = cell_box[4];
  cb.c2 = cell_box[5];
  const long n = (long)g.n0 * g.n1 * g.n2;
  double* zcol = work + work_per_field * nf;
  hipStream_t s = (hipStream_t)stream;
#define CDFT_LIFT(NFV)                                                                                        \
  hipLaunchKernelGGL(lift_chunks_kernel<NFV>, dim3((unsigned)n_chunks), dim3(64), 0, s, x, m, base, lw, cb.lo2, \
                     chunk_node, chunk_zlo, chunk_hz, chunk_off, work_per_field, work);                        \
  rc = check_launch("lift_chunks");                                                                          \
  if (rc) return rc;                                                                                         \
  hipLaunchKernelGGL(lift_columns_kernel<NFV>, dim3(blocks_for(64L * (cb.c2 + 7), 256), (unsigned)(cb.c0 * cb.c1)), \
                     dim3(256), 0, s, work, cb,                                                                \
                     chunk_zlo, chunk_hz, chunk_off, col_chunk, work_per_field, zcol);                        \
  rc = check_launch("lift_columns");                                                                         \
  if (rc) return rc;                                                                                         \
  hipLaunchKernelGGL(lift_combine_kernel<NFV>, dim3(blocks_for(n, 256)), dim3(256), 0, s, zcol, cb, g, add, out); \
  return check_launch("lift_combine")
  int rc = CDFT_OK;
  switch (nf) {
    case 1: CDFT_LIFT(1);
    case 2: CDFT_LIFT(2);
    case 3: CDFT_LIFT(3);
    default: CDFT_LIFT(4);
  }
#undef CDFT_LIFT
}

// coef: DEVICE array of the 2P+1 central first-derivative weights divided by h.
extern "C" int cdft_gradient_f64(const double* field, int nf, const int* shape, double h,
                                 const double* origin, const double* coef, int order, double* out,
                                 void* stream) {
  if (!field || !coef || !out || nf < 1) return fail(CDFT_E_ARG, "gradient: bad arguments");
  GridDesc g = make_grid(shape, h, origin);
  const long n = (long)g.n0 * g.n1 * g.n2;
  const unsigned blocks = blocks_for(n, 256);
  hipStream_t s = (hipStream_t)stream;
  switch (order) {
    case 2: hipLaunchKernelGGL(gradient_kernel<1>, dim3(blocks), dim3(256), 0, s, field, nf, g, coef, out); break;
    case 4: hipLaunchKernelGGL(gradient_kernel<2>, dim3(blocks), dim3(256), 0, s, field, nf, g, coef, out); break;
    case 6: hipLaunchKernelGGL(gradient_kernel<3>, dim3(blocks), dim3(256), 0, s, field, nf, g, coef, out); break;
    case 8: hipLaunchKernelGGL(gradient_kernel<4>, dim3(blocks), dim3(256), 0, s, field, nf, g, coef, out); break;
    case 12: hipLaunchKernelGGL(gradient_kernel<6>, dim3(blocks), dim3(256), 0, s, field, nf, g, coef, out); break;
    default: return fail(CDFT_E_ARG, "gradient: unsupported order %d", order);
  }
  return check_launch("gradient");
}

// pts: device (m x 3) or NULL for the grid points (then m must be n0*n1*n2). dens/v1s: (na, m), either may be NULL.
extern "C" int cdft_hydrogenic_fields_f64(const double* pts, long m, const int* shape, double h,
                                          const double* origin, int n_atoms, const double* charges,
                                          const double* positions, double* dens, double* v1s,
                                          void* stream) {
  int rc = check_nuclei(n_atoms);
  if (rc) return rc;
  if (m < 1 || (!dens && !v1s)) return fail(CDFT_E_ARG, "hydrogenic_fields: bad arguments");
  GridDesc g = make_grid(shape, h, origin);
  if (!pts && m != (long)g.n0 * g.n1 * g.n2) return fail(CDFT_E_ARG, "hydrogenic_fields: m must be the grid size");
  hipLaunchKernelGGL(hydrogenic_fields_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, (hipStream_t)stream, pts, m, g,
                     make_nuclei(n_atoms, charges, positions), dens, v1s);
  return check_launch("hydrogenic_fields");
}
