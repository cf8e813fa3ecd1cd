// The cusp quadrature's node table, on the host (reference src/cdft/operators/quadrature.py:129-180 for
// the spheres, the base-cell numbering is this build's own).
//
// Each sphere is a composite Gauss-Legendre radius rule times a Gauss-Legendre (cos theta) x uniform
// phi direction rule whose resolution grows with the radius. The caller computes the per-shell radii,
// radial weights and direction counts (a few dozen values per sphere) and the unit directions of every
// direction count once per process; this function places the nodes, r * d + centre and wr * dw, with
// each product and sum rounded on its own (no contraction), numbers them in the order of their
// interpolation base cell floor((p - origin) / h) by a stable counting sort, and writes the table in
// that order. A numpy construction of the same table takes about 2 ms of a solve for 42,000 nodes;
// this takes a tenth of that.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.h"

using namespace cdft;

// shells: n_shells radii r, radial weights wr, direction counts n_theta (each >= 4) and owning
// atom; centres (3 per atom); directions: for each direction count n in [4, max_theta], the rows of
// its 2 n^2 unit directions, x then y then z, at dir_offset[n] in dir_x / dir_y / dir_z, and their
// solid-angle weights at the same offset in dir_w. m = the node count (sum of 2 n_theta^2).
// Outputs (node order = ascending base-cell key, stable): points (3, m) rows, owners (m), bare weights
// (m), base cells (3, m) rows, perm (m; sorted position -> enumeration index), keys (m; the sorted
// keys) and box6 = {lo0, lo1, lo2, ext0, ext1, ext2} of the base cells.
extern "C" int cdft_node_table(int n_shells, const double* r, const double* wr, const int* n_theta, const int* atom,
                               const double* centres, double spacing, const double* origin, int max_theta,
                               const long* dir_offset, const double* dir_x, const double* dir_y, const double* dir_z,
                               const double* dir_w, long m, double* points, int* owners, double* bare, int* base,
                               long* perm, long* keys, int* box6) {
#pragma clang fp contract(off)
  if (n_shells < 0 || m < 1 || !r || !wr || !n_theta || !atom || !centres || !origin || !dir_offset || !dir_x ||
      !dir_y || !dir_z || !dir_w || !points || !owners || !bare || !base || !perm || !keys || !box6 || !(spacing > 0))
    return fail(CDFT_E_ARG, "node_table: bad arguments");
  // One record per node, so that writing the table in sorted order gathers one or two cache lines per
  // node; scratch reused across calls (fresh arrays per call spend more time in first-touch page
  // faults than in the arithmetic).
  struct Node {
    double x, y, z, w;
    int own, bx, by, bz;
  };
  static thread_local std::vector<Node> nodes;
  static thread_local std::vector<long> key, start;
  if ((long)nodes.size() < m) {
    nodes.resize(m);
    key.resize(m);
  }
  // floor of a quotient in int range, without a library call
  auto ifloor = [](double q) {
    const int t = (int)q;
    return q < (double)t ? t - 1 : t;
  };
  long k = 0;
  for (int s = 0; s < n_shells; ++s) {
    const int n = n_theta[s];
    if (n < 4 || n > max_theta) return fail(CDFT_E_ARG, "node_table: direction count %d outside [4, %d]", n, max_theta);
    const long count = 2L * n * n;
    if (k + count > m) return fail(CDFT_E_ARG, "node_table: more nodes than m = %ld", m);
    const long off = dir_offset[n];
    const double rs = r[s], ws = wr[s];
    const double* c = centres + 3 * atom[s];
    for (long j = 0; j < count; ++j, ++k) {
      Node& p = nodes[k];
      const double x = dir_x[off + j] * rs;
      const double y = dir_y[off + j] * rs;
      const double z = dir_z[off + j] * rs;
      p.x = x + c[0];
      p.y = y + c[1];
      p.z = z + c[2];
      p.w = ws * dir_w[off + j];
      p.own = atom[s];
      const double qx = p.x - origin[0], qy = p.y - origin[1], qz = p.z - origin[2];
      p.bx = ifloor(qx / spacing);
      p.by = ifloor(qy / spacing);
      p.bz = ifloor(qz / spacing);
    }
  }
  if (k != m) return fail(CDFT_E_ARG, "node_table: %ld nodes, m = %ld", k, m);
  int lo[3] = {nodes[0].bx, nodes[0].by, nodes[0].bz}, hi[3] = {lo[0], lo[1], lo[2]};
  for (long i = 1; i < m; ++i) {
    const Node& p = nodes[i];
    lo[0] = std::min(lo[0], p.bx);
    lo[1] = std::min(lo[1], p.by);
    lo[2] = std::min(lo[2], p.bz);
    hi[0] = std::max(hi[0], p.bx);
    hi[1] = std::max(hi[1], p.by);
    hi[2] = std::max(hi[2], p.bz);
  }
  const long e0 = hi[0] - lo[0] + 1, e1 = hi[1] - lo[1] + 1, e2 = hi[2] - lo[2] + 1;
  const long cells = e0 * e1 * e2;
  if (cells > (1L << 31)) return fail(CDFT_E_ARG, "node_table: base-cell box of %ld cells", cells);
  start.assign(cells + 1, 0);
  for (long i = 0; i < m; ++i) {
    const Node& p = nodes[i];
    key[i] = ((long)(p.bx - lo[0]) * e1 + (p.by - lo[1])) * e2 + (p.bz - lo[2]);
    ++start[key[i] + 1];
  }
  for (long c = 0; c < cells; ++c) start[c + 1] += start[c];
  for (long i = 0; i < m; ++i) perm[start[key[i]]++] = i;  // stable: ascending enumeration within a cell
  for (long t = 0; t < m; ++t) {
    const long i = perm[t];
    const Node& p = nodes[i];
    points[t] = p.x;
    points[m + t] = p.y;
    points[2 * m + t] = p.z;
    owners[t] = p.own;
    bare[t] = p.w;
    base[t] = p.bx;
    base[m + t] = p.by;
    base[2 * m + t] = p.bz;
    keys[t] = key[i];
  }
  box6[0] = lo[0];
  box6[1] = lo[1];
  box6[2] = lo[2];
  box6[3] = (int)e0;
  box6[4] = (int)e1;
  box6[5] = (int)e2;
  return CDFT_OK;
}
