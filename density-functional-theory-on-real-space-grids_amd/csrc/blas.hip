// K3 / K4 and the small dense kernels of Rayleigh-Ritz and CheFSI, all float64 and deterministic.
//
//   wgram   : S_ij = scale * sum_p q_p a_ip b_jp   (block partials in a fixed order, second pass)
//   rotate  : Y = R X  (R is kout x kin on the device, row-major; or R^T)
//   lincomb : out = sum_t c_t x_t with c_t a host constant optionally times / over a device scalar
//   resnorm : r_i = sqrt(max(sum_p q_p (Av_ip - e_i v_ip)^2, 0))
//   chol_inv: T = L^{-1}, S = L L^T, flag on a non-positive pivot
//   eigh    : cyclic Jacobi on a symmetrised k x k matrix, eigenvalues ascending
//
// Reductions never use floating-point atomics: every block writes its partial sums, and one block
// sums the partials in block order, so a repeat is bitwise identical.
#include <mutex>

#include "common.h"
#include "ritz_tail.h"

namespace cdft {

constexpr int RB = 256;          // threads per reduction block
constexpr int KMAX = 8;          // largest tile of a Gram product per launch
constexpr int MAX_PARTIALS = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sums of COUNT values per thread (a block of RB = 256 threads), each handed to
// store(entry, sum) by lane 0 of the wavefront that forms it. The values go through LDS by entry
// (row e holds every thread's value e), then wavefront w reduces entries w, w + 4, ...: lane l adds
// threads l, l + 64, l + 128, l + 192 in that order, then a fixed butterfly. A butterfly per entry
// and wavefront followed by a cross-wavefront pass, the obvious alternative, spends 25 butterflies
// per wavefront on a 5 x 5 Gram matrix instead of 7 and took about 5 us more per launch. Entries go
// through the tile GROUP at a time, which bounds its LDS.
template <int COUNT, typename Store>
__device__ __forceinline__ void block_partials(const double (&acc)[COUNT], Store store) {
  static_assert(RB == 256, "four wavefronts per block");
  constexpr int GROUP = COUNT < 32 ? COUNT : 32;
  __shared__ double tile[GROUP][RB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int g0 = 0; g0 < COUNT; g0 += GROUP) {
    if (g0 > 0) __syncthreads();  // the previous group's reads are done
#pragma unroll
    for (int e = 0; e < GROUP; ++e)
      if (g0 + e < COUNT) tile[e][threadIdx.x] = acc[g0 + e];
    __syncthreads();
    for (int e = wave; e < GROUP && g0 + e < COUNT; e += RB / 64) {
      double v = ((tile[e][lane] + tile[e][lane + 64]) + tile[e][lane + 128]) + tile[e][lane + 192];
      v = wave_sum(v);
      if (lane == 0) store(g0 + e, v);
    }
  }
}

// KA x KB accumulators per thread (compile-time; runtime ka <= KA, kb <= KB guard the loads).
template <int KA, int KB>
__global__ __launch_bounds__(RB) void wgram_partial_kernel(const double* __restrict__ a, int ka,
                                                           const double* __restrict__ b, int kb,
                                                           const double* __restrict__ q, long n,
                                                           double* __restrict__ partials) {
  double acc[KA * KB];
#pragma unroll
  for (int i = 0; i < KA * KB; ++i) acc[i] = 0.0;
  const long stride = (long)gridDim.x * RB;
  const bool same = a == b && ka == kb;  // a Gram matrix: each column is read once
  // U points per pass, all loads issued before the products (one pass covers a thread's points at
  // the usual launch); the points are accumulated in ascending order, as a plain loop would.
  constexpr int U = KA * KB <= 25 ? 4 : (KA * KB <= 36 ? 2 : 1);  // register budget
  for (long p0 = (long)blockIdx.x * RB + threadIdx.x; p0 < n; p0 += U * stride) {
    double w[U], bv[U][KB];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long p = p0 + u * stride;
      const bool in = p < n;
      w[u] = in ? (q ? q[p] : 1.0) : 0.0;
#pragma unroll
      for (int j = 0; j < KB; ++j) bv[u][j] = (in && j < kb) ? b[(long)j * n + p] : 0.0;
    }
    double avs[U][KA];
    if (!same) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long p = p0 + u * stride;
#pragma unroll
        for (int i = 0; i < KA; ++i) avs[u][i] = (p < n && i < ka) ? a[(long)i * n + p] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (p0 + u * stride < n) {
#pragma unroll
        for (int i = 0; i < KA; ++i) {
          if (i < ka) {
            const double av = (same && i < KB ? bv[u][i] : avs[u][i]) * w[u];
#pragma unroll
            for (int j = 0; j < KB; ++j) acc[i * KB + j] += av * bv[u][j];
          }
        }
      }
    }
  }
  double* row = partials + (long)blockIdx.x * KMAX * KMAX;
  block_partials<KA * KB>(acc, [&](int e, double v) {
    const int i = e / KB, j = e - (e / KB) * KB;
    if (i < ka && j < kb) row[i * KMAX + j] = v;
  });
}

// One wavefront per output entry: lane l sums partials l, l+64, ... in order, then a fixed
// butterfly. post: 0 plain, 1 sqrt(max(x, 0)).
__global__ __launch_bounds__(64) void wgram_final_kernel(const double* __restrict__ partials, int nblocks,
                                                         int ka, int kb, double scale, int post,
                                                         double* __restrict__ out, int ldo, int row0,
                                                         int col0) {
  const int e = blockIdx.x;
  if (e >= ka * kb) return;
  const int i = e / kb, j = e % kb;
  double s = 0.0;
  for (int blk = threadIdx.x; blk < nblocks; blk += 64) s += partials[(long)blk * KMAX * KMAX + i * KMAX + j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) {
    s *= scale;
    if (post == 1) s = sqrt(s > 0.0 ? s : 0.0);
    out[(row0 + i) * ldo + col0 + j] = s;
  }
}

static int reduction_blocks(long n) {
  // four points per thread: the block-sum epilogue (k^2 wavefront reductions) outweighs the
  // loads at one point per thread
  long blocks = (n + 4 * RB - 1) / (4 * RB);
  if (blocks < 1) blocks = 1;
  if (blocks > MAX_PARTIALS) blocks = MAX_PARTIALS;
  return (int)blocks;
}

// ---- rotation ----

constexpr int KROT = 16;  // largest subspace handled by the rotation and small dense kernels

__global__ __launch_bounds__(256) void rotate_kernel(const double* __restrict__ r, int kin, int kout,
                                                     int transpose, const double* __restrict__ x,
                                                     double* __restrict__ y, long n) {
  __shared__ double sr[KROT * KROT];
  for (int e = threadIdx.x; e < kin * kout; e += blockDim.x) {
    const int i = e / kin, j = e % kin;  // output i, input j
    sr[i * KROT + j] = transpose ? r[j * kout + i] : r[i * kin + j];
  }
  __syncthreads();
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    double xv[KROT];
#pragma unroll
    for (int j = 0; j < KROT; ++j) xv[j] = (j < kin) ? x[(long)j * n + p] : 0.0;
    for (int i = 0; i < kout; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < KROT; ++j)
        if (j < kin) s += sr[i * KROT + j] * xv[j];
      y[(long)i * n + p] = s;
    }
  }
}

// The rotation y = R^T x (k x k, R on the device) fused with the weighted Gram matrix of its output,
// sum_p q_p y_ip y_jp: the rotation is rotate_kernel's (same sums), the products and block partials
// are wgram_partial_kernel's for a Gram matrix at the same launch (same points per thread, same
// order), so y and the Gram matrix are bitwise those of the two separate kernels, with y read once.
template <int K>
__global__ __launch_bounds__(RB) void rotate_gram_kernel(const double* __restrict__ r, const double* __restrict__ x,
                                                         double* __restrict__ y, const double* __restrict__ q, long n,
                                                         double* __restrict__ partials) {
  __shared__ double sr[KROT * KROT];
  for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
    const int i = e / K, j = e % K;  // output i, input j
    sr[i * KROT + j] = r[j * K + i];
  }
  __syncthreads();
  double acc[K * K];
#pragma unroll
  for (int i = 0; i < K * K; ++i) acc[i] = 0.0;
  const long stride = (long)gridDim.x * RB;
  constexpr int U = K * K <= 25 ? 4 : (K * K <= 36 ? 2 : 1);
  for (long p0 = (long)blockIdx.x * RB + threadIdx.x; p0 < n; p0 += U * stride) {
    double w[U], yv[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long p = p0 + u * stride;
      const bool in = p < n;
      w[u] = in ? (q ? q[p] : 1.0) : 0.0;
      double xv[KROT];
#pragma unroll
      for (int j = 0; j < KROT; ++j) xv[j] = (in && j < K) ? x[(long)j * n + p] : 0.0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < KROT; ++j)
          if (j < K) sum += sr[i * KROT + j] * xv[j];
        yv[u][i] = sum;
        if (in) y[(long)i * n + p] = sum;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (p0 + u * stride < n) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const double av = yv[u][i] * w[u];
#pragma unroll
          for (int j = 0; j < K; ++j) acc[i * K + j] += av * yv[u][j];
        }
      }
    }
  }
  double* row = partials + (long)blockIdx.x * KMAX * KMAX;
  block_partials<K * K>(acc, [&](int e, double v) { row[(e / K) * KMAX + e % K] = v; });
}

// ---- linear combination ----

struct Terms {
  const double* x[16];
  double c[16];
  int mode[16];  // 0: c; 1: c * dev[idx]; 2: c / dev[idx]
  int idx[16];
};

__device__ __forceinline__ double term_coef(const Terms& t, int i, const double* dev) {
  double c = t.c[i];
  if (t.mode[i] == 1) c = c * dev[t.idx[i]];
  if (t.mode[i] == 2) c = c / dev[t.idx[i]];
  return c;
}

__global__ __launch_bounds__(256) void lincomb_kernel(double* __restrict__ out, int nterms, Terms t,
                                                      const double* __restrict__ dev, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    double s = term_coef(t, 0, dev) * t.x[0][p];
    for (int i = 1; i < nterms; ++i) s += term_coef(t, i, dev) * t.x[i][p];
    out[p] = s;
  }
}

// A linear combination fused with the weighted dot products of its result that follow it (the
// Lanczos updates): out = sum_t c_t x_t (lincomb_kernel's coefficients and term order), then the block
// partials of sum_p q_p out_p b_jp for ND <= 2 operands b_j (b_j == out gives the squared norm), as
// entries (0, j) of the partial tile. The thread-to-point mapping, the products ((out q) b) and the
// block epilogue are wgram_partial_kernel<1, 1>'s, so the result and the sums are bitwise lincomb
// followed by separate wgram calls.
template <int ND>
__global__ __launch_bounds__(RB) void lincomb_dots_kernel(double* __restrict__ out, int nterms, Terms t,
                                                          const double* __restrict__ dev, long n,
                                                          const double* __restrict__ q, const double* b0,
                                                          const double* b1, double* __restrict__ partials) {
  double acc[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j) acc[j] = 0.0;
  const double* bs[2] = {b0, b1};
  const long stride = (long)gridDim.x * RB;
  constexpr int U = 4;  // wgram_partial_kernel<1, 1>'s points per pass
  for (long p0 = (long)blockIdx.x * RB + threadIdx.x; p0 < n; p0 += U * stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long p = p0 + u * stride;
      if (p < n) {
        double v = term_coef(t, 0, dev) * t.x[0][p];
        for (int i = 1; i < nterms; ++i) v += term_coef(t, i, dev) * t.x[i][p];
        out[p] = v;
        const double av = v * (q ? q[p] : 1.0);
#pragma unroll
        for (int j = 0; j < ND; ++j) acc[j] += av * (bs[j] == out ? v : bs[j][p]);
      }
    }
  }
  double* row = partials + (long)blockIdx.x * KMAX * KMAX;
  block_partials<ND>(acc, [&](int j, double v) { row[j] = v; });
}

// ---- residual norms ----

__global__ __launch_bounds__(RB) void resnorm_partial_kernel(const double* __restrict__ av,
                                                             const double* __restrict__ v,
                                                             const double* __restrict__ evals,
                                                             const double* __restrict__ q, int k,
                                                             long n, double* __restrict__ partials) {
  double acc[KMAX];
  double e[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    acc[i] = 0.0;
    e[i] = (i < k) ? evals[i] : 0.0;
  }
  const long stride = (long)gridDim.x * RB;
  for (long p = (long)blockIdx.x * RB + threadIdx.x; p < n; p += stride) {
    const double w = q ? q[p] : 1.0;
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      if (i < k) {
        const double r = av[(long)i * n + p] - e[i] * v[(long)i * n + p];
        acc[i] += (r * r) * w;
      }
    }
  }
  double* row = partials + (long)blockIdx.x * KMAX * KMAX;
  block_partials<KMAX>(acc, [&](int i, double v) {
    if (i < k) row[i * KMAX] = v;
  });
}

// ---- small dense kernels (one thread: k is the subspace size, <= KROT) ----

__global__ __launch_bounds__(64) void chol_inv_kernel(const double* __restrict__ s, int k, double* __restrict__ t,
                                int* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l[KROT * KROT];
  int ok = 1;
  for (int i = 0; i < k * k; ++i) l[i] = 0.0;
  for (int j = 0; j < k; ++j) {
    double d = s[j * k + j];
    for (int m = 0; m < j; ++m) d -= l[j * k + m] * l[j * k + m];
    if (!(d > 0.0)) {
      ok = 0;
      break;
    }
    const double ljj = sqrt(d);
    l[j * k + j] = ljj;
    for (int i = j + 1; i < k; ++i) {
      double v = s[i * k + j];
      for (int m = 0; m < j; ++m) v -= l[i * k + m] * l[j * k + m];
      l[i * k + j] = v / ljj;
    }
  }
  // T = L^{-1}, forward substitution column by column of the identity.
  for (int i = 0; i < k * k; ++i) t[i] = 0.0;
  if (ok) {
    for (int c = 0; c < k; ++c) {
      for (int i = c; i < k; ++i) {
        double v = (i == c) ? 1.0 : 0.0;
        for (int m = c; m < i; ++m) v -= l[i * k + m] * t[m * k + c];
        t[i * k + c] = v / l[i * k + i];
      }
    }
  }
  flag[0] = ok ? 0 : 1;
}

// The Jacobi rotation (cs, sn) that zeroes apq in the symmetric 2 x 2 block [[app, apq], [apq, aqq]]:
// t = tan(angle) is the root of smaller magnitude of t^2 + 2 theta t - 1 = 0 with
// theta = (aqq - app) / (2 apq), i.e. t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)). Multiplying
// through by 2 |apq| gives t = 2 apq sgn(d) / (|d| + sqrt(d^2 + 4 apq^2)), d = aqq - app: one square
// root and one divide where theta's form takes two of each, then cs = rsqrt(1 + t^2), sn = t cs.
// These long-latency operations are the critical path of the one-thread kernels below. When
// d = 0 and apq^2 underflows, the denominator is zero and t = 1 (the 45-degree rotation of d = 0).
// The rotated block is diagonal with app - t apq and aqq + t apq (Golub and Van Loan, the
// symmetric Schur decomposition), and rows r != p, q of columns p and q rotate by (cs, sn); the
// kernels write each rotated entry to both of its symmetric places.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double& cs, double& sn,
                                                double& t) {
  const double d = aqq - app;
  const double den = fabs(d) + sqrt(d * d + 4.0 * apq * apq);
  t = den > 0.0 ? (d >= 0.0 ? 2.0 * apq : -2.0 * apq) / den : 1.0;
  cs = rsqrt(t * t + 1.0);
  sn = t * cs;
}

// Cyclic Jacobi on 0.5 (M + M^T). Writes eigenvalues ascending and C with eigenvectors in columns
// (C[i*k + j] = component i of eigenvector j), each column's largest entry made positive.
__global__ __launch_bounds__(64) void eigh_kernel(const double* __restrict__ m_in, int k, double* __restrict__ evals,
                            double* __restrict__ c_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a[KROT * KROT], v[KROT * KROT];
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      a[i * k + j] = 0.5 * (m_in[i * k + j] + m_in[j * k + i]);
      v[i * k + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < k; ++i) {
      diag += a[i * k + i] * a[i * k + i];
      for (int j = i + 1; j < k; ++j) off += a[i * k + j] * a[i * k + j];
    }
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int p = 0; p < k - 1; ++p) {
      for (int q = p + 1; q < k; ++q) {
        const double apq = a[p * k + q];
        if (apq == 0.0) continue;
        const double app = a[p * k + p], aqq = a[q * k + q];
        double cs, sn, t;
        jacobi_rotation(app, aqq, apq, cs, sn, t);
        for (int r = 0; r < k; ++r) {
          if (r == p || r == q) continue;
          const double arp = a[r * k + p], arq = a[r * k + q];
          a[r * k + p] = a[p * k + r] = cs * arp - sn * arq;
          a[r * k + q] = a[q * k + r] = sn * arp + cs * arq;
        }
        a[p * k + p] = app - t * apq;
        a[q * k + q] = aqq + t * apq;
        a[p * k + q] = a[q * k + p] = 0.0;
        for (int r = 0; r < k; ++r) {
          const double vrp = v[r * k + p], vrq = v[r * k + q];
          v[r * k + p] = cs * vrp - sn * vrq;
          v[r * k + q] = sn * vrp + cs * vrq;
        }
      }
    }
  }
  int order[KROT];
  for (int i = 0; i < k; ++i) order[i] = i;
  for (int i = 1; i < k; ++i) {  // insertion sort, stable
    int o = order[i];
    int j = i - 1;
    while (j >= 0 && a[order[j] * k + order[j]] > a[o * k + o]) {
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = o;
  }
  for (int jj = 0; jj < k; ++jj) {
    const int j = order[jj];
    evals[jj] = a[j * k + j];
    int big = 0;
    for (int i = 1; i < k; ++i)
      if (fabs(v[i * k + j]) > fabs(v[big * k + j])) big = i;
    const double sgn = v[big * k + j] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < k; ++i) c_out[i * k + jj] = sgn * v[i * k + j];
  }
}

// The same two kernels for a compile-time K <= 8: every index is a constant, so the matrices live
// in registers (the generic versions above keep them in per-lane scratch). These one-thread kernels
// are launched with 64 threads and say so (__launch_bounds__(64)): without it the compiler budgets
// registers for 1024-thread blocks, caps them at 128 and spills the K >= 6 matrices to scratch (the
// K = 8 Jacobi took 370 us instead of 90).
template <int K>
__global__ __launch_bounds__(64) void chol_inv_small_kernel(const double* __restrict__ s, double* __restrict__ t, int* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double l[K][K], ti[K][K];
  bool ok = true;
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) l[i][j] = ti[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double d = s[j * K + j];
#pragma unroll
    for (int m = 0; m < j; ++m) d -= l[j][m] * l[j][m];
    ok = ok && (d > 0.0);
    const double ljj = sqrt(d > 0.0 ? d : 1.0);
    l[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < K; ++i) {
      double v = s[i * K + j];
#pragma unroll
      for (int m = 0; m < j; ++m) v -= l[i][m] * l[j][m];
      l[i][j] = v / ljj;
    }
  }
#pragma unroll
  for (int c = 0; c < K; ++c)
#pragma unroll
    for (int i = c; i < K; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int m = c; m < i; ++m) v -= l[i][m] * ti[m][c];
      ti[i][c] = v / l[i][i];
    }
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) t[i * K + j] = ok ? ti[i][j] : 0.0;
  flag[0] = ok ? 0 : 1;
}

template <int K>
__global__ __launch_bounds__(64) void eigh_small_kernel(const double* __restrict__ m_in, double* __restrict__ evals, double* __restrict__ c_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a[K][K], v[K][K];
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) {
      a[i][j] = 0.5 * (m_in[i * K + j] + m_in[j * K + i]);
      v[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      diag += a[i][i] * a[i][i];
#pragma unroll
      for (int j = i + 1; j < K; ++j) off += a[i][j] * a[i][j];
    }
    if (off <= 1e-36 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < K - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < K; ++q) {
        const double apq = a[p][q];
        if (apq != 0.0) {
          const double app = a[p][p], aqq = a[q][q];
          double cs, sn, t;
          jacobi_rotation(app, aqq, apq, cs, sn, t);
#pragma unroll
          for (int r = 0; r < K; ++r) {
            if (r == p || r == q) continue;
            const double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = cs * arp - sn * arq;
            a[r][q] = a[q][r] = sn * arp + cs * arq;
          }
          a[p][p] = app - t * apq;
          a[q][q] = aqq + t * apq;
          a[p][q] = a[q][p] = 0.0;
#pragma unroll
          for (int r = 0; r < K; ++r) {
            const double vrp = v[r][p], vrq = v[r][q];
            v[r][p] = cs * vrp - sn * vrq;
            v[r][q] = sn * vrp + cs * vrq;
          }
        }
      }
    }
  }
  double e[K];
#pragma unroll
  for (int i = 0; i < K; ++i) e[i] = a[i][i];
  // Stable ascending sort of (e, columns of v) by an unrolled exchange network.
#pragma unroll
  for (int pass = 0; pass < K; ++pass)
#pragma unroll
    for (int j = 0; j + 1 < K; ++j)
      if (e[j] > e[j + 1]) {
        const double te = e[j];
        e[j] = e[j + 1];
        e[j + 1] = te;
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const double tv = v[r][j];
          v[r][j] = v[r][j + 1];
          v[r][j + 1] = tv;
        }
      }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    evals[j] = e[j];
    double big = v[0][j];
#pragma unroll
    for (int i = 1; i < K; ++i)
      if (fabs(v[i][j]) > fabs(big)) big = v[i][j];
    const double sgn = big < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int i = 0; i < K; ++i) c_out[i * K + j] = sgn * v[i][j];
  }
}

// The filter's Chebyshev table and the orthonormality error, one thread each (bodies in ritz_tail.h).
__global__ __launch_bounds__(64) void cheb_table_kernel(const double* __restrict__ evals, int index, const double* __restrict__ bounds,
                                  int degree, double* __restrict__ table) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  cheb_table_rows(evals, index, bounds, degree, table);
}

__global__ __launch_bounds__(64) void max_abs_minus_identity_kernel(const double* __restrict__ m, int k, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  max_abs_minus_identity(m, k, out);
}

}  // namespace cdft

using namespace cdft;

namespace cdft {

// The weighted Gram product of cdft_wgram_f64 with its block partials in `partials`.
static int wgram_impl(const double* a, int ka, const double* b, int kb, const double* q, long n, double scale,
                      int post, double* partials, long partials_len, double* out, hipStream_t s) {
  if (!a || !b || !out || !partials) return fail(CDFT_E_ARG, "wgram: null pointer argument");
  if (ka < 1 || kb < 1 || n < 1) return fail(CDFT_E_ARG, "wgram: sizes must be positive");
  if (post != 0 && post != 1) return fail(CDFT_E_ARG, "wgram: post must be 0 or 1");
  const int blocks = reduction_blocks(n);
  if (partials_len < (long)blocks * KMAX * KMAX)
    return fail(CDFT_E_CAPACITY, "wgram: partials workspace holds %ld doubles, needs %ld", partials_len,
                (long)blocks * KMAX * KMAX);
  for (int i0 = 0; i0 < ka; i0 += KMAX) {
    for (int j0 = 0; j0 < kb; j0 += KMAX) {
      const int ti = ka - i0 < KMAX ? ka - i0 : KMAX;
      const int tj = kb - j0 < KMAX ? kb - j0 : KMAX;
      const double* ap = a + (long)i0 * n;
      const double* bp = b + (long)j0 * n;
      if (ti == tj && ti <= 8) {
        switch (ti) {
          case 1: hipLaunchKernelGGL((wgram_partial_kernel<1, 1>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 2: hipLaunchKernelGGL((wgram_partial_kernel<2, 2>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 3: hipLaunchKernelGGL((wgram_partial_kernel<3, 3>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 4: hipLaunchKernelGGL((wgram_partial_kernel<4, 4>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 5: hipLaunchKernelGGL((wgram_partial_kernel<5, 5>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 6: hipLaunchKernelGGL((wgram_partial_kernel<6, 6>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          case 7: hipLaunchKernelGGL((wgram_partial_kernel<7, 7>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
          default: hipLaunchKernelGGL((wgram_partial_kernel<8, 8>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials); break;
        }
      } else {
        hipLaunchKernelGGL((wgram_partial_kernel<KMAX, KMAX>), dim3(blocks), dim3(RB), 0, s, ap, ti, bp, tj, q, n, partials);
      }
      int rc = check_launch("wgram_partial");
      if (rc) return rc;
      hipLaunchKernelGGL(wgram_final_kernel, dim3(ti * tj), dim3(64), 0, s, partials, blocks, ti, tj, scale,
                         post, out, kb, i0, j0);
      rc = check_launch("wgram_final");
      if (rc) return rc;
    }
  }
  return CDFT_OK;
}

// ---- the allocation-free Gram product ----
//
// Its block partials live in `scratch`, a static device array of the library: one copy per GPU,
// loaded with the code object and held until the process ends (512 KiB). The host-side name of the
// array is not a device address; its address on each device is looked up once, with that device
// current, by hipGetSymbolAddress and kept per device. Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU (two streams, or two host threads) race on it.
constexpr long SCRATCH_DOUBLES = (long)MAX_PARTIALS * KMAX * KMAX;
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[SCRATCH_DEVICES] = {};

// The device address of `scratch` on the device `stream` belongs to (the current device for the null
// stream).
static int scratch_for_stream(hipStream_t stream, double** ptr) {
  int device = 0;
  hipError_t e = stream ? hipStreamGetDevice(stream, &device) : hipGetDevice(&device);
  if (e != hipSuccess) return fail((int)e, "wgram_noalloc: cannot find the stream's device: %s", hipGetErrorString(e));
  if (device < 0 || device >= SCRATCH_DEVICES)
    return fail(CDFT_E_ARG, "wgram_noalloc: device %d is beyond the %d devices the library keeps scratch for", device,
                SCRATCH_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    int current = 0;
    e = hipGetDevice(&current);
    if (e != hipSuccess) return fail((int)e, "wgram_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (current != device && (e = hipSetDevice(device)) != hipSuccess)
      return fail((int)e, "wgram_noalloc: hipSetDevice(%d): %s", device, hipGetErrorString(e));
    void* address = nullptr;
    const hipError_t found = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
    if (current != device && (e = hipSetDevice(current)) != hipSuccess)
      return fail((int)e, "wgram_noalloc: hipSetDevice(%d): %s", current, hipGetErrorString(e));
    if (found != hipSuccess || !address)
      return fail(found != hipSuccess ? (int)found : CDFT_E_LAUNCH, "wgram_noalloc: hipGetSymbolAddress(scratch): %s",
                  hipGetErrorString(found));
    g_scratch_ptr[device] = (double*)address;
  }
  *ptr = g_scratch_ptr[device];
  return CDFT_OK;
}

}  // namespace cdft

extern "C" int cdft_wgram_f64(const double* a, int ka, const double* b, int kb, const double* q,
                              long n, double scale, int post, double* partials, long partials_len,
                              double* out, void* stream) {
  return wgram_impl(a, ka, b, kb, q, n, scale, post, partials, partials_len, out, (hipStream_t)stream);
}

extern "C" int cdft_wgram_noalloc_f64(const double* a, int ka, const double* b, int kb, const double* q, long n,
                                      double scale, int post, double* out, void* stream) {
  if (!a || !b || !out) return fail(CDFT_E_ARG, "wgram_noalloc: null pointer argument");
  if (ka < 1 || kb < 1 || n < 1) return fail(CDFT_E_ARG, "wgram_noalloc: sizes must be positive");
  // The reduction takes at most MAX_PARTIALS blocks (its grid-stride loop covers any n), each with one
  // KMAX x KMAX tile of partial sums; a request beyond the static scratch is refused, not truncated.
  const long need = (long)reduction_blocks(n) * KMAX * KMAX;
  if (need > SCRATCH_DOUBLES)
    return fail(CDFT_E_CAPACITY, "wgram_noalloc: %ld partial sums exceed the static scratch of %ld doubles", need,
                SCRATCH_DOUBLES);
  double* partials = nullptr;
  const int rc = scratch_for_stream((hipStream_t)stream, &partials);
  if (rc) return rc;
  return wgram_impl(a, ka, b, kb, q, n, scale, post, partials, SCRATCH_DOUBLES, out, (hipStream_t)stream);
}

// y = R^T x and gram = sum_p q_p y_ip y_jp (k <= 8; q may be NULL): bitwise cdft_rotate_f64 (transpose)
// followed by cdft_wgram_f64(y, k, y, k, q, n, 1, 0, ..).
extern "C" int cdft_rotate_gram_f64(const double* r, int k, const double* x, double* y, const double* q, long n,
                                    double* partials, long partials_len, double* gram, void* stream) {
  if (!r || !x || !y || !partials || !gram) return fail(CDFT_E_ARG, "rotate_gram: null pointer argument");
  if (k < 1 || k > KMAX || n < 1) return fail(CDFT_E_ARG, "rotate_gram: k must be in 1..%d", KMAX);
  if (x == y) return fail(CDFT_E_ARG, "rotate_gram: y must not alias x");
  const int blocks = reduction_blocks(n);
  if (partials_len < (long)blocks * KMAX * KMAX) return fail(CDFT_E_CAPACITY, "rotate_gram: partials too small");
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
    case 1: hipLaunchKernelGGL(rotate_gram_kernel<1>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 2: hipLaunchKernelGGL(rotate_gram_kernel<2>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 3: hipLaunchKernelGGL(rotate_gram_kernel<3>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 4: hipLaunchKernelGGL(rotate_gram_kernel<4>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 5: hipLaunchKernelGGL(rotate_gram_kernel<5>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 6: hipLaunchKernelGGL(rotate_gram_kernel<6>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    case 7: hipLaunchKernelGGL(rotate_gram_kernel<7>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
    default: hipLaunchKernelGGL(rotate_gram_kernel<8>, dim3(blocks), dim3(RB), 0, s, r, x, y, q, n, partials); break;
  }
  int rc = check_launch("rotate_gram");
  if (rc) return rc;
  hipLaunchKernelGGL(wgram_final_kernel, dim3(k * k), dim3(64), 0, s, partials, blocks, k, k, 1.0, 0, gram, k, 0, 0);
  return check_launch("rotate_gram_final");
}

// out = sum_t c_t x_t (as cdft_lincomb_f64, nterms <= 16, out may not alias the b operands unless it is
// one of them) and dots[j] = post(sum_p q_p out_p b_j[p]) for ndots <= 2 (b_j == out: the squared
// norm; post 1 = sqrt), bitwise cdft_lincomb_f64 followed by ndots cdft_wgram_f64(out, 1, b_j, 1, ...)
// calls; dots are written to dots[0], dots[1].
extern "C" int cdft_lincomb_dots_f64(double* out, int nterms, const double* const* x, const double* c, const int* mode,
                                     const int* idx, const double* dev, long n, const double* q, int ndots,
                                     const double* const* b, int post, double* partials, long partials_len,
                                     double* dots, void* stream) {
  if (!out || !x || !c || !mode || !idx || !b || !dots || !partials || nterms < 1 || nterms > 16 || n < 1 ||
      ndots < 1 || ndots > 2 || (post != 0 && post != 1))
    return fail(CDFT_E_ARG, "lincomb_dots: bad arguments (nterms=%d, ndots=%d)", nterms, ndots);
  Terms t{};
  for (int i = 0; i < nterms; ++i) {
    if (!x[i]) return fail(CDFT_E_ARG, "lincomb_dots: term %d has no operand", i);
    if ((mode[i] == 1 || mode[i] == 2) && !dev) return fail(CDFT_E_ARG, "lincomb_dots: device coefficients missing");
    t.x[i] = x[i];
    t.c[i] = c[i];
    t.mode[i] = mode[i];
    t.idx[i] = idx[i];
  }
  for (int j = 0; j < ndots; ++j)
    if (!b[j]) return fail(CDFT_E_ARG, "lincomb_dots: dot operand %d missing", j);
  const int blocks = reduction_blocks(n);
  if (partials_len < (long)blocks * KMAX * KMAX) return fail(CDFT_E_CAPACITY, "lincomb_dots: partials too small");
  hipStream_t s = (hipStream_t)stream;
  if (ndots == 1)
    hipLaunchKernelGGL(lincomb_dots_kernel<1>, dim3(blocks), dim3(RB), 0, s, out, nterms, t, dev, n, q, b[0], b[0],
                       partials);
  else
    hipLaunchKernelGGL(lincomb_dots_kernel<2>, dim3(blocks), dim3(RB), 0, s, out, nterms, t, dev, n, q, b[0], b[1],
                       partials);
  int rc = check_launch("lincomb_dots");
  if (rc) return rc;
  hipLaunchKernelGGL(wgram_final_kernel, dim3(ndots), dim3(64), 0, s, partials, blocks, 1, ndots, 1.0, post, dots,
                     ndots, 0, 0);
  return check_launch("lincomb_dots_final");
}

extern "C" long cdft_reduction_workspace_doubles(long n) {
  return (long)reduction_blocks(n) * KMAX * KMAX;
}

extern "C" int cdft_rotate_f64(const double* r, int kin, int kout, int transpose, const double* x,
                               double* y, long n, void* stream) {
  if (!r || !x || !y) return fail(CDFT_E_ARG, "rotate: null pointer argument");
  if (kin < 1 || kout < 1 || kin > KROT || kout > KROT || n < 1)
    return fail(CDFT_E_ARG, "rotate: kin, kout must be in 1..%d (got %d, %d)", KROT, kin, kout);
  if (x == y) return fail(CDFT_E_ARG, "rotate: y must not alias x");
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(rotate_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, r, kin,
                     kout, transpose, x, y, n);
  return check_launch("rotate");
}

extern "C" int cdft_lincomb_f64(double* out, int nterms, const double* const* x, const double* c,
                                const int* mode, const int* idx, const double* dev, long n,
                                void* stream) {
  if (!out || !x || !c || nterms < 1 || nterms > 16 || n < 1)
    return fail(CDFT_E_ARG, "lincomb: bad arguments (nterms %d)", nterms);
  Terms t;
  for (int i = 0; i < nterms; ++i) {
    if (!x[i]) return fail(CDFT_E_ARG, "lincomb: null term %d", i);
    t.x[i] = x[i];
    t.c[i] = c[i];
    t.mode[i] = mode ? mode[i] : 0;
    t.idx[i] = idx ? idx[i] : 0;
    if (t.mode[i] != 0 && !dev) return fail(CDFT_E_ARG, "lincomb: device scalar requested without dev");
  }
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out,
                     nterms, t, dev, n);
  return check_launch("lincomb");
}

extern "C" int cdft_resnorm_f64(const double* av, const double* v, const double* evals,
                                const double* q, int k, long n, double* partials, long partials_len,
                                double* norms, void* stream) {
  if (!av || !v || !evals || !norms || !partials) return fail(CDFT_E_ARG, "resnorm: null pointer argument");
  if (k < 1 || k > KMAX || n < 1) return fail(CDFT_E_ARG, "resnorm: k must be in 1..%d", KMAX);
  const int blocks = reduction_blocks(n);
  if (partials_len < (long)blocks * KMAX * KMAX) return fail(CDFT_E_CAPACITY, "resnorm: partials too small");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(resnorm_partial_kernel, dim3(blocks), dim3(RB), 0, s, av, v, evals, q, k, n, partials);
  int rc = check_launch("resnorm_partial");
  if (rc) return rc;
  // Reuse the Gram final pass: entries (i, 0) of each partial tile, one column.
  hipLaunchKernelGGL(wgram_final_kernel, dim3(k), dim3(64), 0, s, partials, blocks, k, 1, 1.0, 1, norms, 1, 0, 0);
  return check_launch("resnorm_final");
}

extern "C" int cdft_chol_inv_f64(const double* s, int k, double* t, int* flag, void* stream) {
  if (!s || !t || !flag || k < 1 || k > KROT) return fail(CDFT_E_ARG, "chol_inv: bad arguments (k=%d)", k);
  hipStream_t st = (hipStream_t)stream;
  switch (k) {
    case 1: hipLaunchKernelGGL(chol_inv_small_kernel<1>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 2: hipLaunchKernelGGL(chol_inv_small_kernel<2>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 3: hipLaunchKernelGGL(chol_inv_small_kernel<3>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 4: hipLaunchKernelGGL(chol_inv_small_kernel<4>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 5: hipLaunchKernelGGL(chol_inv_small_kernel<5>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 6: hipLaunchKernelGGL(chol_inv_small_kernel<6>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 7: hipLaunchKernelGGL(chol_inv_small_kernel<7>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    case 8: hipLaunchKernelGGL(chol_inv_small_kernel<8>, dim3(1), dim3(64), 0, st, s, t, flag); break;
    default: hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(64), 0, st, s, k, t, flag);
  }
  return check_launch("chol_inv");
}

extern "C" int cdft_eigh_f64(const double* m, int k, double* evals, double* c, void* stream) {
  if (!m || !evals || !c || k < 1 || k > KROT) return fail(CDFT_E_ARG, "eigh: bad arguments (k=%d)", k);
  hipStream_t st = (hipStream_t)stream;
  switch (k) {
    case 1: hipLaunchKernelGGL(eigh_small_kernel<1>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 2: hipLaunchKernelGGL(eigh_small_kernel<2>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 3: hipLaunchKernelGGL(eigh_small_kernel<3>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 4: hipLaunchKernelGGL(eigh_small_kernel<4>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 5: hipLaunchKernelGGL(eigh_small_kernel<5>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 6: hipLaunchKernelGGL(eigh_small_kernel<6>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 7: hipLaunchKernelGGL(eigh_small_kernel<7>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    case 8: hipLaunchKernelGGL(eigh_small_kernel<8>, dim3(1), dim3(64), 0, st, m, evals, c); break;
    default: hipLaunchKernelGGL(eigh_kernel, dim3(1), dim3(64), 0, st, m, k, evals, c);
  }
  return check_launch("eigh");
}

extern "C" int cdft_cheb_table_f64(const double* evals, int index, const double* bounds, int degree, double* table,
                                   void* stream) {
  if (!evals || !bounds || !table || index < 0 || degree < 1) return fail(CDFT_E_ARG, "cheb_table: bad arguments");
  hipLaunchKernelGGL(cheb_table_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, evals, index, bounds, degree, table);
  return check_launch("cheb_table");
}

extern "C" int cdft_orth_error_f64(const double* m, int k, double* out, void* stream) {
  if (!m || !out || k < 1) return fail(CDFT_E_ARG, "orth_error: bad arguments");
  hipLaunchKernelGGL(max_abs_minus_identity_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, m, k, out);
  return check_launch("orth_error");
}
