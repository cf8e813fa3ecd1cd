/* cdft_hip.h -- C ABI of libcdft_hip.so, the MI355X (gfx950) Kohn-Sham hot path of cdft.
 *
 * Conventions for every entry point:
 *   - returns 0 on success, a negative CDFT_E_* code or a positive hipError_t otherwise;
 *     cdft_last_error() returns the calling thread's message for the last failure;
 *   - every array argument is a raw DEVICE pointer unless its comment says "host";
 *   - every kernel is enqueued on `stream` (a hipStream_t) and nothing synchronises the host;
 *   - blocks of k vectors are row-major (k, N), N = n0*n1*n2, C order with z (n2) fastest;
 *   - face-weight arrays w0, w1, w2 hold f^2 at the staggered faces of axis 0, 1, 2 on the box
 *     extended by `order` along that axis, indexed from face -order/2 (w0 is (n0+order, n1, n2));
 *   - `coef` (host) holds the `order` staggered weights c_k / h, k = -order/2 .. order/2 - 1;
 *   - no entry point allocates device memory on these paths; workspaces are passed in.
 */
#ifndef CDFT_HIP_H
#define CDFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDFT_E_ARG (-1)
#define CDFT_E_LAUNCH (-2)
#define CDFT_E_LIBRARY (-3)
#define CDFT_E_CAPACITY (-4)

/* Message of the calling thread's last failure. */
const char* cdft_last_error(void);
/* ABI version of this header (1). */
int cdft_abi_version(void);

/* The geometry records every stencil entry below also takes (`gp`): for plane m and column (y, z) of
 * the face-extended rows (y < n1 + order, z < n2 + order), four values (w0 at extended plane
 * m + order - 1, w1 at (m, y, z), w2 at (m, y, z), invw at (m, y, z); zero where the array has no
 * such entry), built from w0, w1, w2, invw by cdft_hdiv_pack_* into cdft_hdiv_pack_len(..) elements.
 * A plane's four per-point geometry values are then two 16-byte loads instead of four. Built once per
 * geometry (they do not involve pot). Part of the stencil interface below; no reference counterpart. */
long cdft_hdiv_pack_len(int n0, int n1, int n2, int order);
int cdft_hdiv_pack_f64(const double* w0, const double* w1, const double* w2, const double* invw, int n0,
                       int n1, int n2, int order, double* gp, void* stream);
int cdft_hdiv_pack_f32(const float* w0, const float* w1, const float* w2, const float* invw, int n0,
                       int n1, int n2, int order, float* gp, void* stream);

/* K1: out = 0.5 * invw * sum_d D_d^T [w_d D_d phi] + pot * phi for a block of k vectors.
 * Replaces CuspFactoredHamiltonian.apply / apply_with_potential / kinetic_apply
 * (src/cdft/operators/hamiltonian.py:251-318) and weighted_divergence
 * (src/cdft/operators/divergence.py:153-197, device tier fused.py:353-380). order in {2,4,6,8}, and 12
 * in float64 (the comparison order of the stencil-order gate, src/cdft/gates/tier3.py:290-352). */
int cdft_hdiv_apply_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                        const double* invw, const double* pot, const double* gp, double* out, int k,
                        int n0, int n1, int n2, const double* coef, int order, void* stream);
int cdft_hdiv_apply_f32(const float* phi, const float* w0, const float* w1, const float* w2,
                        const float* invw, const float* pot, const float* gp, float* out, int k,
                        int n0, int n1, int n2, const double* coef, int order, void* stream);

/* The Euclidean transpose of K1 for the backward pass: out = 0.5 sum_d D_d^T[w_d D_d (invw x)]
 * + pot x (invw applied on the input). work: a distinct (k, N) scratch block. No reference
 * counterpart (the reference's inversion is a stub, src/cdft/inversion/__init__.py:1; interface
 * InversionProtocol contract.py:1263-1283). */
int cdft_hdiv_apply_adjoint_f64(const double* x, const double* w0, const double* w1, const double* w2,
                                const double* invw, const double* pot, const double* gp, double* out,
                                double* work, int k, int n0, int n1, int n2, const double* coef, int order,
                                void* stream);
int cdft_hdiv_apply_adjoint_f32(const float* x, const float* w0, const float* w1, const float* w2,
                                const float* invw, const float* pot, const float* gp, float* out,
                                float* work, int k, int n0, int n1, int n2, const double* coef, int order,
                                void* stream);

/* K1 fused with the residual norms of Rayleigh-Ritz: norms[v] = ||A phi_v - evals[v] phi_v||_q
 * (sqrt of sum_p q_p r_p^2), with nothing stored but per-block partials (partials: at least the launch's
 * block count, cdft_hdiv_launch_shape(..)[5] doubles) combined in a fixed order. Replaces the apply and
 * the norm of residual_norms (src/cdft/eigen/rayleigh_ritz.py:105-115). */
int cdft_hdiv_resnorm_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                          const double* invw, const double* pot, const double* gp, const double* evals,
                          const double* q, int k,
                          int n0, int n1, int n2, const double* coef, int order, double* partials, long partials_len,
                          double* norms, void* stream);

/* cdft_hdiv_resnorm_f64 with the end of a Rayleigh-Ritz pass folded into its combine launch: also
 * builds the next filter's Chebyshev table from evals[index] and bounds (as cdft_cheb_table_f64) and
 * writes max |gram - I| of the k x k gram to orth[0] (as cdft_orth_error_f64), bitwise the separate
 * calls. No single reference function: it ends the CheFSI pass of src/cdft/eigen/chefsi.py:499-547
 * (residual_norms, the next filter's rows of chefsi.py:116-136, the orthonormality error). */
int cdft_hdiv_resnorm_tail_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                               const double* invw, const double* pot, const double* gp, const double* evals,
                               const double* q, int k, int n0, int n1, int n2, const double* coef, int order,
                               double* partials, long partials_len, double* norms, int index, const double* bounds,
                               int degree, double* table, const double* gram, double* orth, void* stream);

/* K2: one fused scaled-Chebyshev step, out = (A y - c y) a - b prev with (c, a, b) = table[step]
 * read on the device (table is a (degree, 3) float64 device array), prev ignored at step 0.
 * Replaces chebyshev_update (src/cdft/operators/fused.py:391-417) and the body of
 * chebyshev_recurrence (src/cdft/eigen/chefsi.py:139-160). */
int cdft_cheb_step_f64(const double* y, const double* prev, const double* w0, const double* w1,
                       const double* w2, const double* invw, const double* pot, const double* gp,
                       double* out, int k,
                       int n0, int n1, int n2, const double* coef, int order, const double* table,
                       int step, void* stream);
int cdft_cheb_step_f32(const float* y, const float* prev, const float* w0, const float* w1,
                       const float* w2, const float* invw, const float* pot, const float* gp,
                       float* out, int k,
                       int n0, int n1, int n2, const double* coef, int order, const double* table,
                       int step, void* stream);

/* K2: the whole degree-`degree` filter of x into out, rotating through work1/work2 (each (k, N);
 * x, out, work1, work2 distinct). Replaces chebyshev_filter (src/cdft/eigen/chefsi.py:81-113) and
 * the captured GraphedChebyshevFilter replay (chefsi.py:185-392). */
int cdft_cheb_filter_f64(const double* x, double* out, double* work1, double* work2,
                         const double* w0, const double* w1, const double* w2, const double* invw,
                         const double* pot, const double* gp, int k, int n0, int n1, int n2, const double* coef,
                         int order, const double* table, int degree, void* stream);
int cdft_cheb_filter_f32(const float* x, float* out, float* work1, float* work2, const float* w0,
                         const float* w1, const float* w2, const float* invw, const float* pot,
                         const float* gp, int k, int n0, int n1, int n2, const double* coef, int order,
                         const double* table, int degree, void* stream);

/* ---- the plain operator of the analytic model systems: H = -1/2 lap + V ----
 * Replaces LocalHamiltonian.apply / apply_with_potential (src/cdft/operators/hamiltonian.py:92-125)
 * with the grid's Laplacian (src/cdft/grid.py:714-740, src/cdft/operators/laplacian.py:85-105):
 * coef = the 2R + 1 central second-derivative weights c_j / h^2 (j = -R..R, order = 2R in 2..12),
 * accumulated per axis 0, 1, 2 from -R to +R; boundary 0 = zero halo, 1 = odd reflection about walls
 * one spacing outside the outermost points (grid.py:808-827; every axis needs >= R points). Blocks
 * (k, N) float64 on the device, C order. */
int cdft_local_apply_f64(const double* phi, const double* pot, double* out, int k, int n0, int n1, int n2,
                         const double* coef, int order, int boundary, void* stream);
/* The degree-`degree` scaled Chebyshev filter of x with this operator (as cdft_cheb_filter_f64;
 * src/cdft/eigen/chefsi.py:81-160). */
int cdft_local_filter_f64(const double* x, double* out, double* work1, double* work2, const double* pot, int k,
                          int n0, int n1, int n2, const double* coef, int order, int boundary, const double* table,
                          int degree, void* stream);
/* The doubles `partials` must hold for cdft_local_resnorm_f64 (-1 for bad sizes). */
long cdft_local_partials_len(int n0, int n1, int n2, int k);
/* norms[v] = ||H phi_v - evals[v] phi_v||_q (residual_norms, src/cdft/eigen/rayleigh_ritz.py); with
 * index >= 0 also the next filter's table and max |gram - I| in the combine's launch, as
 * cdft_hdiv_resnorm_tail_f64 (index < 0: bounds, table, gram and orth unused). */
int cdft_local_resnorm_f64(const double* phi, const double* pot, const double* evals, const double* q, int k, int n0,
                           int n1, int n2, const double* coef, int order, int boundary, double* partials,
                           long partials_len, double* norms, int index, const double* bounds, int degree,
                           double* table, const double* gram, double* orth, void* stream);

/* ---- batched grids: many independent operators in one launch ----
 * The batched scan of independent solves that the reference runs as separate processes or streams
 * (scripts/scan_throughput.py:212-313); no single reference function corresponds. */

/* Bytes of the device descriptor table of n_grids grids. */
long cdft_batch_desc_bytes(int n_grids);
/* Host-only planning: fills `desc` (host buffer of cdft_batch_desc_bytes bytes, to be copied to the
 * device unchanged) for n_grids grids of one stencil order and block width k. shapes (host,
 * n_grids x 3), coefs (host, n_grids x order staggered weights / h), ptrs (host, n_grids x 11 device
 * pointers: x, out, work1, work2, w0, w1, w2, invw, pot, geometry records, Chebyshev table (degree x 3)). launch (host,
 * 3 ints) receives {blocks, threads, tile}: tile > 0 is the column kernel's width (13 or 16), tile < 0
 * the vector kernel's (-16, -20, -22, -24 or -28), one for every grid of the batch. Order 8 runs on
 * the vector kernel unless cdft_hdiv_set_kernel(1) or a column-only width (13) is in force. */
int cdft_batch_plan_f64(int n_grids, const int* shapes, int k, int order, const double* coefs,
                        const void* const* ptrs, void* desc, long desc_bytes, int* launch);
/* K2 for every grid of a batch: each grid's degree-`degree` filter of x into out through work1 /
 * work2 with its own table, one launch per step for all grids (bitwise the single-grid filter). */
int cdft_cheb_filter_batched_f64(const void* desc, int n_grids, const int* launch, int order, int degree,
                                 void* stream);
/* K1 for every grid of a batch: out = A x. */
int cdft_hdiv_apply_batched_f64(const void* desc, int n_grids, const int* launch, int order, void* stream);
/* The float32 batch: the same descriptor layout and size, planned by cdft_batch_plan_f32 over float32
 * blocks and geometry (the coefficients and Chebyshev tables stay float64); its filter and apply are
 * the float32 single-grid kernels' steps, bitwise. */
int cdft_batch_plan_f32(int n_grids, const int* shapes, int k, int order, const double* coefs,
                        const void* const* ptrs, void* desc, long desc_bytes, int* launch);
int cdft_cheb_filter_batched_f32(const void* desc, int n_grids, const int* launch, int order, int degree,
                                 void* stream);
int cdft_hdiv_apply_batched_f32(const void* desc, int n_grids, const int* launch, int order, void* stream);

/* ---- K3 / K4: float64 reductions and rotations of Rayleigh-Ritz (deterministic, no atomics) ---- */

/* Doubles of block-partials workspace the reductions below need for length n. */
long cdft_reduction_workspace_doubles(long n);

/* out[i*kb + j] = scale * sum_p q[p] a[i*n + p] b[j*n + p] (q may be NULL for unit weights);
 * post = 1 stores sqrt(max(x, 0)). Replaces Measure.inner/norm/gram/cross
 * (src/cdft/eigen/measure.py:72-96) and orthonormality_error_tensor (eigen/rayleigh_ritz.py:44-48). */
int cdft_wgram_f64(const double* a, int ka, const double* b, int kb, const double* q, long n,
                   double scale, int post, double* partials, long partials_len, double* out,
                   void* stream);

/* cdft_wgram_f64 without a workspace argument and without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The block partials live in a static device array of
 * the library (512 KiB per GPU, held until the process ends; at most 1024 blocks x 8 x 8 partial sums,
 * which covers every n, since the reduction's grid-stride loop caps its blocks at 1024); a request that
 * needed more would return CDFT_E_CAPACITY. Results are bitwise those of cdft_wgram_f64. Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (two streams, two host threads) race. */
int cdft_wgram_noalloc_f64(const double* a, int ka, const double* b, int kb, const double* q, long n,
                           double scale, int post, double* out, void* stream);

/* y = R^T x (k x k, k <= 8) fused with gram = sum_p q_p y_ip y_jp (q may be NULL), bitwise cdft_rotate_f64
 * (transpose = 1) followed by cdft_wgram_f64 of y with itself: the rotation of the Ritz vectors and
 * their orthonormality Gram matrix (src/cdft/eigen/rayleigh_ritz.py:95-102). */
int cdft_rotate_gram_f64(const double* r, int k, const double* x, double* y, const double* q, long n,
                         double* partials, long partials_len, double* gram, void* stream);

/* y[i] = sum_j R[i][j] x[j] (R kout x kin row-major on the device; transpose = 1 uses R^T),
 * kin, kout <= 16. Replaces the rotations `rotation.T @ basis` and the triangular solve of
 * orthonormalise (src/cdft/eigen/rayleigh_ritz.py:51-102, precision.py:292-308). */
int cdft_rotate_f64(const double* r, int kin, int kout, int transpose, const double* x, double* y,
                    long n, void* stream);

/* out = sum_t c_t x_t (x: host array of nterms device pointers; c, mode, idx: host arrays; mode 1
 * multiplies c_t by dev[idx_t], mode 2 divides). nterms <= 16; out may alias a term. Replaces the
 * Lanczos updates (src/cdft/operators/hamiltonian.py:55-65) and the Pulay combination
 * (src/cdft/scf/mixing.py:158-162). */
int cdft_lincomb_f64(double* out, int nterms, const double* const* x, const double* c,
                     const int* mode, const int* idx, const double* dev, long n, void* stream);

/* cdft_lincomb_f64 into out, then dots[j] = post(sum_p q_p out_p b_j[p]) for ndots <= 2 operands b (host
 * array of device pointers; b[j] == out gives the squared norm; post 1 = sqrt; q may be NULL), in one
 * pass with one final launch: bitwise the lincomb followed by the ndots cdft_wgram_f64 calls. The
 * Lanczos update and its reorthogonalisation / normalisation products (src/cdft/operators/hamiltonian.py:
 * 55-65). */
int cdft_lincomb_dots_f64(double* out, int nterms, const double* const* x, const double* c, const int* mode,
                          const int* idx, const double* dev, long n, const double* q, int ndots,
                          const double* const* b, int post, double* partials, long partials_len,
                          double* dots, void* stream);

/* norms[i] = sqrt(max(sum_p q (av_i - evals_i v_i)^2, 0)), k <= 8. Replaces residual_norms
 * (src/cdft/eigen/rayleigh_ritz.py:105-115). */
int cdft_resnorm_f64(const double* av, const double* v, const double* evals, const double* q, int k,
                     long n, double* partials, long partials_len, double* norms, void* stream);

/* T = L^{-1} with S = L L^T (k <= 16); flag[0] = 1 if S is not positive definite. Replaces the
 * Cholesky step of orthonormalise (src/cdft/eigen/rayleigh_ritz.py:63-72). */
int cdft_chol_inv_f64(const double* s, int k, double* t, int* flag, void* stream);

/* Symmetric eigendecomposition of 0.5 (M + M^T) (k <= 16) by cyclic Jacobi: evals ascending,
 * eigenvectors in the columns of c. Replaces torch.linalg.eigh of the Rayleigh-Ritz matrix
 * (src/cdft/eigen/rayleigh_ritz.py:96-99). */
int cdft_eigh_f64(const double* m, int k, double* evals, double* c, void* stream);

/* out[0] = max |M - I| over a k x k matrix (src/cdft/eigen/rayleigh_ritz.py:44-48). */
int cdft_orth_error_f64(const double* m, int k, double* out, void* stream);
/* The (degree, 3) scaled-Chebyshev table of a filter built on the device, bitwise the host rows of
 * chebyshev_coefficients (src/cdft/eigen/chefsi.py:116-136): lower = evals[index] clamped into
 * (bounds[0] + 1e-8, bounds[1] - 1e-8), upper = bounds[1], spectrum_min = bounds[0] (device arrays). A
 * filter graph that starts with it needs no per-pass table upload. */
int cdft_cheb_table_f64(const double* evals, int index, const double* bounds, int degree, double* table,
                        void* stream);

/* ---- stencil launch control (no reference counterpart: tuning of the K1/K2 launch) ---- */

/* Planes per x-segment of the marching stencil kernels (0 = automatic choice). Process-wide. */
int cdft_hdiv_set_segment(int planes);

/* Tile extents of the marching stencil kernels for later launches and batch plans (0 = automatic):
 * the largest tile extent along y, 1..16 (the tiles then split n1 evenly), and the tile width along z:
 * 13 or 16 for the column kernel, and for an order-8 batch plan also a vector-kernel width (16, 20,
 * 22, 24 or 28). Outputs are bitwise independent of the tiling (and of the segment length); a launch
 * tuning knob with no counterpart in the reference, whose stencil is torch slicing
 * (src/cdft/operators/divergence.py:84-197). Process-wide, like cdft_hdiv_set_segment. */
int cdft_hdiv_set_tile(int tile_y, int tile_z);
/* Host-only: the launch the column stencil kernel uses for a (k, n0, n1, n2) block of the given order
 * (the transpose, the residual-norm epilogue, the batched launches, and every order but 8):
 * shape6 (host) = {tile_y, tile_z, longest segment, segments, threads per block, blocks}. */
int cdft_hdiv_launch_shape(int n0, int n1, int n2, int k, int order, int* shape6);

/* The kernel of the order-8 apply and Chebyshev steps (cdft_hdiv_apply_*, cdft_cheb_step_*,
 * cdft_cheb_filter_*): 0 = the vector kernel (each thread owns a 16-byte vector of adjacent
 * z-columns; the default), 1 = the column kernel (one column per thread). Bitwise the same outputs;
 * process-wide. No reference counterpart (a launch choice of the K1/K2 stencil). */
int cdft_hdiv_set_kernel(int which);
/* Launch settings of the vector kernel for one problem shape: blocks of k vectors on an n0 x n1 x n2
 * grid of elem_bytes-byte values (8 or 4) use tiles of at most tile_y rows (1..17; the tiles then
 * split n1 evenly), tile_z columns (16, 20, 22, 24 or 28 in float64; 16, 20, 24 or 28 in float32)
 * and `segment` planes per x-segment (0 = the CU-fill rule). tile_y = 0 removes the entry. A
 * process-wide table of 64 entries (the oldest makes room), read at every launch, so a setting
 * applies to graphs captured after it; outputs do not depend on it. No reference counterpart. */
int cdft_hdiv_vec_set(int n0, int n1, int n2, int k, int elem_bytes, int tile_y, int tile_z, int segment);
/* Host-only: the vector kernel's launch for that shape: shape6 (host) = {tile_y, tile_z, longest
 * segment, segments, threads per block, blocks}. */
int cdft_hdiv_vec_shape(int n0, int n1, int n2, int k, int elem_bytes, int* shape6);

/* ---- K7: geometry of the cusp factor and the cusp-aware quadrature (once per geometry) ----
 * Grid points sit at origin + h * (i0, i1, i2); shape (host, 3 ints) = (n0, n1, n2); origin (host,
 * 3 doubles). charges (host, n_atoms) and positions (host, n_atoms x 3) describe the nuclei. */

/* u = sum_a Z_a |r - R_a|, grad u (3, N), weight = f^2 = exp(-2u), and the transformed potential
 * W = -|grad u|^2 / 2 cell-averaged with the 4^3 Gauss-Legendre rule (gl_offsets: 4 host node
 * offsets in units of h, gl_weights: 64 host weights, gl_total their sum) at points within `reach`
 * of a nucleus. Replaces CuspFactor.__init__ / _build_potential (src/cdft/operators/cusp.py:46-145). */
int cdft_cusp_grid_fields_f64(int n_atoms, const double* charges, const double* positions,
                              const int* shape, double h, const double* origin, double reach,
                              const double* gl_offsets, const double* gl_weights, double gl_total,
                              double* u, double* grad_u, double* weight, double* wpot, void* stream);

/* exp(-2u) at the staggered faces of each axis on the box extended by `order` along that axis
 * (w0, w1, w2 as in the conventions above); umax (device, zeroed by the caller) receives the bits
 * of max u so the host can refuse an underflowing factor. Replaces CuspFactor.face_weights
 * (src/cdft/operators/cusp.py:168-214). */
int cdft_face_weights_f64(int n_atoms, const double* charges, const double* positions,
                          const int* shape, double h, const double* origin, int order, double* w0,
                          double* w1, double* w2, unsigned long long* umax, void* stream);

/* Host-only: the quadrature's node table. For n_shells composite Gauss-Legendre shells (radius r,
 * radial weight wr, direction count n_theta >= 4 and owning atom; centres 3 per atom) places the
 * 2 n_theta^2 nodes of each shell at r d + centre with bare weight wr dw (each product and sum rounded
 * on its own), d and dw the unit directions and weights of count n (rows dir_x/y/z and dir_w from
 * dir_offset[n], n <= max_theta), and numbers all m nodes by their base cell floor((p - origin) /
 * spacing) with a stable counting sort. Outputs in that order: points (3, m) rows, owners, bare
 * weights, base cells (3, m) rows, perm (sorted position -> enumeration index), the sorted cell keys
 * and box6 = {lo, ext} of the base cells. Replaces the node construction of CuspQuadrature
 * (src/cdft/operators/quadrature.py:129-180, sphere_nodes); the base-cell numbering has no
 * counterpart (the reference stores a CSR interpolation matrix instead). */
int cdft_node_table(int n_shells, const double* r, const double* wr, const int* n_theta, const int* atom,
                    const double* centres, double spacing, const double* origin, int max_theta,
                    const long* dir_offset, const double* dir_x, const double* dir_y, const double* dir_z,
                    const double* dir_w, long m, double* points, int* owners, double* bare, int* base, long* perm,
                    long* keys, int* box6);

/* Per quadrature node (pts (m, 3), owner atom, base weight): partition P = (1 - far) * (Becke
 * share of the owner plus the lune of the other cells), weights = base_w * P, f^2, grad u (3, m),
 * W = -|grad u|^2 / 2 and P itself. spheres (host): the sphere table of the quadrature (centres,
 * radii, taper width). Replaces CuspQuadrature._partition_weights / _becke_cells / _bump
 * (src/cdft/operators/quadrature.py:360-419). */
int cdft_node_fields_f64(const double* pts, const int* owner, const double* base_w, long m,
                         int n_atoms, const double* charges, const double* positions,
                         const double* spheres, double cell_width, double* weights, double* f2,
                         double* grad_u, double* wpot, double* partition, void* stream);

/* far[p] = prod_a (1 - taper_a(|r_p - c_a|)) at every grid point. Replaces
 * CuspQuadrature._far_factor (src/cdft/operators/quadrature.py:499-504). */
int cdft_far_factor_f64(const int* shape, double h, const double* origin, int n_spheres,
                        const double* spheres, double* far, void* stream);

/* Degree-7 Lagrange stencil of each point: base (m, 3) = the lower-left grid index of the 8^3
 * stencil's base cell, lw (m, 24) = the 8 weights per axis; bad[0] = 1 if a stencil leaves the
 * grid. Replaces CuspQuadrature._build_stencils / _point_stencil and lagrange_weights
 * (src/cdft/operators/quadrature.py:56-98,506-609,662-700). */
int cdft_node_stencils_f64(const double* pts, long m, const int* shape, double h,
                           const double* origin, int* base, double* lw, int* bad, void* stream);

/* out[f][i] = sum over the 8^3 stencil of node i of lw products times field[f] (nf <= 4 fields of
 * N values each). Replaces CuspQuadrature.interpolate / interpolate_at
 * (src/cdft/operators/quadrature.py:613-620,652-660). */
int cdft_interpolate_f64(const double* field, int nf, const int* base, const double* lw, long m,
                         const int* shape, double h, const double* origin, double* out,
                         void* stream);

/* cdft_interpolate_f64 at every node of a quadrature whose nodes are numbered in base-cell order,
 * bitwise its result: chunk_node (device, int32, n_chunks + 1 node offsets, the last = m) splits the
 * nodes into chunks of at most 16 consecutive nodes that share one base cell; each chunk's 8^3 stencil
 * block is loaded once. Replaces CuspQuadrature.interpolate (src/cdft/operators/quadrature.py:613-620). */
int cdft_interpolate_chunks_f64(const double* field, int nf, const int* chunk_node, int n_chunks,
                                const int* base, const double* lw, long m, const int* shape, double h,
                                const double* origin, double* out, void* stream);

/* The transpose of interpolate: out[f] = add[f] + L^T x[f] (add may be NULL), deterministic and
 * atomic-free. Nodes must be numbered in base-cell order over the cell box (host: lo0, lo1, lo2,
 * c0, c1, c2). The walk is split into chunks of at most 64 consecutive nodes of one z-column; the
 * chunk tables (device, int32) are chunk_node (n_chunks + 1 node offsets), chunk_zlo / chunk_hz
 * (each chunk's first z-cell relative to lo2 and its slab height), chunk_off (n_chunks + 1 slab
 * offsets in doubles per field; a slab is 64 x height) and col_chunk (c0 c1 + 1 chunk offsets per
 * column); work holds nf * (work_per_field + c0 c1 64 (c2 + 7)) doubles, work_per_field =
 * chunk_off[n_chunks]. Replaces
 * CuspQuadrature.lift and the precomputed L^T CSR product
 * (src/cdft/operators/quadrature.py:582-609,628-638). */
int cdft_lift_f64(const double* x, int nf, long m, const int* base, const double* lw, const int* shape,
                  double h, const double* origin, const int* cell_box, int n_chunks, const int* chunk_node,
                  const int* chunk_zlo, const int* chunk_hz, const int* chunk_off, const int* col_chunk,
                  long work_per_field, const double* add, double* work, long work_len, double* out,
                  void* stream);

/* Lumped far mass weights of overlapping spheres (src/cdft/operators/quadrature.py:421-497), three
 * steps. (1) out (f0, f1, f2) = (h_f^3 far) exp(-2u) at the points (fx[i], fy[j], fz[l]) of the
 * refined lattice (fx, fy, fz: device coordinate arrays; spheres (host): cx, cy, cz, radius,
 * taper centre, taper width per sphere, n_spheres <= 8; charges (host)). */
int cdft_far_fine_f64(int n_spheres, const double* charges, const double* spheres, const int* fine_shape,
                      const double* fx, const double* fy, const double* fz, double volume, double* out,
                      void* stream);
/* (2) The transposed one-dimensional Lagrange interpolation along `axis` of a (dims) array onto
 * n_coarse points: out[.., c, ..] = sum_{f in [range[2c], range[2c+1])} w[8 f + c - first[f]]
 * in[.., f, ..] (range, first, w: device tables of the axis). */
int cdft_interp_transpose_axis_f64(const double* in, const int* dims, int axis, int n_coarse,
                                   const int* range, const int* first, const double* w, double* out,
                                   void* stream);
/* (3) out = plain + interior (lumped - plain), interior = erfc((margin - d_face) / width) / 2. */
int cdft_far_blend_f64(const int* shape, double h, const double* origin, double margin, double width,
                       const double* plain, const double* lumped, double* out, void* stream);

/* Host only: the MT19937 state (key[624 i ..], index[i] = next word, 624 = twist due) of the
 * stream seeded `seed` (init_genrand, low 32 bits) after positions[i] 32-bit outputs (ascending).
 * torch's CPU generator runs this stream; from these states a float64 normal draw is split into
 * chunks of 16k values drawn on separate host threads, bitwise equal to the serial draw. Serves the
 * CPU draw of seeded_randn (src/cdft/precision.py:246-263). */
int cdft_mt19937_states(uint64_t seed, int n, const int64_t* positions, uint32_t* key, int32_t* index);

/* Central order-`order` first derivatives of nf fields (coef: DEVICE array of order+1 weights
 * divided by h), out (nf, 3, N), zero halo. Replaces gradient_from_padded
 * (src/cdft/operators/laplacian.py:133-150) as used by the energy terms. */
int cdft_gradient_f64(const double* field, int nf, const int* shape, double h, const double* origin,
                      const double* coef, int order, double* out, void* stream);

/* Per nucleus a: the unit cusp density exp(-2 Z_a s_a) and its closed-form potential at the grid
 * points (pts NULL, m = N) or at pts (m, 3); dens, v1s (n_atoms, m), either may be NULL. Replaces
 * hydrogenic_1s_potential and StepGeometry.build (src/cdft/operators/quadrature.py:805-820,
 * src/cdft/scf/step.py:128-165). */
int cdft_hydrogenic_fields_f64(const double* pts, long m, const int* shape, double h,
                               const double* origin, int n_atoms, const double* charges,
                               const double* positions, double* dens, double* v1s, void* stream);

/* ---- K6 / K8: XC, densities, pointwise algebra and reductions of the Kohn-Sham map ---- */

/* Spin-unpolarised Slater exchange + VWN5 correlation: e = n eps_xc(n) and v = d e / d n
 * (analytic), zero below n = 1e-14. Replaces LDA(VWN5) through SemiLocalFunctional.evaluate
 * (src/cdft/xc/lda.py:122-159,201-217; src/cdft/xc/base.py:115-210). */
int cdft_lda_vwn_f64(const double* n, long len, double* e, double* v, void* stream);

/* Spin-unpolarised PBE: e = e_x + e_c per volume at (n, sigma = |grad n|^2) with v_n = de/dn and
 * v_sigma = de/dsigma (analytic), zero below n = 1e-14, sigma / 4 floored at 1e-40 as per spin.
 * Replaces GGA(PBEExchange, PBECorrelation) through SemiLocalFunctional.evaluate
 * (src/cdft/xc/gga.py:43-90; src/cdft/xc/lda.py:64-108 for the modified PW92; xc/base.py:115-210). */
int cdft_pbe_f64(const double* n, const double* sigma, long len, double* e, double* vn, double* vs, void* stream);

/* The pointwise sigma path of the GGA potentials at len points (the nodes, or the grid points):
 * n = f2 rho, grad n = f2 (grad_rho - 2 rho grad_u), sigma = |grad n|^2, PBE at (n, sigma), then
 * e, aw = w_s f2 v_n, crossw = w_s (-2 flux . grad_u) and fluxw = w_f flux with flux = 2 v_sigma f2 grad n;
 * grad_rho, grad_u and fluxw are (3, len). Replaces the flux and cross terms of
 * KohnShamStep.potentials (src/cdft/scf/step.py:253-279) with QuadratureOperator.density_gradient_at_nodes /
 * density_gradient_on_grid (src/cdft/operators/quadrature.py:757-773). */
int cdft_gga_points_f64(const double* rho, const double* grad_rho, const double* f2, const double* grad_u,
                        const double* w_s, const double* w_f, long len, double* e, double* aw, double* crossw,
                        double* fluxw, void* stream);

/* out = d0 F0 + d1 F1 + d2 F2 of a (3, N) grid field (central order-`order` taps / h on the device, zero
 * halo, axes added in order). The divergence of lift_gradient and lumped_gradient_weights
 * (src/cdft/operators/quadrature.py:640-650,718-729; grid.partial_derivative, grid.py:748-758). */
int cdft_divergence_f64(const double* field, const int* shape, const double* coef, int order, double* out,
                        void* stream);

/* out = (a + (-near_div - h3 far_div)) + cross: the GGA XC lumped weights from their scalar, gradient
 * and cross parts, in the reference's order of addition (src/cdft/scf/step.py:264-271). */
int cdft_gga_combine_f64(const double* a, const double* near_div, const double* far_div, const double* cross,
                         double h3, long len, double* out, void* stream);

/* out = op(a, b, c, s) elementwise over len values: 0 a b; 1 (a b) c; 2 s / a; 3 a / b; 4 a s;
 * 5 (a + b) / c; 6 a - b; 7 (a / s) s; 8 s; 9 a + b; 10 a (b . c) with b, c of shape (3, len);
 * 11 (a b_r) c for the two rows r of b and out (shape (2, len)).
 * The pointwise steps of KohnShamStep.potentials and energy_terms
 * (src/cdft/scf/step.py:224-313,398-439). */
int cdft_ewise_f64(int op, const double* a, const double* b, const double* c, double s, double* out,
                   long len, void* stream);

/* rho = sum_i occ_i phi_i^2 and n = f^2 rho (occ: host, k <= 4). Replaces
 * KohnShamStep.output_density and CuspFactor.density (src/cdft/scf/step.py:366-370,
 * src/cdft/operators/cusp.py:325-331). */
int cdft_density_f64(const double* phi, int k, long len, const double* f2, const double* occ,
                     double* rho, double* n, void* stream);

/* out = base + sign * sum_a c[a] fields[a] (c: device, na values; fields (na, len)). The cusp
 * split of the Hartree term (src/cdft/scf/step.py:239-250). */
int cdft_atom_combine_f64(const double* base, const double* c, const double* fields, int na,
                          long len, double sign, double* out, void* stream);

/* x = A^-1 (b * scale) for a small dense system on the device (na <= 8, partial pivoting). The
 * cusp coefficients c = S^-1 n(R_a) (src/cdft/scf/step.py:239-250). */
/* Pulay/DIIS coefficients on the device: gram (ldg x ldg, the residual history's Gram matrix in the
 * mixing inner product), slots (HOST, m <= 16 active history slots, oldest first); solves the
 * Tikhonov-regularised (1e-12 max diag) bordered system and writes coef[2i] = c_i,
 * coef[2i+1] = alpha c_i. Replaces the host solve of PeriodicPulayMixer.mix
 * (src/cdft/scf/mixing.py:86-162). */
int cdft_pulay_coefficients_f64(const double* gram, int ldg, const int* slots, int m, double alpha, double* coef,
                                void* stream);
int cdft_small_solve_f64(const double* a, const double* scale, const double* b, int na, double* x,
                         void* stream);

/* Doubles of block-partials workspace cdft_reduce_f64 needs for len values. */
long cdft_reduce_workspace_doubles(long len);
/* out[slot] = scale * R (or += with accumulate), R by mode: 0 sum a b w (b, w may be NULL);
 * 1 sum w |a|; 2 max |a - b|. Deterministic two-pass reduction. The integrals of
 * CuspQuadrature.integrate, the energy terms and the SCF residual norms
 * (src/cdft/operators/quadrature.py:701-709; src/cdft/scf/step.py:398-439;
 * src/cdft/scf/loop.py:345-412). */
int cdft_reduce_f64(int mode, const double* a, const double* b, const double* w, long len,
                    double scale, int accumulate, double* partials, long partials_len, double* out,
                    int slot, void* stream);
/* Up to four mode-0 reductions over the same len values in one pass: out[slots[i]] = scale[i] *
 * sum a_i b_i w_i (+= with accumulate[i]); a, b, w, scale, accumulate, slots are host arrays of nd
 * entries (b, w entries may be NULL); partials hold nd x cdft_reduce_workspace_doubles(len) doubles.
 * Bitwise the nd separate cdft_reduce_f64 calls. The same energy and potential integrals
 * (src/cdft/scf/step.py:224-313,398-439). */
int cdft_reduce_dots_f64(int nd, const double* const* a, const double* const* b, const double* const* w,
                         long len, const double* scale, const int* accumulate, const int* slots,
                         double* partials, long partials_len, double* out, void* stream);

/* ---- K5: free-space Hartree potential by the Coulomb-cutoff FFT (rocFFT through hipFFT) ---- */

/* Plans for an (n0, n1, n2) grid zero-padded to (p0, p1, p2); *handle owns them. Replaces
 * CoulombCutoffPoisson.__init__ (src/cdft/operators/poisson.py:27-42,110-120). */
int cdft_poisson_create(int n0, int n1, int n2, int p0, int p1, int p2, void** handle);
int cdft_poisson_destroy(void* handle);
/* kernel (p0 x p1 x (p2/2+1)) = K(G) / (p0 p1 p2), K = 4 pi (1 - cos G R) / G^2, K(0) = 2 pi R^2.
 * Replaces CoulombCutoffPoisson._build_kernel (src/cdft/operators/poisson.py:43-64). */
int cdft_poisson_kernel_f64(void* handle, double h, double radius, double* kernel, void* stream);
/* potential = the cut-off Coulomb potential of density, with the transforms pruned to the box: (y, z)
 * transforms on the n0 occupied x-slabs only, the x transform on the whole padded spectrum. Caller
 * workspace: padded_in and padded_out (n0 p1 p2 doubles; padded_in zero outside the box), slabs and
 * spectrum (p0 p1 (p2/2+1) complex doubles; slabs zero beyond its first n0 x-slabs); the zero parts
 * are never written. The forward x transform, the multiply by kernel and the inverse x transform run
 * as one hand-written column kernel (Stockham radix-2/3/4/5 FFTs in LDS) when p0 factors into 2, 3
 * and 5 and a column fits its LDS budget, else as rocFFT strided transforms around a multiply.
 * Replaces CoulombCutoffPoisson.solve (src/cdft/operators/poisson.py:66-84). */
int cdft_poisson_solve_f64(void* handle, const double* density, double* padded_in, void* slabs, void* spectrum,
                           double* padded_out, const double* kernel, double* potential, void* stream);
/* Which x path cdft_poisson_solve_f64 takes, for every later call in the process: 0 = the column
 * kernel where the plan has one (default), 1 = the rocFFT transforms (same results to rounding).
 * No counterpart in the reference: a switch for comparing the two paths. */
int cdft_poisson_set_x_path(int which);

/* ---- the LOBPCG cross-check's kinetic preconditioner (rocFFT through hipFFT) ---- */

/* 3-D real-to-complex and complex-to-real plans of an (n0, n1, n2) box; *handle owns them. Replaces
 * the lazily built spectrum of LOBPCG._apply_preconditioner (src/cdft/eigen/lobpcg.py:207-228). */
int cdft_tpa_create(int n0, int n1, int n2, void** handle);
int cdft_tpa_destroy(void* handle);

/* The Teter-Payne-Allan filter K(x) = P / (P + 16 x^4), P = 27 + 18x + 12x^2 + 8x^3, x = (G^2/2) / E_ref,
 * on the n0 x n1 x (n2/2+1) half spectrum, divided by n0 n1 n2 (so the inverse transform needs no
 * normalisation); E_ref = max(mean of G^2/2 over the half spectrum, 1e-6), written to *e_ref when it
 * is not NULL (src/cdft/eigen/lobpcg.py:219-236). */
int cdft_tpa_filter_f64(void* handle, double h, double* filter, double* e_ref, void* stream);

/* out[i] = f^-1 K (f in[i]) for the k rows of `in`, f = max(exp(-u), 1e-150) (the similarity frame,
 * src/cdft/eigen/lobpcg.py:129-133,181-192); u = NULL applies K alone (the "transformed" frame).
 * real_work: n0 n1 n2 doubles; spectrum: n0 n1 (n2/2+1) complex doubles; out may alias in. Replaces
 * LOBPCG._apply_preconditioner (src/cdft/eigen/lobpcg.py:207-239). */
int cdft_tpa_apply_f64(void* handle, const double* in, int k, const double* u, const double* filter,
                       double* real_work, void* spectrum, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CDFT_HIP_H */
