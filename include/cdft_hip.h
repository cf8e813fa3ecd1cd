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

/* K1: out = 0.5 * invw * sum_d D_d^T [w_d D_d phi] + pot * phi for a block of k vectors.
 * Replaces CuspFactoredHamiltonian.apply / apply_with_potential / kinetic_apply
 * (src/cdft/operators/hamiltonian.py:251-318) and weighted_divergence
 * (src/cdft/operators/divergence.py:153-197, device tier fused.py:353-380). order in {2,4,6,8}. */
int cdft_hdiv_apply_f64(const double* phi, const double* w0, const double* w1, const double* w2,
                        const double* invw, const double* pot, double* out, int k, int n0, int n1,
                        int n2, const double* coef, int order, void* stream);
int cdft_hdiv_apply_f32(const float* phi, const float* w0, const float* w1, const float* w2,
                        const float* invw, const float* pot, float* out, int k, int n0, int n1,
                        int n2, const double* coef, int order, void* stream);

/* K2: one fused scaled-Chebyshev step, out = (A y - c y) a - b prev with (c, a, b) = table[step]
 * read on the device (table is a (degree, 3) float64 device array), prev ignored at step 0.
 * Replaces chebyshev_update (src/cdft/operators/fused.py:391-417) and the body of
 * chebyshev_recurrence (src/cdft/eigen/chefsi.py:139-160). */
int cdft_cheb_step_f64(const double* y, const double* prev, const double* w0, const double* w1,
                       const double* w2, const double* invw, const double* pot, double* out, int k,
                       int n0, int n1, int n2, const double* coef, int order, const double* table,
                       int step, void* stream);
int cdft_cheb_step_f32(const float* y, const float* prev, const float* w0, const float* w1,
                       const float* w2, const float* invw, const float* pot, float* out, int k,
                       int n0, int n1, int n2, const double* coef, int order, const double* table,
                       int step, void* stream);

/* K2: the whole degree-`degree` filter of x into out, rotating through work1/work2 (each (k, N);
 * x, out, work1, work2 distinct). Replaces chebyshev_filter (src/cdft/eigen/chefsi.py:81-113) and
 * the captured GraphedChebyshevFilter replay (chefsi.py:185-392). */
int cdft_cheb_filter_f64(const double* x, double* out, double* work1, double* work2,
                         const double* w0, const double* w1, const double* w2, const double* invw,
                         const double* pot, int k, int n0, int n1, int n2, const double* coef,
                         int order, const double* table, int degree, void* stream);
int cdft_cheb_filter_f32(const float* x, float* out, float* work1, float* work2, const float* w0,
                         const float* w1, const float* w2, const float* invw, const float* pot,
                         int k, int n0, int n1, int n2, const double* coef, int order,
                         const double* table, int degree, void* stream);

/* ---- K3 / K4: float64 reductions and rotations of Rayleigh-Ritz (deterministic, no atomics) ---- */

/* Doubles of block-partials workspace the reductions below need for length n. */
long cdft_reduction_workspace_doubles(long n);

/* out[i*kb + j] = scale * sum_p q[p] a[i*n + p] b[j*n + p] (q may be NULL for unit weights);
 * post = 1 stores sqrt(max(x, 0)). Replaces Measure.inner/norm/gram/cross
 * (src/cdft/eigen/measure.py:72-96) and orthonormality_error_tensor (eigen/rayleigh_ritz.py:44-48). */
int cdft_wgram_f64(const double* a, int ka, const double* b, int kb, const double* q, long n,
                   double scale, int post, double* partials, long partials_len, double* out,
                   void* stream);

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

#ifdef __cplusplus
}
#endif

#endif /* CDFT_HIP_H */
