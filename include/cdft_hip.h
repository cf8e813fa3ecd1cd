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

#ifdef __cplusplus
}
#endif

#endif /* CDFT_HIP_H */
