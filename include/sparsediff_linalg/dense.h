/* sparsediff_linalg drop-in: dense helpers.  Same signatures as the reference's
 * include/sparsediff_linalg/dense.h:22-41 (src/support/mini_numpy.c:22-105).
 * Arrays are host arrays; each call runs on the GPU and returns when the
 * result is in `result` / `AT` / `v`.  The Kronecker products sum every output
 * from +0.0 over individually rounded products in the reference's order. */
#ifndef SDL_AMD_DENSE_H
#define SDL_AMD_DENSE_H

#ifdef __cplusplus
extern "C" {
#endif

/* reference dense.h:22: result = [a[0] x repeats, a[1] x repeats, ...]. */
void repeat(double *result, const double *a, int len, int repeats);
/* reference dense.h:25: result = a tiled `tiles` times. */
void tile_double(double *result, const double *a, int len, int tiles);
/* reference dense.h:28: result[0..size) = value. */
void scaled_ones(double *result, int size, double value);
/* reference dense.h:31: AT[j*m + i] = A[i*n + j] for a row-major m x n A. */
void A_transpose(double *AT, const double *A, int m, int n);
/* reference dense.h:36: v = (Y kron I_m) w, Y k x n col-major, len(w) = m*n. */
void Y_kron_I_vec(int m, int k, int n, const double *Y, const double *w, double *v);
/* reference dense.h:41: v = (I_n kron X^T) w, X m x k col-major, len(w) = m*n. */
void I_kron_XT_vec(int m, int k, int n, const double *X, const double *w, double *v);

#ifdef __cplusplus
}
#endif

#endif
