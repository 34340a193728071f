/* sparsediff_linalg drop-in: Kronecker-structured products.  Same signatures
 * as the reference's include/sparsediff_linalg/matmul_kron.h:31-55. */
#ifndef SDL_AMD_MATMUL_KRON_H
#define SDL_AMD_MATMUL_KRON_H

#include "sparsediff_linalg/matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* y = kron(I_p, A) x = [A x_1; ...; A x_p]; A sparse or full PD. */
void BA_kron_vec_matrices(const matrix *A, const double *x, double *y, int p);

/* C = kron(A, I_m) J; A a full PD; C is stacked_pd. */
matrix *BA_dense_kron_I_matrices_alloc(const matrix *A, int m, matrix *J);
void BA_dense_kron_I_matrices_fill_values(const matrix *A, int m, const matrix *J,
                                          matrix *C);

/* C = kron(I_p, A) J; A full PD or sparse; J->m == p * A->n. */
matrix *BA_kron_matrices_alloc(const matrix *A, int p, matrix *J);
void BA_kron_matrices_fill_values(const matrix *A, int p, matrix *J, matrix *C);

#ifdef __cplusplus
}
#endif

#endif
