/* sparsediff_linalg drop-in: sums of same-shape matrices of any kinds into a
 * sparse matrix.  Same signatures as the reference's
 * include/sparsediff_linalg/matrix_sum.h:24-57.  The alloc builds the union
 * structure on the GPU; every fill sums each output entry from +0.0 over its
 * operands in operand order on the GPU (the reference's accumulator order,
 * src/sparse/CSR_sum.c:150-157), writes C's host values and bumps C's
 * values_version once. */
#ifndef SDL_AMD_MATRIX_SUM_H
#define SDL_AMD_MATRIX_SUM_H

#include "sparsediff_linalg/matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference matrix_sum.h:24: structure of C = A + B into the caller's sparse C
   (C's column and value arrays are resized to the union). */
void sum_matrices_alloc(matrix *A, matrix *B, matrix *C);
/* reference matrix_sum.h:27: values of C = A + B. */
void sum_matrices_fill_values(matrix *A, matrix *B, matrix *C);
/* reference matrix_sum.h:30-31: C = diag(d1) A + diag(d2) B, C from
   sum_matrices_alloc; each term is the rounded product d[row] * a. */
void sum_scaled_matrices_fill_values(matrix *A, matrix *B, matrix *C, const double *d1,
                                     const double *d2);
/* reference matrix_sum.h:34,37-38: C = A + B + D + E. */
matrix *sum_4_matrices_alloc(matrix *A, matrix *B, matrix *D, matrix *E);
void sum_4_matrices_fill_values(const matrix *A, const matrix *B, const matrix *D,
                                const matrix *E, matrix *C);
/* reference matrix_sum.h:42,46: C = ops[0] + ... + ops[n-1], n >= 1. */
matrix *sum_n_matrices_alloc(matrix *const *ops, int n);
void sum_n_matrices_fill_values(matrix *const *ops, int n, matrix *C);
/* reference matrix_sum.h:53,57: incremental fill -- zero C, then add operand k
   right after its values are final (stays correct when a later operand's
   evaluation overwrites an earlier one's values). */
void sum_n_matrices_fill_zero(matrix *C);
void sum_n_matrices_fill_add(const matrix *op, int k, matrix *C);

#ifdef __cplusplus
}
#endif

#endif
