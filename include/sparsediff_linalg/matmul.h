/* sparsediff_linalg drop-in: product dispatchers.  Same signatures and output
 * kind table as the reference's include/sparsediff_linalg/matmul.h:26-69:
 *   B PD -> PD;  B stacked_pd -> stacked_pd;  B sparse -> A's kind.
 * Note the argument order: BTA_matrices_alloc(A, B) builds B^T A.  Every fill
 * bumps C's values_version once and allocates nothing. */
#ifndef SDL_AMD_MATMUL_H
#define SDL_AMD_MATMUL_H

#include "sparsediff_linalg/matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference matmul.h:38-40 (matmul_dispatchers.c:48-112) */
matrix *BTA_matrices_alloc(matrix *A, matrix *B);
void BTA_matrices_fill_values(matrix *A, matrix *B, matrix *C);
void BTDA_matrices_fill_values(matrix *A, const double *d, matrix *B, matrix *C);

/* reference matmul.h:44 (matmul_dispatchers.c:431-449): C = diag(d) A. */
void DA_matrices_fill_values(const double *d, const matrix *A, matrix *C);

/* reference matmul.h:52-53 (matmul_dispatchers.c:490-527): C = A^T diag(d) A,
   the full product (both triangles). */
matrix *ATA_matrices_alloc(matrix *A);
void ATDA_matrices_fill_values(const matrix *A, const double *d, matrix *C);

/* reference matmul.h:67-68 (matmul_dispatchers.c:197-223): C = B A; a
   stacked_pd B is rejected. */
matrix *BA_matrices_alloc(const matrix *B, matrix *A);
void BA_matrices_fill_values(const matrix *B, const matrix *A, matrix *C);

#ifdef __cplusplus
}
#endif

#endif
