/* sparsediff_linalg drop-in: compressed sparse column matrices and the raw
 * B^T D A entry points.  Same layout/signatures as the reference's
 * include/sparsediff_linalg/CSC_matrix.h:33-61. */
#ifndef SDL_AMD_CSC_MATRIX_H
#define SDL_AMD_CSC_MATRIX_H

#include "sparsediff_linalg/CSR_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* m x n matrix: p[n+1] column starts, i[nnz] row indices, x[nnz] values. */
typedef struct CSC_matrix
{
    int *p;
    int *i;
    double *x;
    int m;
    int n;
    int nnz;
} CSC_matrix;

/* reference CSC_matrix.h:44-45 (CSC_matrix.c:26-60) */
CSC_matrix *new_CSC_matrix(int m, int n, int nnz);
void free_CSC_matrix(CSC_matrix *matrix);

/* reference CSC_matrix.h:48 (CSC_matrix.c:191-196): structure of C = B^T A,
   computed by the GPU symbolic SpGEMM; rows sorted, exact. */
CSR_matrix *BTA_alloc(const CSR_matrix *A, const CSC_matrix *B);

/* reference CSC_matrix.h:52-53 (CSC_matrix.c:198-203): C.x = B^T diag(d) A, NULL d
   is the identity.  acc is accepted for compatibility and unused.  Values are
   bit-identical to the reference's accumulator fill. */
void BTDA_fill_values(const CSR_matrix *A, const CSC_matrix *B, const double *d,
                      double *acc, CSR_matrix *C);

/* reference src/internal/CSC_matrix.h (CSC_matrix.c:205-217): A^T D A with A
   given both as CSC and CSR. */
CSR_matrix *ATA_alloc(const CSC_matrix *A_csc, const CSR_matrix *A);
void ATDA_fill_values(const CSC_matrix *A_csc, const double *d, const CSR_matrix *A,
                      double *acc, CSR_matrix *C);

/* reference CSC_matrix.h:56-61 (CSC_matrix.c:62-189): stable transposition. */
CSC_matrix *csr_to_csc_alloc(const CSR_matrix *A, int *iwork);
void csr_to_csc_fill_values(const CSR_matrix *A, CSC_matrix *C, int *iwork);
CSR_matrix *csc_to_csr_alloc(const CSC_matrix *A, int *iwork);
void csc_to_csr_fill_values(const CSC_matrix *A, CSR_matrix *C, int *iwork);

#ifdef __cplusplus
}
#endif

#endif
