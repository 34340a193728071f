/* sparsediff_linalg drop-in: compressed sparse row matrices.
 *
 * Same struct layout and signatures as the reference's
 * include/sparsediff_linalg/CSR_matrix.h:31-55, so callers compiled against the
 * reference link against libsdl_amd.so unchanged.  Arrays are host memory (the
 * ABI truth); every computation runs on the GPU. */
#ifndef SDL_AMD_CSR_MATRIX_H
#define SDL_AMD_CSR_MATRIX_H

#ifdef __cplusplus
extern "C" {
#endif

/* m x n matrix: p[m+1] row starts, i[nnz] column indices (ascending, unique
   per row), x[nnz] values.  Field order is part of the ABI. */
typedef struct CSR_matrix
{
    int *p;
    int *i;
    double *x;
    int m;
    int n;
    int nnz;
} CSR_matrix;

/* reference CSR_matrix.h:42-45 (CSR_matrix.c:27-65) */
CSR_matrix *new_CSR_matrix(int m, int n, int nnz);
CSR_matrix *new_csr(const CSR_matrix *A);
void free_CSR_matrix(CSR_matrix *matrix);
void copy_CSR_matrix(const CSR_matrix *A, CSR_matrix *C);

/* reference CSR_matrix.h:50 (CSR_matrix.c:56-63): shrink i/x to p[m]. */
void CSR_trim(CSR_matrix *A);

/* reference CSR_matrix.h:53-55 (CSR_matrix.c:99-205): A^T as CSR.  iwork (length
   A->n) is accepted for signature compatibility; the GPU path does not use it. */
CSR_matrix *transpose(const CSR_matrix *A, int *iwork);
CSR_matrix *AT_alloc(const CSR_matrix *A, int *iwork);
void AT_fill_values(const CSR_matrix *A, CSR_matrix *AT, int *iwork);

#ifdef __cplusplus
}
#endif

#endif
