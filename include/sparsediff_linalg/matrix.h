/* sparsediff_linalg drop-in: the polymorphic matrix handle.
 *
 * The struct layout (field order, the vtable slot order) and the constructor
 * signatures are the reference's include/sparsediff_linalg/matrix.h:38-168,
 * because callers read m/n/nnz/x/kind/values_version and call the vtable slots
 * directly.  Each handle keeps a device mirror of its values, uploaded iff
 * values_version moved since the last upload; every fill runs on the GPU and
 * writes its result back to the host x it exposes. */
#ifndef SDL_AMD_MATRIX_H
#define SDL_AMD_MATRIX_H

#include "sparsediff_linalg/CSC_matrix.h"
#include "sparsediff_linalg/CSR_matrix.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct matrix matrix;

/* reference matrix.h:38-43; MATRIX_SPARSE is 0 so a zeroed handle is sparse. */
typedef enum
{
    MATRIX_SPARSE = 0,
    MATRIX_PERMUTED_DENSE,
    MATRIX_STACKED_PD
} matrix_kind;

typedef CSR_matrix *(*matrix_to_csr_fn)(matrix *A);
/* C = A[map, :]; C keeps map, so the fill takes none. */
typedef matrix *(*matrix_row_gather_alloc_fn)(const matrix *A, const int *map, int m_out);
typedef void (*matrix_row_gather_fill_values_fn)(const matrix *A, matrix *C);
/* C[g, :] = sum of rows i with group[i] == g; C keeps the reduction map. */
typedef matrix *(*matrix_row_reduce_alloc_fn)(const matrix *A, const int *group, int m_out);
typedef void (*matrix_row_reduce_fill_values_fn)(const matrix *A, matrix *C);

/* reference matrix.h:64-88 */
struct matrix
{
    int m, n, nnz;
    double *x;                /* non-owning pointer to the host value buffer */
    matrix_kind kind;
    uint64_t values_version;  /* bump after writing x directly (on the owner) */
    matrix_to_csr_fn to_csr;
    matrix_row_gather_alloc_fn row_gather_alloc;
    matrix_row_gather_fill_values_fn row_gather_fill_values;
    matrix_row_reduce_alloc_fn row_reduce_alloc;
    matrix_row_reduce_fill_values_fn row_reduce_fill_values;
};

/* reference matrix.h:91-94 */
static inline void matrix_values_changed(matrix *A)
{
    A->values_version++;
}

/* reference matrix.h:97-110 */
void free_matrix(matrix *m);
matrix *copy_sparsity_matrices_alloc(const matrix *A);
matrix *transpose_matrices_alloc(const matrix *A);
void transpose_matrices_fill_values(const matrix *A, matrix *AT);

/* reference matrix.h:117-123: takes ownership of A. */
matrix *new_sparse_matrix(CSR_matrix *A);
matrix *new_sparse_matrix_alloc(int m, int n, int nnz);

/* reference matrix.h:125-147: a dense m0 x n0 block X (row-major) at strictly
   increasing row_perm / col_perm of an m x n matrix. */
typedef struct permuted_dense permuted_dense;
matrix *new_permuted_dense(int m, int n, int m0, int n0, const int *row_perm,
                           const int *col_perm, const double *X_data);
matrix *new_permuted_dense_full(int m, int n, const double *data);

/* reference matrix.h:149-166: a row-disjoint vertical stack of permuted_dense
   blocks; takes ownership of the blocks and absorbs their values into one
   block-major buffer. */
matrix *new_stacked_pd(int m, int n, int n_blocks, permuted_dense **blocks,
                       const int *src_block_idx_p, const int *src_block_idx);

#ifdef __cplusplus
}
#endif

#endif
