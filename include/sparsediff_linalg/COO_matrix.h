/* sparsediff_linalg drop-in: coordinate-format matrices.  Same layout and
 * signatures as the reference's include/sparsediff_linalg/COO_matrix.h:35-57,
 * plus one new alloc/fill pair that coalesces triplets into CSR. */
#ifndef SDL_AMD_COO_MATRIX_H
#define SDL_AMD_COO_MATRIX_H

#include "sparsediff_linalg/CSR_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct COO_matrix
{
    int *rows;
    int *cols;
    double *x;
    int *value_map; /* lower-triangular export: CSR position of each entry */
    int m;
    int n;
    int nnz;
} COO_matrix;

/* reference COO_matrix.h:47-57 (COO_matrix.c:23-110) */
COO_matrix *new_COO_matrix(const CSR_matrix *A);
COO_matrix *new_COO_matrix_lower_triangular(const CSR_matrix *A);
void refresh_lower_triangular_coo(COO_matrix *coo, const double *vals);
void free_COO_matrix(COO_matrix *matrix);

/* New (no reference entry point; semantics of the reference's row_reduce,
   src/sparse/sparse_matrix.c:136-240, with every triplet a one-entry row grouped
   by its row index): C is the m x n CSR with columns ascending and unique per
   row; the fill sets each entry to ((+0.0 + x_t1) + x_t2) + ... over its
   duplicate triplets in input order.  The fill reads A->x only and must get a
   COO with the same rows/cols as the alloc. */
CSR_matrix *COO_coalesce_to_csr_alloc(const COO_matrix *A);
void COO_coalesce_to_csr_fill_values(const COO_matrix *A, CSR_matrix *C);

#ifdef __cplusplus
}
#endif

#endif
