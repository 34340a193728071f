// Dispatcher hooks of the permuted_dense / stacked_pd kinds (api_pd.hip).
// The sparse-kind layer (api_matrix.hip) routes every call whose operands are
// not both sparse through these.
#pragma once

#include "api_matrix.h"

namespace sdl {

void pd_kind_free(matrix *m);
matrix *pd_copy_sparsity(const matrix *A);
void pd_DA_fill(const double *d, const matrix *A, matrix *C);
matrix *pd_transpose_alloc(const matrix *A);
void pd_transpose_fill(const matrix *A, matrix *AT);
matrix *pd_BTA_alloc(matrix *A, matrix *B);
void pd_BTDA_fill(matrix *A, const double *d, matrix *B, matrix *C);
matrix *pd_ATA_alloc(matrix *A);
void pd_ATDA_fill(const matrix *A, const double *d, matrix *C);
matrix *pd_BA_alloc(const matrix *B, matrix *A);
void pd_BA_fill(const matrix *B, const matrix *A, matrix *C);
void pd_kron_vec(const matrix *A, const double *x, double *y, int p);
matrix *pd_dense_kron_alloc(const matrix *A, int p, matrix *J, bool kron_A_I);
void pd_dense_kron_fill(const matrix *A, int p, matrix *J, matrix *C, bool kron_A_I);

// A sparse_matrix whose CSR is the CSR view of M (M itself when sparse).
// _fresh additionally brings the view's values up to date (fill phase).
matrix *pd_csr_operand(matrix *M);
matrix *pd_csr_operand_fresh(matrix *M);

} // namespace sdl
