// permuted_dense / stacked_pd kinds (in progress: sparse-only stubs).
#include "api_pd.h"

namespace sdl {

void pd_kind_free(matrix *m) { unsupported("free_matrix", m, nullptr); }
matrix *pd_copy_sparsity(const matrix *A) { unsupported("copy_sparsity_matrices_alloc", A, nullptr); }
void pd_DA_fill(const double *, const matrix *A, matrix *C) { unsupported("DA_matrices_fill_values", A, C); }
matrix *pd_transpose_alloc(const matrix *A) { unsupported("transpose_matrices_alloc", A, nullptr); }
void pd_transpose_fill(const matrix *A, matrix *AT) { unsupported("transpose_matrices_fill_values", A, AT); }
matrix *pd_BTA_alloc(matrix *A, matrix *B) { unsupported("BTA_matrices_alloc", A, B); }
void pd_BTDA_fill(matrix *A, const double *, matrix *B, matrix *) { unsupported("BTDA_matrices_fill_values", A, B); }
matrix *pd_ATA_alloc(matrix *A) { unsupported("ATA_matrices_alloc", A, nullptr); }
void pd_ATDA_fill(const matrix *A, const double *, matrix *C) { unsupported("ATDA_matrices_fill_values", A, C); }
matrix *pd_BA_alloc(const matrix *B, matrix *A) { unsupported("BA_matrices_alloc", B, A); }
void pd_BA_fill(const matrix *B, const matrix *A, matrix *) { unsupported("BA_matrices_fill_values", B, A); }
void pd_kron_vec(const matrix *A, const double *, double *, int) { unsupported("BA_kron_vec_matrices", A, nullptr); }
matrix *pd_dense_kron_alloc(const matrix *A, int, matrix *J, bool) { unsupported("BA_dense_kron", A, J); }
void pd_dense_kron_fill(const matrix *A, int, matrix *J, matrix *, bool) { unsupported("BA_dense_kron", A, J); }
matrix *pd_csr_operand(matrix *M)
{
    if (M->kind == MATRIX_SPARSE) return M;
    unsupported("csr view", M, nullptr);
}
matrix *pd_csr_operand_fresh(matrix *M) { return pd_csr_operand(M); }

} // namespace sdl

extern "C" {
#pragma GCC visibility push(default)
matrix *new_permuted_dense(int, int, int, int, const int *, const int *, const double *)
{
    sdl::api_exit("new_permuted_dense", "not built yet");
}
matrix *new_permuted_dense_full(int, int, const double *) { sdl::api_exit("new_permuted_dense_full", "not built yet"); }
matrix *new_stacked_pd(int, int, int, permuted_dense **, const int *, const int *)
{
    sdl::api_exit("new_stacked_pd", "not built yet");
}
#pragma GCC visibility pop
}
