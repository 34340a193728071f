"""ctypes signatures of the reference-compatible drop-in entry points
(include/sparsediff_linalg/*.h)."""
import ctypes as C

from ._lib import CSR_p, CSC_p, COO_p, MAT_p, c_int_p, c_double_p

ip, dp = c_int_p, c_double_p
PD_pp = C.POINTER(MAT_p)  # permuted_dense ** (opaque handles)

SIGNATURES = {
    # CSR_matrix.h
    "new_CSR_matrix": (CSR_p, [C.c_int, C.c_int, C.c_int]),
    "new_csr": (CSR_p, [CSR_p]),
    "free_CSR_matrix": (None, [CSR_p]),
    "copy_CSR_matrix": (None, [CSR_p, CSR_p]),
    "CSR_trim": (None, [CSR_p]),
    "transpose": (CSR_p, [CSR_p, ip]),
    "AT_alloc": (CSR_p, [CSR_p, ip]),
    "AT_fill_values": (None, [CSR_p, CSR_p, ip]),
    # CSC_matrix.h
    "new_CSC_matrix": (CSC_p, [C.c_int, C.c_int, C.c_int]),
    "free_CSC_matrix": (None, [CSC_p]),
    "BTA_alloc": (CSR_p, [CSR_p, CSC_p]),
    "BTDA_fill_values": (None, [CSR_p, CSC_p, dp, dp, CSR_p]),
    "ATA_alloc": (CSR_p, [CSC_p, CSR_p]),
    "ATDA_fill_values": (None, [CSC_p, dp, CSR_p, dp, CSR_p]),
    "csr_to_csc_alloc": (CSC_p, [CSR_p, ip]),
    "csr_to_csc_fill_values": (None, [CSR_p, CSC_p, ip]),
    "csc_to_csr_alloc": (CSR_p, [CSC_p, ip]),
    "csc_to_csr_fill_values": (None, [CSC_p, CSR_p, ip]),
    # COO_matrix.h
    "new_COO_matrix": (COO_p, [CSR_p]),
    "new_COO_matrix_lower_triangular": (COO_p, [CSR_p]),
    "refresh_lower_triangular_coo": (None, [COO_p, dp]),
    "free_COO_matrix": (None, [COO_p]),
    "COO_coalesce_to_csr_alloc": (CSR_p, [COO_p]),
    "COO_coalesce_to_csr_fill_values": (None, [COO_p, CSR_p]),
    # matrix.h
    "free_matrix": (None, [MAT_p]),
    "copy_sparsity_matrices_alloc": (MAT_p, [MAT_p]),
    "transpose_matrices_alloc": (MAT_p, [MAT_p]),
    "transpose_matrices_fill_values": (None, [MAT_p, MAT_p]),
    "new_sparse_matrix": (MAT_p, [CSR_p]),
    "new_sparse_matrix_alloc": (MAT_p, [C.c_int, C.c_int, C.c_int]),
    "new_permuted_dense": (MAT_p, [C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp]),
    "new_permuted_dense_full": (MAT_p, [C.c_int, C.c_int, dp]),
    "new_stacked_pd": (MAT_p, [C.c_int, C.c_int, C.c_int, PD_pp, ip, ip]),
    # matmul.h
    "BTA_matrices_alloc": (MAT_p, [MAT_p, MAT_p]),
    "BTA_matrices_fill_values": (None, [MAT_p, MAT_p, MAT_p]),
    "BTDA_matrices_fill_values": (None, [MAT_p, dp, MAT_p, MAT_p]),
    "DA_matrices_fill_values": (None, [dp, MAT_p, MAT_p]),
    "ATA_matrices_alloc": (MAT_p, [MAT_p]),
    "ATDA_matrices_fill_values": (None, [MAT_p, dp, MAT_p]),
    "BA_matrices_alloc": (MAT_p, [MAT_p, MAT_p]),
    "BA_matrices_fill_values": (None, [MAT_p, MAT_p, MAT_p]),
    # matmul_kron.h
    "BA_kron_vec_matrices": (None, [MAT_p, dp, dp, C.c_int]),
    "BA_dense_kron_I_matrices_alloc": (MAT_p, [MAT_p, C.c_int, MAT_p]),
    "BA_dense_kron_I_matrices_fill_values": (None, [MAT_p, C.c_int, MAT_p, MAT_p]),
    "BA_kron_matrices_alloc": (MAT_p, [MAT_p, C.c_int, MAT_p]),
    "BA_kron_matrices_fill_values": (None, [MAT_p, C.c_int, MAT_p, MAT_p]),
    # dense.h
    "repeat": (None, [dp, dp, C.c_int, C.c_int]),
    "tile_double": (None, [dp, dp, C.c_int, C.c_int]),
    "scaled_ones": (None, [dp, C.c_int, C.c_double]),
    "A_transpose": (None, [dp, dp, C.c_int, C.c_int]),
    "Y_kron_I_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "I_kron_XT_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    # matrix_sum.h
    "sum_matrices_alloc": (None, [MAT_p, MAT_p, MAT_p]),
    "sum_matrices_fill_values": (None, [MAT_p, MAT_p, MAT_p]),
    "sum_scaled_matrices_fill_values": (None, [MAT_p, MAT_p, MAT_p, dp, dp]),
    "sum_4_matrices_alloc": (MAT_p, [MAT_p, MAT_p, MAT_p, MAT_p]),
    "sum_4_matrices_fill_values": (None, [MAT_p, MAT_p, MAT_p, MAT_p, MAT_p]),
    "sum_n_matrices_alloc": (MAT_p, [C.POINTER(MAT_p), C.c_int]),
    "sum_n_matrices_fill_values": (None, [C.POINTER(MAT_p), C.c_int, MAT_p]),
    "sum_n_matrices_fill_zero": (None, [MAT_p]),
    "sum_n_matrices_fill_add": (None, [MAT_p, C.c_int, MAT_p]),
    # matrix_vstack.h
    "vstack_matrices_alloc": (MAT_p, [C.POINTER(MAT_p), C.c_int]),
    "vstack_matrices_fill_values": (None, [C.POINTER(MAT_p), C.c_int, MAT_p]),
}
