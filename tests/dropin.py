"""Test-side helpers driving the drop-in C ABI through ctypes: build CSR /
CSC / COO structs and matrix handles from numpy arrays, read them back."""
import ctypes as C

import numpy as np

from sparsediff_linalg_amd import load
from sparsediff_linalg_amd._lib import MAT_p


def lib():
    return load()


def new_csr(p, i, x, m, n):
    L = lib()
    p = np.ascontiguousarray(p, np.int32)
    i = np.ascontiguousarray(i, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    nnz = int(p[m])
    A = L.new_CSR_matrix(m, n, nnz)
    C.memmove(A.contents.p, p.ctypes.data, (m + 1) * 4)
    if nnz:
        C.memmove(A.contents.i, i.ctypes.data, nnz * 4)
        C.memmove(A.contents.x, x.ctypes.data, nnz * 8)
    return A


def new_csc(p, i, x, m, n):
    L = lib()
    p = np.ascontiguousarray(p, np.int32)
    i = np.ascontiguousarray(i, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    nnz = int(p[n])
    A = L.new_CSC_matrix(m, n, nnz)
    C.memmove(A.contents.p, p.ctypes.data, (n + 1) * 4)
    if nnz:
        C.memmove(A.contents.i, i.ctypes.data, nnz * 4)
        C.memmove(A.contents.x, x.ctypes.data, nnz * 8)
    return A


def read_compressed(S, n_ptr):
    s = S.contents
    p = np.ctypeslib.as_array(s.p, (n_ptr + 1,)).copy()
    nnz = int(p[-1])
    i = np.ctypeslib.as_array(s.i, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    x = np.ctypeslib.as_array(s.x, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0)
    return p, i, x


def read_csr(S):
    return read_compressed(S, S.contents.m)


def read_csc(S):
    return read_compressed(S, S.contents.n)


def sparse(p, i, x, m, n):
    """A sparse_matrix handle (takes ownership of a fresh CSR)."""
    return lib().new_sparse_matrix(new_csr(p, i, x, m, n))


def mat_csr(M):
    """(p, i, x) of a handle's CSR view."""
    csr = M.contents.to_csr(M)
    return read_csr(csr)


def mat_x(M):
    nnz = M.contents.nnz
    if nnz == 0:
        return np.zeros(0)
    return np.ctypeslib.as_array(M.contents.x, (nnz,)).copy()


def set_x(M, x):
    """Write values directly into a handle's host buffer and bump its version
    (the caller contract of matrix_values_changed)."""
    x = np.ascontiguousarray(x, np.float64)
    C.memmove(M.contents.x, x.ctypes.data, x.nbytes)
    M.contents.values_version += 1


def ptr(a, t):
    return np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))


def dptr(a):
    return None if a is None else ptr(np.ascontiguousarray(a, np.float64), C.c_double)


def iptr(a):
    return ptr(np.ascontiguousarray(a, np.int32), C.c_int)


def pd(m, n, row_perm, col_perm, X):
    row_perm = np.ascontiguousarray(row_perm, np.int32)
    col_perm = np.ascontiguousarray(col_perm, np.int32)
    X = np.ascontiguousarray(X, np.float64)
    return lib().new_permuted_dense(m, n, len(row_perm), len(col_perm),
                                    iptr(row_perm) if len(row_perm) else None,
                                    iptr(col_perm) if len(col_perm) else None,
                                    dptr(X) if X.size else None)


def spd(m, n, blocks):
    arr = (MAT_p * max(len(blocks), 1))(*blocks)
    return lib().new_stacked_pd(m, n, len(blocks), C.cast(arr, C.POINTER(MAT_p)), None, None)


def ops_array(ops):
    return (MAT_p * len(ops))(*ops)
