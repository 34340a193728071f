"""ORACLE (test infrastructure only): ctypes binding of liboracle.so, the
single-threaded CPU restatement of the reference's sparse-product algorithms,
and of oracle/_ref/libref_sparse.so, the reference's own sparse sources built
in this repository's container.  Only tests, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use this module, and only as the checker."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref_sparse.so")

ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_longlong)

_SIGS = {
    "orc_spgemm_symbolic": (C.c_longlong, [C.c_int, ip, ip, C.c_int, C.c_int, C.c_int, ip, ip,
                                           lp, ip]),
    "orc_spgemm_numeric": (None, [C.c_int, ip, ip, dp, C.c_int, dp, C.c_int, C.c_int, ip, ip,
                                  dp, lp, ip, dp, dp]),
    "orc_spgemm_plan": (C.c_longlong, [C.c_int, ip, ip, C.c_int, C.c_int, C.c_int, ip, ip, lp,
                                       ip, C.c_longlong, lp]),
    "orc_spgemm_numeric_planned": (None, [C.c_int, ip, ip, dp, C.c_int, dp, C.c_int, ip, dp,
                                          lp, C.c_longlong, dp]),
    "orc_spgemm_row": (C.c_int, [C.c_int, C.c_int, ip, ip, dp, C.c_int, dp, C.c_int, C.c_int,
                                 ip, ip, dp, ip, dp, ip, dp]),
    "orc_csr_transpose": (None, [C.c_int, C.c_int, ip, ip, dp, ip, ip, dp, ip]),
    "orc_spgemm_range": (C.c_longlong, [C.c_int, C.c_int, C.c_int, ip, ip, dp, C.c_int, dp,
                                        C.c_int, C.c_int, ip, ip, dp, ip, dp, C.c_longlong, lp,
                                        ip, dp]),
    "orc_row_reduce_alloc": (C.c_int, [C.c_int, C.c_int, ip, ip, ip, C.c_int, ip, ip, ip]),
    "orc_row_reduce_fill": (None, [C.c_int, dp, ip, C.c_int, dp]),
    "orc_coo_coalesce": (C.c_longlong, [C.c_longlong, ip, ip, dp, C.c_int, C.c_int, lp, ip,
                                        dp]),
    "orc_row_gather": (C.c_longlong, [C.c_int, ip, ip, ip, dp, ip, ip, dp]),
    "orc_da": (None, [C.c_int, ip, dp, dp, dp]),
    "orc_kron_spmv": (None, [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_int]),
    "orc_coo_lower": (C.c_int, [C.c_int, ip, ip, dp, ip, ip, dp, ip]),
    "orc_y_kron_i_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "orc_i_kron_xt_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "orc_sum_n_alloc": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip, ip, ip, ip, ip]),
    "orc_sum_n_fill": (None, [C.c_int, C.c_int, ip, ip, dp, dp, ip, C.c_int, dp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(ORACLE_PATH, mode=os.RTLD_LOCAL)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a, t):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))


def I(a):  # noqa: E743
    return _p(a, C.c_int)


def D(a):
    return _p(a, C.c_double)


def L(a):
    return _p(a, C.c_longlong)


def spgemm(Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax, planned=None):
    """C = kron(I_blocks, L) diag(d) A by the restated Gustavson (gustavson.c).
    Returns (Cp int64, Ci int32, Cx float64, terms).  planned=None follows the
    reference's plan-or-accumulator policy; True/False forces one."""
    o = lib()
    Lp, Li, Ap, Ai = (np.ascontiguousarray(v, np.int32) for v in (Lp, Li, Ap, Ai))
    m_out = Lm * blocks
    Cp = np.zeros(m_out + 1, np.int64)
    nnz = o.orc_spgemm_symbolic(Lm, I(Lp), I(Li), blocks, Am, An, I(Ap), I(Ai), L(Cp), None)
    Ci = np.zeros(max(nnz, 1), np.int32)
    o.orc_spgemm_symbolic(Lm, I(Lp), I(Li), blocks, Am, An, I(Ap), I(Ai), L(Cp), I(Ci))
    Ci = Ci[:nnz]
    Cx = np.zeros(max(nnz, 1), np.float64)
    terms = o.orc_spgemm_plan(Lm, I(Lp), I(Li), blocks, Am, An, I(Ap), I(Ai), L(Cp), I(Ci),
                              -1, None)
    use_plan = planned if planned is not None else (0 < terms <= 3 * nnz)
    Lx = np.ascontiguousarray(Lx, np.float64)
    Ax = np.ascontiguousarray(Ax, np.float64)
    dd = None if d is None else np.ascontiguousarray(d, np.float64)
    if use_plan and terms > 0:
        plan = np.zeros(terms, np.int64)
        o.orc_spgemm_plan(Lm, I(Lp), I(Li), blocks, Am, An, I(Ap), I(Ai), L(Cp), I(Ci), -1,
                          L(plan))
        o.orc_spgemm_numeric_planned(Lm, I(Lp), I(Li), D(Lx), blocks, D(dd), Am, I(Ap), D(Ax),
                                     L(plan), nnz, D(Cx))
    else:
        acc = np.zeros(max(An, 1), np.float64)
        o.orc_spgemm_numeric(Lm, I(Lp), I(Li), D(Lx), blocks, D(dd), Am, An, I(Ap), I(Ai),
                             D(Ax), L(Cp), I(Ci), D(acc), D(Cx))
    return Cp, Ci, Cx[:nnz], terms


def transpose(m, n, Ap, Ai, Ax=None):
    """Stable counting-sort transpose of an m x n CSR -> (ATp, ATi, ATx)."""
    o = lib()
    Ap, Ai = np.ascontiguousarray(Ap, np.int32), np.ascontiguousarray(Ai, np.int32)
    nnz = int(Ap[m])
    ATp = np.zeros(n + 1, np.int32)
    ATi = np.zeros(max(nnz, 1), np.int32)
    ATx = np.zeros(max(nnz, 1), np.float64) if Ax is not None else None
    iwork = np.zeros(max(n, 1), np.int32)
    o.orc_csr_transpose(m, n, I(Ap), I(Ai), D(Ax), I(ATp), I(ATi), D(ATx), I(iwork))
    return ATp, ATi[:nnz], (ATx[:nnz] if Ax is not None else None)


def btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d):
    """C = B^T diag(d) A through B's CSC mirror (matmul_dispatchers.c:335-375)."""
    Lp, Li, Lx = transpose(m, nB, Bp, Bi, Bx)
    return spgemm(nB, Lp, Li, Lx, 1, d, m, nA, Ap, Ai, Ax)


def spgemm_rows(rows, Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax):
    """Selected output rows of C = kron(I_blocks, L) diag(d) A, one at a time
    (gustavson.c:43-68, :78-101, :117-141 for one row): a list of
    (cols int32, vals float64).  For products too large to form whole."""
    o = lib()
    Lp, Li, Ap, Ai = (np.ascontiguousarray(v, np.int32) for v in (Lp, Li, Ap, Ai))
    Lx, Ax = np.ascontiguousarray(Lx, np.float64), np.ascontiguousarray(Ax, np.float64)
    mark = np.full(An, -1, np.int32)
    acc = np.zeros(An, np.float64)
    out = []
    for r in rows:
        b, rr = divmod(int(r), Lm)
        bound = int(sum(Ap[k + b * (Am // blocks) + 1] - Ap[k + b * (Am // blocks)]
                        for k in Li[Lp[rr]:Lp[rr + 1]]))
        cols = np.zeros(max(min(bound, An), 1), np.int32)
        vals = np.zeros(max(min(bound, An), 1), np.float64)
        cnt = o.orc_spgemm_row(int(r), Lm, I(Lp), I(Li), D(Lx), blocks,
                               D(np.ascontiguousarray(d, np.float64)) if d is not None else None,
                               Am, An, I(Ap), I(Ai), D(Ax), I(mark), D(acc), I(cols), D(vals))
        out.append((cols[:cnt].copy(), vals[:cnt].copy()))
    return out


def spgemm_range(r0, r1, Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax, cap):
    """Output rows [r0, r1) of C = kron(I_blocks, L) diag(d) A (the oracle's
    row routine over a range; arrays must already be contiguous int32 /
    float64): (Cp int64 from 0, Ci, Cx), or None if nnz would exceed cap.
    Releases the GIL during the C call, so threads may check blocks in
    parallel."""
    o = lib()
    mark = np.full(max(An, 1), -1, np.int32)
    acc = np.zeros(max(An, 1), np.float64)
    Cp = np.zeros(r1 - r0 + 1, np.int64)
    Ci = np.zeros(max(cap, 1), np.int32)
    Cx = np.zeros(max(cap, 1), np.float64)
    nnz = o.orc_spgemm_range(r0, r1, Lm, I(Lp), I(Li), D(Lx), blocks,
                             D(d) if d is not None else None, Am, An, I(Ap), I(Ai), D(Ax),
                             I(mark), D(acc), cap, L(Cp), I(Ci), D(Cx))
    if nnz < 0:
        return None
    return Cp, Ci[:nnz], Cx[:nnz]


def coo_coalesce(rows, cols, x, m, n):
    o = lib()
    rows, cols = np.ascontiguousarray(rows, np.int32), np.ascontiguousarray(cols, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    nnz = len(rows)
    Cp = np.zeros(m + 1, np.int64)
    Ci = np.zeros(max(nnz, 1), np.int32)
    Cx = np.zeros(max(nnz, 1), np.float64)
    out = o.orc_coo_coalesce(nnz, I(rows), I(cols), D(x), m, n, L(Cp), I(Ci), D(Cx))
    return Cp, Ci[:out], Cx[:out]


def row_reduce(m, n, Ap, Ai, Ax, group, m_out):
    o = lib()
    Ap, Ai, group = (np.ascontiguousarray(v, np.int32) for v in (Ap, Ai, group))
    nnzA = int(Ap[m])
    Cp = np.zeros(m_out + 1, np.int32)
    Ci = np.zeros(max(nnzA, 1), np.int32)
    mp = np.zeros(max(nnzA, 1), np.int32)
    nnz = o.orc_row_reduce_alloc(m, n, I(Ap), I(Ai), I(group), m_out, I(Cp), I(Ci), I(mp))
    Cx = np.zeros(max(nnz, 1), np.float64)
    o.orc_row_reduce_fill(nnzA, D(np.ascontiguousarray(Ax, np.float64)), I(mp), nnz, D(Cx))
    return Cp, Ci[:nnz], Cx[:nnz], mp[:nnzA]


def sum_n(m, n, ops, d=None):
    """Sum of same-shape CSR operands ops = [(p, i, x), ...]: returns (Cp, Ci,
    Cx, map) with map the output position of every operand entry, operands
    back to back.  d (len(ops) x m row scales) gives the diag-scaled sum."""
    o = lib()
    nops = len(ops)
    P = np.ascontiguousarray(np.concatenate([np.asarray(p, np.int32) for p, _, _ in ops]),
                             np.int32)
    nz = [int(p[m]) for p, _, _ in ops]
    off = np.zeros(nops + 1, np.int32)
    off[1:] = np.cumsum(nz)
    tot = int(off[-1])
    cat = (lambda arrs, t: np.ascontiguousarray(np.concatenate(arrs) if tot else np.zeros(1, t), t))
    Iall = cat([np.asarray(i, np.int32)[:k] for (_, i, _), k in zip(ops, nz)], np.int32)
    Xall = cat([np.asarray(x, np.float64)[:k] for (_, _, x), k in zip(ops, nz)], np.float64)
    Cp = np.zeros(m + 1, np.int32)
    Ci = np.zeros(max(tot, 1), np.int32)
    mp = np.zeros(max(tot, 1), np.int32)
    nnz = o.orc_sum_n_alloc(m, n, nops, I(P), I(Iall), I(off), I(Cp), I(Ci), I(mp))
    Cx = np.zeros(max(nnz, 1), np.float64)
    dd = None if d is None else np.ascontiguousarray(d, np.float64).reshape(-1)
    o.orc_sum_n_fill(m, nops, I(P), I(off), D(Xall), D(dd) if dd is not None else None, I(mp),
                     nnz, D(Cx))
    return Cp, Ci[:nnz], Cx[:nnz], mp[:tot]


def y_kron_i_vec(m, k, n, Y, w, lib_fn=None):
    """v = (Y kron I_m) w (Y k x n column-major, len(w) = m*n), the oracle's C
    restatement; lib_fn = ref_lib().Y_kron_I_vec runs the reference build."""
    Y = np.ascontiguousarray(Y, np.float64).reshape(-1)
    w = np.ascontiguousarray(w, np.float64).reshape(-1)
    v = np.zeros(max(m * k, 1))
    (lib_fn or lib().orc_y_kron_i_vec)(m, k, n, D(Y), D(w), D(v))
    return v[:m * k]


def i_kron_xt_vec(m, k, n, X, w, lib_fn=None):
    """v = (I_n kron X^T) w (X m x k column-major, len(w) = m*n)."""
    X = np.ascontiguousarray(X, np.float64).reshape(-1)
    w = np.ascontiguousarray(w, np.float64).reshape(-1)
    v = np.zeros(max(k * n, 1))
    (lib_fn or lib().orc_i_kron_xt_vec)(m, k, n, D(X), D(w), D(v))
    return v[:k * n]


def row_gather(Ap, Ai, Ax, rmap):
    o = lib()
    rmap = np.ascontiguousarray(rmap, np.int32)
    Ap, Ai = np.ascontiguousarray(Ap, np.int32), np.ascontiguousarray(Ai, np.int32)
    m_out = len(rmap)
    tot = int(sum(int(Ap[r + 1] - Ap[r]) for r in rmap))
    Cp = np.zeros(m_out + 1, np.int32)
    Ci = np.zeros(max(tot, 1), np.int32)
    Cx = np.zeros(max(tot, 1), np.float64)
    o.orc_row_gather(m_out, I(rmap), I(Ap), I(Ai), D(np.ascontiguousarray(Ax, np.float64)),
                     I(Cp), I(Ci), D(Cx))
    return Cp, Ci[:tot], Cx[:tot]


def da(m, Ap, Ax, d):
    o = lib()
    Ap = np.ascontiguousarray(Ap, np.int32)
    Cx = np.zeros(max(int(Ap[m]), 1), np.float64)
    o.orc_da(m, I(Ap), D(np.ascontiguousarray(Ax, np.float64)), D(np.ascontiguousarray(d)), D(Cx))
    return Cx[:int(Ap[m])]


def kron_spmv(m, n, Ap, Ai, Ax, x, p):
    o = lib()
    y = np.zeros(m * p, np.float64)
    o.orc_kron_spmv(m, n, I(np.ascontiguousarray(Ap, np.int32)),
                    I(np.ascontiguousarray(Ai, np.int32)), D(np.ascontiguousarray(Ax)),
                    D(np.ascontiguousarray(x)), D(y), p)
    return y


def coo_lower(m, Ap, Ai, Ax):
    o = lib()
    nnz = int(Ap[m])
    rows = np.zeros(max(nnz, 1), np.int32)
    cols = np.zeros(max(nnz, 1), np.int32)
    x = np.zeros(max(nnz, 1), np.float64)
    vm = np.zeros(max(nnz, 1), np.int32)
    c = o.orc_coo_lower(m, I(np.ascontiguousarray(Ap, np.int32)),
                        I(np.ascontiguousarray(Ai, np.int32)), D(np.ascontiguousarray(Ax)),
                        I(rows), I(cols), D(x), I(vm))
    return rows[:c], cols[:c], x[:c], vm[:c]


# ---------------------------------------------------------------------------
# oracle/_ref: the reference's own sparse sources (gustavson.c, CSC_matrix.c,
# CSR_matrix.c, sparse_matrix.c, ...) built by oracle/build_ref.sh.

class _CSR(C.Structure):
    _fields_ = [("p", ip), ("i", ip), ("x", dp), ("m", C.c_int), ("n", C.c_int),
                ("nnz", C.c_int)]


class _COO(C.Structure):  # COO_matrix, include/sparsediff_linalg/COO_matrix.h:35-44
    _fields_ = [("rows", ip), ("cols", ip), ("x", dp), ("value_map", ip), ("m", C.c_int),
                ("n", C.c_int), ("nnz", C.c_int)]


class _Rows(C.Structure):  # sparse_rows, src/internal/gustavson.h:26-32
    _fields_ = [("p", ip), ("i", ip), ("x", dp), ("m", C.c_int)]


class _Mat(C.Structure):  # struct matrix, include/sparsediff_linalg/matrix.h:64-88
    pass


_Mat._fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int), ("x", dp),
                 ("kind", C.c_int), ("values_version", C.c_uint64),
                 ("to_csr", C.CFUNCTYPE(C.POINTER(_CSR), C.POINTER(_Mat))),
                 ("row_gather_alloc", C.CFUNCTYPE(C.POINTER(_Mat), C.POINTER(_Mat), ip,
                                                  C.c_int)),
                 ("row_gather_fill_values", C.CFUNCTYPE(None, C.POINTER(_Mat),
                                                        C.POINTER(_Mat))),
                 ("row_reduce_alloc", C.CFUNCTYPE(C.POINTER(_Mat), C.POINTER(_Mat), ip,
                                                  C.c_int)),
                 ("row_reduce_fill_values", C.CFUNCTYPE(None, C.POINTER(_Mat),
                                                        C.POINTER(_Mat)))]

_REF_SIGS = {
    "csr_to_csc_alloc": (C.POINTER(_CSR), [C.POINTER(_CSR), ip]),
    "csr_to_csc_fill_values": (None, [C.POINTER(_CSR), C.POINTER(_CSR), ip]),
    "free_CSC_matrix": (None, [C.POINTER(_CSR)]),
    "free_CSR_matrix": (None, [C.POINTER(_CSR)]),
    "sparse_product_alloc": (C.POINTER(_Mat), [C.POINTER(_Rows), C.c_int, C.POINTER(_CSR)]),
    "sparse_product_fill_values": (None, [C.POINTER(_Rows), C.c_int, dp, C.POINTER(_CSR),
                                          C.POINTER(_Mat)]),
    "sparse_matrix_free": (None, [C.POINTER(_Mat)]),
    "new_sparse_matrix": (C.POINTER(_Mat), [C.POINTER(_CSR)]),
    "new_CSR_matrix": (C.POINTER(_CSR), [C.c_int, C.c_int, C.c_int]),
    "sum_csr_alloc": (None, [C.POINTER(_CSR), C.POINTER(_CSR), C.POINTER(_CSR)]),
    "sum_csr_fill_values": (None, [C.POINTER(_CSR), C.POINTER(_CSR), C.POINTER(_CSR)]),
    "sum_scaled_csr_matrices_fill_values": (None, [C.POINTER(_CSR), C.POINTER(_CSR),
                                                   C.POINTER(_CSR), dp, dp]),
    "sum_4_csr_alloc": (C.POINTER(_CSR), [C.POINTER(_CSR), C.POINTER(_CSR), C.POINTER(_CSR),
                                          C.POINTER(_CSR), C.POINTER(ip)]),
    "sum_n_csr_alloc": (C.POINTER(_CSR), [C.POINTER(C.POINTER(_CSR)), C.c_int, ip]),
    "accumulator": (None, [dp, C.c_int, ip, dp]),
    "Y_kron_I_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "I_kron_XT_vec": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    "new_COO_matrix_lower_triangular": (C.POINTER(_COO), [C.POINTER(_CSR)]),
    "refresh_lower_triangular_coo": (None, [C.POINTER(_COO), dp]),
    "free_COO_matrix": (None, [C.POINTER(_COO)]),
}

_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH, mode=os.RTLD_LOCAL)
        for name, (res, args) in _REF_SIGS.items():
            f = getattr(_ref, name)
            f.restype = res
            f.argtypes = args
    return _ref


def _csr_view(m, n, p, i, x, keep):
    p = np.ascontiguousarray(p, np.int32)
    i = np.ascontiguousarray(i, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    keep.extend([p, i, x])
    return _CSR(I(p), I(i), D(x), m, n, int(p[m]))


def ref_btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d, timing=None):
    """The reference's own BTA/BTDA sparse path (matmul_dispatchers.c:335-375
    minus the dispatch): CSC mirror of B, sparse_product_alloc (Gustavson +
    plan-or-accumulator policy), sparse_product_fill_values.  timing, if a
    dict, receives 'alloc_s' and 'fill_s' wall times of the two phases."""
    import time
    r = ref_lib()
    keep = []
    A = _csr_view(m, nA, Ap, Ai, Ax, keep)
    B = _csr_view(m, nB, Bp, Bi, Bx, keep)
    iwork = np.zeros(max(nB, 1), np.int32)
    t0 = time.perf_counter()
    Bc = r.csr_to_csc_alloc(C.byref(B), I(iwork))
    r.csr_to_csc_fill_values(C.byref(B), Bc, I(iwork))
    Lr = _Rows(Bc.contents.p, Bc.contents.i, Bc.contents.x, Bc.contents.n)
    Cm = r.sparse_product_alloc(C.byref(Lr), 1, C.byref(A))
    t1 = time.perf_counter()
    dd = np.ascontiguousarray(d, np.float64) if d is not None else None
    r.sparse_product_fill_values(C.byref(Lr), 1, D(dd), C.byref(A), Cm)
    t2 = time.perf_counter()
    if timing is not None:
        timing["alloc_s"] = t1 - t0
        timing["fill_s"] = t2 - t1
    Cc = Cm.contents.to_csr(Cm).contents
    nnz = Cc.nnz
    Cp = np.ctypeslib.as_array(Cc.p, (nB + 1,)).copy()
    Ci = np.ctypeslib.as_array(Cc.i, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    Cx = np.ctypeslib.as_array(Cc.x, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0)
    r.sparse_matrix_free(Cm)
    r.free_CSC_matrix(Bc)
    return Cp, Ci, Cx


def _ref_out(Cc, m):
    nnz = Cc.nnz
    Cp = np.ctypeslib.as_array(Cc.p, (m + 1,)).copy()
    Ci = np.ctypeslib.as_array(Cc.i, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    return Cp, Ci


def ref_sum(m, n, ops, d=None):
    """The reference's own CSR sums (CSR_sum.c): two operands through
    sum_csr_alloc and sum_csr_fill_values (sum_scaled_csr_matrices_fill_values
    with d = (d1, d2)), four through sum_4_csr_alloc and accumulator, any other
    count through sum_n_csr_alloc and accumulator.  Returns (Cp, Ci, Cx)."""
    r = ref_lib()
    keep = []
    views = [_csr_view(m, n, p, i, x, keep) for p, i, x in ops]
    if len(ops) == 2:
        cap = views[0].nnz + views[1].nnz
        Cc = r.new_CSR_matrix(m, n, max(cap, 1))
        r.sum_csr_alloc(C.byref(views[0]), C.byref(views[1]), Cc)
        if d is None:
            r.sum_csr_fill_values(C.byref(views[0]), C.byref(views[1]), Cc)
        else:
            d1, d2 = (np.ascontiguousarray(v, np.float64) for v in d)
            keep.extend([d1, d2])
            r.sum_scaled_csr_matrices_fill_values(C.byref(views[0]), C.byref(views[1]), Cc,
                                                  D(d1), D(d2))
        Cp, Ci = _ref_out(Cc.contents, m)
        nnz = int(Cp[m])
        Cx = np.ctypeslib.as_array(Cc.contents.x, (max(nnz, 1),))[:nnz].copy()
        r.free_CSR_matrix(Cc)
        return Cp, Ci, Cx
    assert d is None
    tot = sum(v.nnz for v in views)
    maps = [np.zeros(max(v.nnz, 1), np.int32) for v in views]
    if len(ops) == 4:
        marr = (ip * 4)(*[I(mp) for mp in maps])
        Cc = r.sum_4_csr_alloc(*[C.byref(v) for v in views], marr)
        # sum_4_csr_alloc allocates the maps itself (CSR_sum.c:170-173)
        maps = [np.ctypeslib.as_array(marr[k], (views[k].nnz,)).copy() if views[k].nnz
                else np.zeros(1, np.int32) for k in range(4)]
    else:
        flat = np.zeros(max(tot, 1), np.int32)
        parr = (C.POINTER(_CSR) * len(views))(*[C.pointer(v) for v in views])
        Cc = r.sum_n_csr_alloc(parr, len(views), I(flat))
        o = 0
        for k, v in enumerate(views):
            maps[k] = flat[o:o + v.nnz].copy()
            o += v.nnz
    Cp, Ci = _ref_out(Cc.contents, m)
    nnz = int(Cp[m])
    Cx = np.zeros(max(nnz, 1), np.float64)
    for (p, i, x), v, mp in zip(ops, views, maps):
        if v.nnz:
            xv = np.ascontiguousarray(x, np.float64)
            mp = np.ascontiguousarray(mp, np.int32)
            r.accumulator(D(xv), v.nnz, I(mp), D(Cx))
    r.free_CSR_matrix(Cc)
    return Cp, Ci, Cx[:nnz]


def ref_lower_coo_timer(m, n, p, i, x, seconds):
    """The reference's lower-triangular export (COO_matrix.c:48-90) once, then
    its refresh (COO_matrix.c:92-98) repeated for about `seconds` on one core.
    Returns (exported entries, refreshes, refresh seconds)."""
    import time
    r = ref_lib()
    keep = []
    A = _csr_view(m, n, p, i, x, keep)
    coo = r.new_COO_matrix_lower_triangular(C.byref(A))
    cnt = coo.contents.nnz
    xv = keep[2]
    fills, t = 0, 0.0
    while t < seconds or fills == 0:
        t0 = time.perf_counter()
        r.refresh_lower_triangular_coo(coo, D(xv))
        t += time.perf_counter() - t0
        fills += 1
    r.free_COO_matrix(coo)
    return cnt, fills, t


def ref_fill_timer(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, ds):
    """Times the reference's alloc once and its fill once per d in ds (the
    single-thread CPU baseline).  Returns (alloc_s, [fill_s...], last Cx)."""
    import time
    r = ref_lib()
    keep = []
    A = _csr_view(m, nA, Ap, Ai, Ax, keep)
    B = _csr_view(m, nB, Bp, Bi, Bx, keep)
    iwork = np.zeros(max(nB, 1), np.int32)
    t0 = time.perf_counter()
    Bc = r.csr_to_csc_alloc(C.byref(B), I(iwork))
    r.csr_to_csc_fill_values(C.byref(B), Bc, I(iwork))
    Lr = _Rows(Bc.contents.p, Bc.contents.i, Bc.contents.x, Bc.contents.n)
    Cm = r.sparse_product_alloc(C.byref(Lr), 1, C.byref(A))
    alloc_s = time.perf_counter() - t0
    fills = []
    for d in ds:
        dd = np.ascontiguousarray(d, np.float64)
        t1 = time.perf_counter()
        r.sparse_product_fill_values(C.byref(Lr), 1, D(dd), C.byref(A), Cm)
        fills.append(time.perf_counter() - t1)
    Cc = Cm.contents.to_csr(Cm).contents
    Cx = np.ctypeslib.as_array(Cc.x, (max(Cc.nnz, 1),))[:Cc.nnz].copy()
    r.sparse_matrix_free(Cm)
    r.free_CSC_matrix(Bc)
    return alloc_s, fills, Cx


def port_fill_timer(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, ds):
    """Times the C restatement: alloc (transpose + symbolic + plan) once, then
    one fill per d in ds through the plan or accumulator (the reference's
    policy).  Returns (alloc_s, [fill_s...])."""
    import time
    o = lib()
    t0 = time.perf_counter()
    Lp, Li, Lx = transpose(m, nB, Bp, Bi, Bx)
    Lp, Li = np.ascontiguousarray(Lp, np.int32), np.ascontiguousarray(Li, np.int32)
    Ap, Ai = np.ascontiguousarray(Ap, np.int32), np.ascontiguousarray(Ai, np.int32)
    Cp = np.zeros(nB + 1, np.int64)
    nnz = o.orc_spgemm_symbolic(nB, I(Lp), I(Li), 1, m, nA, I(Ap), I(Ai), L(Cp), None)
    Ci = np.zeros(max(nnz, 1), np.int32)
    o.orc_spgemm_symbolic(nB, I(Lp), I(Li), 1, m, nA, I(Ap), I(Ai), L(Cp), I(Ci))
    terms = o.orc_spgemm_plan(nB, I(Lp), I(Li), 1, m, nA, I(Ap), I(Ai), L(Cp), I(Ci), 3 * nnz,
                              None)
    plan = None
    if terms > 0:
        plan = np.zeros(terms, np.int64)
        o.orc_spgemm_plan(nB, I(Lp), I(Li), 1, m, nA, I(Ap), I(Ai), L(Cp), I(Ci), 3 * nnz, L(plan))
    alloc_s = time.perf_counter() - t0
    Cx = np.zeros(max(nnz, 1))
    acc = np.zeros(max(nA, 1))
    Lx = np.ascontiguousarray(Lx)
    Ax = np.ascontiguousarray(Ax, np.float64)
    fills = []
    for d in ds:
        dd = np.ascontiguousarray(d, np.float64)
        t1 = time.perf_counter()
        if plan is not None:
            o.orc_spgemm_numeric_planned(nB, I(Lp), I(Li), D(Lx), 1, D(dd), m, I(Ap), D(Ax),
                                         L(plan), nnz, D(Cx))
        else:
            o.orc_spgemm_numeric(nB, I(Lp), I(Li), D(Lx), 1, D(dd), m, nA, I(Ap), I(Ai), D(Ax),
                                 L(Cp), I(Ci), D(acc), D(Cx))
        fills.append(time.perf_counter() - t1)
    return alloc_s, fills
