"""The reference-compatible C API (include/sparsediff_linalg/*.h) on the GPU,
sparse kind: known-answer vectors of the reference's tests, bit-exact
agreement with the oracle on random operands (empty rows/columns, dense
columns), the alloc/fill contracts (fills bump the output version once,
refresh operands iff their version moved, allocate no device memory)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import dropin as D
from tests.helpers import bits_equal, random_csr

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_literals.json")))


# ------------------------------------------------------------ raw structs ----
def test_raw_btda_literal_and_random():
    L = D.lib()
    case = GOLD["btda"][0]
    A, B = case["A"], case["B"]
    # B as CSC from its CSR arrays
    Bcp, Bci, Bcx = O.transpose(B["m"], B["n"], B["p"], B["i"], np.array(B["x"], float))
    Ac = D.new_csr(A["p"], A["i"], A["x"], A["m"], A["n"])
    Bc = D.new_csc(Bcp, Bci, Bcx, B["m"], B["n"])
    Cc = L.BTA_alloc(Ac, Bc)
    p, i, _ = D.read_csr(Cc)
    assert list(p) == case["C"]["p"] and list(i) == case["C"]["i"]
    L.BTDA_fill_values(Ac, Bc, D.dptr(case["d"]), None, Cc)
    assert list(D.read_csr(Cc)[2]) == case["C"]["x"]
    for S in (Cc, Ac):
        L.free_CSR_matrix(S)
    L.free_CSC_matrix(Bc)

    rng = np.random.default_rng(1)
    m, nA, nB = 120, 70, 50
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.06, dense_col=3, empty_row=0)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.08, empty_col=1)
    d = rng.uniform(0.1, 10, m)
    Cp, Ci, Cx, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    Bcp, Bci, Bcx = O.transpose(m, nB, Bp, Bi, Bx)
    Ac, Bc = D.new_csr(Ap, Ai, Ax, m, nA), D.new_csc(Bcp, Bci, Bcx, m, nB)
    Cc = L.BTA_alloc(Ac, Bc)
    L.BTDA_fill_values(Ac, Bc, D.dptr(d), None, Cc)
    p, i, x = D.read_csr(Cc)
    assert list(p) == list(Cp) and list(i) == list(Ci) and bits_equal(x, Cx)
    L.BTDA_fill_values(Ac, Bc, None, None, Cc)  # d = NULL is the identity
    _, _, Cx1, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, None)
    assert bits_equal(D.read_csr(Cc)[2], Cx1)
    L.free_CSR_matrix(Cc)
    L.free_CSR_matrix(Ac)
    L.free_CSC_matrix(Bc)


@pytest.mark.parametrize("nA,ratio_pull", [(400, True), (12, False)])
def test_raw_btda_unsorted_csc_rows(nA, ratio_pull):
    """A caller-made CSC B whose row indices are shuffled within each column
    (one column also repeats a row index): BTA_alloc / BTDA_fill_values fold
    every entry's terms in B's storage order, as the reference does
    (gustavson.c:127-135), through the pull schedule (terms <= 3 nnz) and the
    row accumulator (terms > 3 nnz), bit for bit against the oracle walking the
    same storage order.  Entries with three or more terms make the order
    visible in the last bits."""
    L = D.lib()
    rng = np.random.default_rng(31 + nA)
    m, nB = 600, 40
    Ap, Ai, Ax = random_csr(rng, m, nA, min(0.5, 6.0 / nA))
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.12)
    Bcp, Bci, Bcx = (np.array(v) for v in O.transpose(m, nB, Bp, Bi, Bx))
    for c in range(nB):  # shuffle the row indices of every column
        s0, s1 = Bcp[c], Bcp[c + 1]
        perm = rng.permutation(s1 - s0) + s0
        Bci[s0:s1], Bcx[s0:s1] = Bci[perm].copy(), Bcx[perm].copy()
    c = int(np.argmax(np.diff(Bcp)))  # a repeated row index in the longest column
    Bci[Bcp[c] + 1] = Bci[Bcp[c]]
    Bcx = Bcx * rng.uniform(0.5, 2.0, len(Bcx))
    d = rng.uniform(0.1, 10, m)
    Cp, Ci, Cx, terms = O.spgemm(nB, Bcp, Bci, Bcx, 1, d, m, nA, Ap, Ai, Ax)
    assert (terms <= 3 * len(Ci)) == ratio_pull
    assert np.max(np.diff(Cp)) > 0
    Ac, Bc = D.new_csr(Ap, Ai, Ax, m, nA), D.new_csc(Bcp, Bci, Bcx, m, nB)
    Cc = L.BTA_alloc(Ac, Bc)
    L.BTDA_fill_values(Ac, Bc, D.dptr(d), None, Cc)
    p, i, x = D.read_csr(Cc)
    assert list(p) == list(Cp) and list(i) == list(Ci) and bits_equal(x, Cx)
    # the order is visible: folding each entry in ascending k (columns sorted)
    # gives different last bits somewhere
    Sci, Scx = Bci.copy(), Bcx.copy()
    for c in range(nB):
        s0, s1 = Bcp[c], Bcp[c + 1]
        o = np.argsort(Bci[s0:s1], kind="stable") + s0
        Sci[s0:s1], Scx[s0:s1] = Bci[o], Bcx[o]
    _, _, Sx, _ = O.spgemm(nB, Bcp, Sci, Scx, 1, d, m, nA, Ap, Ai, Ax)
    assert not bits_equal(Sx, Cx)
    L.free_CSR_matrix(Cc)
    L.free_CSR_matrix(Ac)
    L.free_CSC_matrix(Bc)


@pytest.mark.parametrize("case", GOLD["ata"], ids=lambda c: c["name"])
def test_raw_ata_literals(case):
    L = D.lib()
    Ac = case["A_csc"]
    m, n = Ac["m"], Ac["n"]
    Ap, Ai, Ax = O.transpose(n, m, Ac["p"], Ac["i"], np.array(Ac["x"], float))
    csc = D.new_csc(Ac["p"], Ac["i"], Ac["x"], m, n)
    csr = D.new_csr(Ap, Ai, Ax, m, n)
    Cm = L.ATA_alloc(csc, csr)
    p, i, _ = D.read_csr(Cm)
    assert list(p) == case["C"]["p"] and list(i) == case["C"]["i"]
    if case["d"] is not None:
        L.ATDA_fill_values(csc, D.dptr(case["d"]), csr, None, Cm)
        x = D.read_csr(Cm)[2]
        exp = np.array(case["C"]["x"])
        if case["tol"] == 0:
            assert list(x) == list(exp)
        else:
            assert np.all(np.abs(x - exp) <= np.maximum(case["tol"], case["tol"] * np.abs(exp)))
        _, _, ox, _ = O.spgemm(n, Ac["p"], Ac["i"], np.array(Ac["x"], float), 1, case["d"], m, n,
                               Ap, Ai, Ax)
        assert bits_equal(x, ox)
    L.free_CSR_matrix(Cm)
    L.free_CSR_matrix(csr)
    L.free_CSC_matrix(csc)


@pytest.mark.parametrize("case", GOLD["transpose"], ids=lambda c: c["name"])
def test_raw_transposes(case):
    L = D.lib()
    A = case["A"]
    m, n = A["m"], A["n"]
    x = np.array(A["x"], float)
    csr = D.new_csr(A["p"], A["i"], x, m, n)
    # transpose / AT_alloc + AT_fill (CSR_matrix.c:99-205)
    T = L.transpose(csr, None)
    Tp, Ti, Tx = D.read_csr(T)
    Op, Oi, Ox = O.transpose(m, n, A["p"], A["i"], x)
    assert list(Tp) == list(Op) and list(Ti) == list(Oi) and bits_equal(Tx, Ox)
    if case["AT"] is not None:
        assert list(Tp) == case["AT"]["p"] and list(Ti) == case["AT"]["i"]
        assert list(Tx) == case["AT"]["x"]
    AT = L.AT_alloc(csr, None)
    L.AT_fill_values(csr, AT, None)
    assert list(D.read_csr(AT)[0]) == list(Op) and bits_equal(D.read_csr(AT)[2], Ox)
    # csr_to_csc then csc_to_csr round trip
    csc = L.csr_to_csc_alloc(csr, None)
    L.csr_to_csc_fill_values(csr, csc, None)
    cp, ci, cx = D.read_csc(csc)
    assert list(cp) == list(Op) and list(ci) == list(Oi) and bits_equal(cx, Ox)
    back = L.csc_to_csr_alloc(csc, None)
    L.csc_to_csr_fill_values(csc, back, None)
    bp, bi, bx = D.read_csr(back)
    assert list(bp) == A["p"] and list(bi) == A["i"] and bits_equal(bx, x)
    for S in (T, AT, back, csr):
        L.free_CSR_matrix(S)
    L.free_CSC_matrix(csc)


def test_raw_coo():
    L = D.lib()
    c0 = GOLD["coo"][0]
    A = c0["A"]
    csr = D.new_csr(A["p"], A["i"], A["x"], A["m"], A["n"])
    coo = L.new_COO_matrix(csr).contents
    nnz = coo.nnz
    assert list(np.ctypeslib.as_array(coo.rows, (nnz,))) == c0["rows"]
    assert list(np.ctypeslib.as_array(coo.cols, (nnz,))) == c0["cols"]
    c1 = GOLD["coo"][1]
    A = c1["A"]
    csr2 = D.new_csr(A["p"], A["i"], A["x"], A["m"], A["n"])
    lowp = L.new_COO_matrix_lower_triangular(csr2)
    low = lowp.contents
    k = low.nnz
    assert list(np.ctypeslib.as_array(low.rows, (k,))) == c1["lower"]["rows"]
    assert list(np.ctypeslib.as_array(low.cols, (k,))) == c1["lower"]["cols"]
    assert list(np.ctypeslib.as_array(low.value_map, (k,))) == c1["lower"]["value_map"]
    assert list(np.ctypeslib.as_array(low.x, (k,))) == list(map(float, c1["lower"]["x"]))
    L.refresh_lower_triangular_coo(lowp, D.dptr(c1["refresh_vals"]))
    assert list(np.ctypeslib.as_array(low.x, (k,))) == list(map(float, c1["refresh_x"]))
    L.free_COO_matrix(lowp)


@pytest.mark.parametrize("seed", range(3))
def test_raw_coo_coalesce(seed):
    L = D.lib()
    rng = np.random.default_rng(seed)
    m, n, nnz = 200 + 100 * seed, 150, 3000
    rows = rng.integers(0, m, nnz).astype(np.int32)
    cols = rng.integers(0, n, nnz).astype(np.int32)
    # 30% duplicates of earlier keys
    dup = rng.random(nnz) < 0.3
    src = rng.integers(0, nnz, nnz)
    rows[dup] = rows[src[dup] % max(1, nnz)]
    cols[dup] = cols[src[dup] % max(1, nnz)]
    x = rng.standard_normal(nnz)
    Cp, Ci, Cx = O.coo_coalesce(rows, cols, x, m, n)
    from sparsediff_linalg_amd._lib import COO_matrix
    coo = COO_matrix(D.iptr(rows), D.iptr(cols), D.dptr(x), None, m, n, nnz)
    keep = (rows, cols, x)
    Cm = L.COO_coalesce_to_csr_alloc(C.byref(coo))
    L.COO_coalesce_to_csr_fill_values(C.byref(coo), Cm)
    p, i, v = D.read_csr(Cm)
    assert list(p) == list(Cp) and list(i) == list(Ci) and bits_equal(v, Cx)
    x2 = x * 3.0 - 1.0
    coo.x = D.dptr(x2)
    keep = keep + (x2,)
    L.COO_coalesce_to_csr_fill_values(C.byref(coo), Cm)
    _, _, Cx2 = O.coo_coalesce(rows, cols, x2, m, n)
    assert bits_equal(D.read_csr(Cm)[2], Cx2)
    del keep
    L.free_CSR_matrix(Cm)


# ---------------------------------------------------------- matrix handles ----
def _rand_pair(seed, m=90, nA=60, nB=45):
    rng = np.random.default_rng(seed)
    A = random_csr(rng, m, nA, 0.07, dense_col=2, empty_row=3, empty_col=5)
    B = random_csr(rng, m, nB, 0.09, dense_col=0, empty_col=7)
    d = rng.uniform(0.1, 10, m)
    return rng, A, B, d


@pytest.mark.parametrize("seed", range(3))
def test_matrix_btda_bta_version_and_no_alloc(seed):
    L = D.lib()
    rng, (Ap, Ai, Ax), (Bp, Bi, Bx), d = _rand_pair(seed)
    m, nA, nB = 90, 60, 45
    A, B = D.sparse(Ap, Ai, Ax, m, nA), D.sparse(Bp, Bi, Bx, m, nB)
    Cm = L.BTA_matrices_alloc(A, B)  # builds B^T A
    Cp, Ci, Cx, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    p, i, _ = D.mat_csr(Cm)
    assert list(p) == list(Cp) and list(i) == list(Ci)
    v0 = Cm.contents.values_version
    allocs = L.sdl_device_alloc_count()
    L.BTDA_matrices_fill_values(A, D.dptr(d), B, Cm)
    assert Cm.contents.values_version == v0 + 1
    assert bits_equal(D.mat_x(Cm), Cx)
    # BTA = BTDA with d = 1 (test_matmul_dispatchers.h:2140-2327)
    L.BTA_matrices_fill_values(A, B, Cm)
    _, _, Cx1, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, None)
    assert bits_equal(D.mat_x(Cm), Cx1)
    # mutate operands' host values, bump their versions: the fill refreshes
    Ax2, Bx2 = Ax * 1.5 - 0.25, -Bx
    D.set_x(A, Ax2)
    D.set_x(B, Bx2)
    L.BTDA_matrices_fill_values(A, D.dptr(d), B, Cm)
    _, _, Cx2, _ = O.btda(m, nA, nB, Ap, Ai, Ax2, Bp, Bi, Bx2, d)
    assert bits_equal(D.mat_x(Cm), Cx2)
    assert L.sdl_device_alloc_count() == allocs, "a fill allocated device memory"
    for M in (Cm, A, B):
        L.free_matrix(M)


def test_matrix_ata_ba_da_transpose():
    L = D.lib()
    rng, (Ap, Ai, Ax), (Bp, Bi, Bx), d = _rand_pair(7)
    m, nA, nB = 90, 60, 45
    A = D.sparse(Ap, Ai, Ax, m, nA)
    # ATA / ATDA (matmul_dispatchers.c:490-527): full product through A's CSC mirror
    Cm = L.ATA_matrices_alloc(A)
    L.ATDA_matrices_fill_values(A, D.dptr(d), Cm)
    ATp, ATi, ATx = O.transpose(m, nA, Ap, Ai, Ax)
    Op, Oi, Ox, _ = O.spgemm(nA, ATp, ATi, ATx, 1, d, m, nA, Ap, Ai, Ax)
    p, i, x = D.mat_csr(Cm)
    assert list(p) == list(Op) and list(i) == list(Oi) and bits_equal(x, Ox)
    L.free_matrix(Cm)
    # BA: C = B A with B sparse (nA x k) read through its CSR rows
    Gp, Gi, Gx = random_csr(rng, 30, m, 0.05, dense_col=4)
    G = D.sparse(Gp, Gi, Gx, 30, m)
    Cm = L.BA_matrices_alloc(G, A)
    L.BA_matrices_fill_values(G, A, Cm)
    Op, Oi, Ox, _ = O.spgemm(30, Gp, Gi, Gx, 1, None, m, nA, Ap, Ai, Ax)
    p, i, x = D.mat_csr(Cm)
    assert list(p) == list(Op) and list(i) == list(Oi) and bits_equal(x, Ox)
    L.free_matrix(Cm)
    # DA
    Cm = L.copy_sparsity_matrices_alloc(A)
    L.DA_matrices_fill_values(D.dptr(d), A, Cm)
    assert bits_equal(D.mat_x(Cm), O.da(m, Ap, Ax, d))
    L.free_matrix(Cm)
    # transpose
    T = L.transpose_matrices_alloc(A)
    L.transpose_matrices_fill_values(A, T)
    p, i, x = D.mat_csr(T)
    assert list(p) == list(ATp) and list(i) == list(ATi) and bits_equal(x, ATx)
    L.free_matrix(T)
    L.free_matrix(G)
    L.free_matrix(A)


def test_matrix_row_gather_and_reduce_literals():
    L = D.lib()
    g = GOLD["row_gather"][0]
    A = g["A"]
    M = D.sparse(A["p"], A["i"], A["x"], A["m"], A["n"])
    mp = np.array(g["map"], np.int32)
    Cm = M.contents.row_gather_alloc(M, D.iptr(mp), len(mp))
    p, i, _ = D.mat_csr(Cm)
    assert list(p) == g["C"]["p"] and list(i) == g["C"]["i"]
    mp[:] = -7  # the map is bound at alloc time
    Cm.contents.row_gather_fill_values(M, Cm)
    assert list(D.mat_x(Cm)) == g["C"]["x"]
    D.set_x(M, np.array(A["x"]) * g["refill_scale"])
    Cm.contents.row_gather_fill_values(M, Cm)
    assert list(D.mat_x(Cm)) == [v * g["refill_scale"] for v in g["C"]["x"]]
    L.free_matrix(Cm)
    L.free_matrix(M)

    r = GOLD["row_reduce"][0]
    A = r["A"]
    M = D.sparse(A["p"], A["i"], A["x"], A["m"], A["n"])
    grp = np.array(r["group"], np.int32)
    Cm = M.contents.row_reduce_alloc(M, D.iptr(grp), r["m_out"])
    p, i, _ = D.mat_csr(Cm)
    assert list(p) == r["C"]["p"] and list(i) == r["C"]["i"]
    D.set_x(Cm, np.full(Cm.contents.nnz, np.nan))  # poisoned: the fill overwrites
    Cm.contents.row_reduce_fill_values(M, Cm)
    assert list(D.mat_x(Cm)) == r["C"]["x"]
    L.free_matrix(Cm)
    L.free_matrix(M)


@pytest.mark.parametrize("seed", range(3))
def test_matrix_row_reduce_random(seed):
    L = D.lib()
    rng = np.random.default_rng(seed)
    m, n, m_out = 80, 33, 25
    Ap, Ai, Ax = random_csr(rng, m, n, 0.1, dense_col=1, empty_row=2)
    grp = rng.integers(0, m_out, m).astype(np.int32)
    Cp, Ci, Cx, _ = O.row_reduce(m, n, Ap, Ai, Ax, grp, m_out)
    M = D.sparse(Ap, Ai, Ax, m, n)
    Cm = M.contents.row_reduce_alloc(M, D.iptr(grp), m_out)
    Cm.contents.row_reduce_fill_values(M, Cm)
    p, i, x = D.mat_csr(Cm)
    assert list(p) == list(Cp) and list(i) == list(Ci) and bits_equal(x, Cx)
    L.free_matrix(Cm)
    L.free_matrix(M)


def test_matrix_vstack_sparse_and_versions():
    L = D.lib()
    rng = np.random.default_rng(4)
    parts = [random_csr(rng, mm, 12, 0.3) for mm in (5, 0, 7, 3)]
    ops = [D.sparse(p, i, x, len(p) - 1, 12) for (p, i, x) in parts]
    arr = D.ops_array(ops)
    Cm = L.vstack_matrices_alloc(arr, len(ops))
    v0 = Cm.contents.values_version
    L.vstack_matrices_fill_values(arr, len(ops), Cm)
    assert Cm.contents.values_version == v0 + 1
    ref_x = np.concatenate([x for (_, _, x) in parts])
    assert bits_equal(D.mat_x(Cm), ref_x)
    p, i, _ = D.mat_csr(Cm)
    ref_i = np.concatenate([i for (_, i, _) in parts])
    assert list(i) == list(ref_i) and p[-1] == len(ref_i)
    # nothing moved: no copy and no version bump
    L.vstack_matrices_fill_values(arr, len(ops), Cm)
    assert Cm.contents.values_version == v0 + 1
    # one operand moved: only it is copied, C bumps once
    newx = parts[2][2] * -2.0
    D.set_x(ops[2], newx)
    L.vstack_matrices_fill_values(arr, len(ops), Cm)
    assert Cm.contents.values_version == v0 + 2
    ref_x = np.concatenate([parts[0][2], parts[1][2], newx, parts[3][2]])
    assert bits_equal(D.mat_x(Cm), ref_x)
    L.free_matrix(Cm)
    for M in ops:
        L.free_matrix(M)


def test_matrix_kron_sparse():
    L = D.lib()
    rng = np.random.default_rng(9)
    m, n, p = 11, 8, 3
    Ap, Ai, Ax = random_csr(rng, m, n, 0.3)
    A = D.sparse(Ap, Ai, Ax, m, n)
    x = rng.standard_normal(n * p)
    y = np.zeros(m * p)
    L.BA_kron_vec_matrices(A, D.dptr(x), y.ctypes.data_as(C.POINTER(C.c_double)), p)
    assert bits_equal(y, O.kron_spmv(m, n, Ap, Ai, Ax, x, p))
    Jp, Ji, Jx = random_csr(rng, n * p, 10, 0.25, dense_col=2)
    J = D.sparse(Jp, Ji, Jx, n * p, 10)
    Cm = L.BA_kron_matrices_alloc(A, p, J)
    L.BA_kron_matrices_fill_values(A, p, J, Cm)
    Op, Oi, Ox, _ = O.spgemm(m, Ap, Ai, Ax, p, None, n * p, 10, Jp, Ji, Jx)
    q, i, xx = D.mat_csr(Cm)
    assert list(q) == list(Op) and list(i) == list(Oi) and bits_equal(xx, Ox)
    L.free_matrix(Cm)
    L.free_matrix(J)
    L.free_matrix(A)


# ------------------------------------------------- sizes and overflow ----
@pytest.mark.parametrize("shape", ["20000x5000", "2000000x500000"])
def test_transposes_at_size(shape):
    """The public transposes through the C ABI at the benchmark operands'
    sizes (CSC_matrix.c:62-189, CSR_matrix.c:99-205): csr_to_csc_alloc /
    _fill, csc_to_csr_alloc / _fill, transpose, AT_alloc / AT_fill_values,
    bit for bit against the oracle's stable counting-sort transpose, the
    fill repeated after the values change."""
    from sparsediff_linalg_amd import workloads as W
    L = D.lib()
    if shape == "20000x5000":
        w = W.btda_20000x5000(1)
        m, n, (p, i, x) = w["m"], w["nB"], (w["Bp"], w["Bi"], w["Bx"])
    else:
        m, n = 2_000_000, 500_000
        p, i, x = W.zipf_csr(2, m, n)
    Tp, Ti, Tx = O.transpose(m, n, p, i, x)
    A = D.new_csr(p, i, x, m, n)
    Cc = L.csr_to_csc_alloc(A, None)
    L.csr_to_csc_fill_values(A, Cc, None)
    cp, ci, cx = D.read_csc(Cc)
    assert np.array_equal(cp, Tp) and np.array_equal(ci, Ti) and bits_equal(cx, Tx)
    R = L.csc_to_csr_alloc(Cc, None)
    L.csc_to_csr_fill_values(Cc, R, None)
    rp, ri, rx = D.read_csr(R)
    assert np.array_equal(rp, p) and np.array_equal(ri, i) and bits_equal(rx, x)
    T = L.transpose(A, None)
    tp, ti, tx = D.read_csr(T)
    assert np.array_equal(tp, Tp) and np.array_equal(ti, Ti) and bits_equal(tx, Tx)
    AT = L.AT_alloc(A, None)
    x2 = -2.0 * np.asarray(x)
    C.memmove(A.contents.x, x2.ctypes.data, x2.nbytes)
    L.AT_fill_values(A, AT, None)
    L.csr_to_csc_fill_values(A, Cc, None)
    _, _, Tx2 = O.transpose(m, n, p, i, x2)
    assert bits_equal(D.read_csr(AT)[2], Tx2) and bits_equal(D.read_csc(Cc)[2], Tx2)
    for S in (R, T, AT, A):
        L.free_CSR_matrix(S)
    L.free_CSC_matrix(Cc)


def test_row_gather_alloc_no_int_overflow():
    """row_gather_alloc where n_idxs * n = 25000 * 100000 overflows int32
    (test_alloc_overflow.h:31-80): a valid CSR with the gathered nnz."""
    L = D.lib()
    n_idxs, n_cols = 25000, 100000
    p = np.arange(n_idxs + 1, dtype=np.int32)
    i = np.arange(n_idxs, dtype=np.int32)
    M = D.sparse(p, i, np.ones(n_idxs), n_idxs, n_cols)
    idx = np.arange(n_idxs, dtype=np.int32)
    out = M.contents.row_gather_alloc(M, D.iptr(idx), n_idxs)
    M.contents.row_gather_fill_values(M, out)
    cp, ci, cx = D.mat_csr(out)
    assert out.contents.nnz == n_idxs and len(ci) == n_idxs
    assert np.array_equal(cp, p) and np.array_equal(ci, i) and np.all(cx == 1.0)
    L.free_matrix(out)
    L.free_matrix(M)


def test_ata_alloc_beyond_int_max_exits():
    """ATA_matrices_alloc of the 2,000,000 x 500,000 rank-Zipf A: nnz(A^T A)
    = 6.7e9 does not fit the int32 CSR of the ABI, so the call prints the
    error and exits(1), as the reference does on int overflow
    (sparse_matrix.c:92-97); run in a child process."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from sparsediff_linalg_amd import workloads as W\n"
        "from tests import dropin as D\n"
        "p, i, x = W.zipf_csr(2, 2_000_000, 500_000)\n"
        "M = D.sparse(p, i, x, 2_000_000, 500_000)\n"
        "print('calling', flush=True)\n"
        "D.lib().ATA_matrices_alloc(M)\n"
        "print('returned', flush=True)\n") % (os.path.dirname(os.path.dirname(__file__)),)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 1, (r.returncode, r.stdout[-500:], r.stderr[-500:])
    assert "calling" in r.stdout and "returned" not in r.stdout
    assert "ATA_matrices_alloc" in r.stderr and "INT_MAX" in r.stderr, r.stderr[-500:]


_MISMATCH_CHILD = r'''
import ctypes as C, sys
sys.path.insert(0, %(root)r)
import numpy as np
from tests import dropin as D
from sparsediff_linalg_amd._lib import COO_matrix
L = D.lib()
case = %(case)r
m, n = 6, 5
p = np.array([0, 2, 3, 5, 6, 8, 9], np.int32)
i1 = np.array([0, 3, 1, 2, 4, 0, 1, 3, 2], np.int32)
i2 = i1.copy(); i2[3] = 3               # same pointers and nnz, one column moved
x = np.arange(1.0, 10.0)
d = np.ones(m)
A1, A2 = D.new_csr(p, i1, x, m, n), D.new_csr(p, i2, x, m, n)
if case == "AT":
    T = L.AT_alloc(A1, None); fill = lambda: L.AT_fill_values(A2, T, None)
elif case == "csr_to_csc":
    T = L.csr_to_csc_alloc(A1, None); fill = lambda: L.csr_to_csc_fill_values(A2, T, None)
elif case == "csc_to_csr":
    S1, S2 = D.new_csc(p[:6], i1[:8], x[:8], 5, 5), D.new_csc(p[:6], i2[:8], x[:8], 5, 5)
    T = L.csc_to_csr_alloc(S1, None); fill = lambda: L.csc_to_csr_fill_values(S2, T, None)
elif case == "BTDA":
    B1 = L.csr_to_csc_alloc(A1, None); L.csr_to_csc_fill_values(A1, B1, None)
    B2 = L.csr_to_csc_alloc(A2, None); L.csr_to_csc_fill_values(A2, B2, None)
    T = L.BTA_alloc(A1, B1)
    L.BTDA_fill_values(A1, B1, D.dptr(d), None, T)         # the alloc's structure: fine
    fill = lambda: L.BTDA_fill_values(A1, B2, D.dptr(d), None, T)
elif case == "ATDA":
    B1 = L.csr_to_csc_alloc(A1, None); L.csr_to_csc_fill_values(A1, B1, None)
    T = L.ATA_alloc(B1, A1)
    L.ATDA_fill_values(B1, D.dptr(d), A1, None, T)
    fill = lambda: L.ATDA_fill_values(B1, D.dptr(d), A2, None, T)
else:
    r = np.array([0, 1, 1, 2, 0], np.int32); c1 = np.array([0, 1, 1, 2, 3], np.int32)
    c2 = c1.copy(); c2[4] = 4
    v = np.ones(5)
    k1 = COO_matrix(D.iptr(r), D.iptr(c1), D.dptr(v), None, 3, 5, 5)
    k2 = COO_matrix(D.iptr(r), D.iptr(c2), D.dptr(v), None, 3, 5, 5)
    T = L.COO_coalesce_to_csr_alloc(C.byref(k1))
    L.COO_coalesce_to_csr_fill_values(C.byref(k1), T)
    fill = lambda: L.COO_coalesce_to_csr_fill_values(C.byref(k2), T)
print("calling", flush=True)
fill()
print("returned", flush=True)
'''


@pytest.mark.parametrize("case", ["AT", "csr_to_csc", "csc_to_csr", "BTDA", "ATDA", "COO"])
def test_fill_with_another_structure_exits(case):
    """A fill handed an operand whose structure differs from the alloc's (same
    dimensions, pointers and nnz; one column index moved) exits(1) with the
    alloc-mismatch message instead of gathering through the alloc's
    permutation (the fills check a structure fingerprint); child process."""
    import subprocess
    import sys
    code = _MISMATCH_CHILD % {"root": os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                              "case": case}
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout[-500:], r.stderr[-800:])
    assert "calling" in r.stdout and "returned" not in r.stdout
    assert "match the alloc" in r.stderr, r.stderr[-500:]
