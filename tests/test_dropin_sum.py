"""GPU: the drop-in matrix sums (include/sparsediff_linalg/matrix_sum.h) against
the oracle, bit for bit.  Mirrors the reference's tests of the sums
(test_matmul_dispatchers.h:2905-3137 and test_cache_coherence.h:544-556):
mixed operand kinds before and after mutating operands, sum_n against sum_4,
n = 1 and n = 5 literals, the incremental fill with an operand overwritten
after it was added, and one version bump per fill."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle import pd_oracle as PO
from tests import dropin as D
from tests.helpers import bits_equal, random_csr
from tests.test_oracle import SUM5_OPS

pytestmark = pytest.mark.gpu


def view_csr(M):
    """(p, i, x) of an oracle PD / SPD's CSR view (rows of the union of its
    blocks, columns ascending, values from the dense form)."""
    p, i = PO.csr_pattern(M)
    dn = PO.dense(M)
    rows = np.repeat(np.arange(M["m"]), np.diff(p))
    return np.array(p, np.int32), np.array(i, np.int32), dn[rows, np.array(i, int)]


def s4_fixture():
    """test_matmul_dispatchers.h:2872-2903: A sparse 6x5, B permuted_dense, D
    stacked_pd of two blocks, E sparse with an empty pattern."""
    A = ([0, 2, 3, 3, 5, 6, 7], [0, 4, 2, 1, 3, 0, 4], [1, 2, 3, 4, 5, 6, 7])
    B = PO.pd(6, 5, [1, 4], [0, 2, 4], np.array([10, 20, 30, 40, 50, 60.0]))
    D0 = PO.pd(6, 5, [0, 2], [1, 4], np.array([100, 200, 300, 400.0]))
    D1 = PO.pd(6, 5, [5], [0, 3, 4], np.array([500, 600, 700.0]))
    Dm = PO.spd(6, 5, [D0, D1])
    E = ([0] * 7, [], [])
    return A, B, Dm, E


def handles(A, B, Dm, E):
    hA = D.sparse(*A, 6, 5)
    hB = D.pd(6, 5, B["rp"], B["cp"], B["X"])
    hD = D.spd(6, 5, [D.pd(6, 5, b["rp"], b["cp"], b["X"]) for b in Dm["blocks"]])
    hE = D.sparse(*E, 6, 5)
    return [hA, hB, hD, hE]


def expect(ops_csr, m, n, d=None):
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops_csr, d=d)
    return Cp, Ci, Cx


def assert_same(Cm, exp):
    p, i, x = D.mat_csr(Cm)
    assert list(p) == list(exp[0]) and list(i) == list(exp[1])
    assert bits_equal(x, exp[2])
    assert bits_equal(D.mat_x(Cm), exp[2])


def test_sum_4_and_n_mixed_kinds():
    L = D.lib()
    A, B, Dm, E = s4_fixture()
    hs = handles(A, B, Dm, E)
    C4 = L.sum_4_matrices_alloc(*hs)
    arr = D.ops_array(hs)
    Cn = L.sum_n_matrices_alloc(arr, 4)
    assert C4.contents.kind == 0 and C4.contents.m == 6 and C4.contents.n == 5
    for rnd in range(2):
        if rnd == 1:  # mutate the stacked_pd's and B's values
            for b in Dm["blocks"]:
                b["X"] = -b["X"]
            D.set_x(hs[2], -D.mat_x(hs[2]))
            B["X"] = B["X"] * 0.5
            D.set_x(hs[1], D.mat_x(hs[1]) * 0.5)
        ops = [tuple(np.asarray(v) for v in A), view_csr(B), view_csr(Dm),
               (np.zeros(7, np.int32), np.zeros(0, np.int32), np.zeros(0))]
        exp = expect(ops, 6, 5)
        L.sum_4_matrices_fill_values(*hs, C4)
        L.sum_n_matrices_fill_values(arr, 4, Cn)
        assert_same(C4, exp)
        assert_same(Cn, exp)
        got = np.zeros((6, 5))
        p, i, x = D.mat_csr(Cn)
        for r in range(6):
            got[r, i[p[r]:p[r + 1]]] = x[p[r]:p[r + 1]]
        dense_ref = (np.zeros((6, 5)) + _dense(A) + PO.dense(B) + PO.dense(Dm))
        assert np.array_equal(got, dense_ref)
    for M in (C4, Cn, *hs):
        L.free_matrix(M)


def _dense(A):
    p, i, x = A
    out = np.zeros((6, 5))
    for r in range(6):
        out[r, i[p[r]:p[r + 1]]] = x[p[r]:p[r + 1]]
    return out


def test_sum_n_reference_literal_n1_n5():
    L = D.lib()
    hs = [D.sparse(p, i, x, 4, 6) for p, i, x in SUM5_OPS]
    hs[2] = hs[0]  # the reference passes M0 twice
    C1 = L.sum_n_matrices_alloc(D.ops_array(hs[:1]), 1)
    L.sum_n_matrices_fill_values(D.ops_array(hs[:1]), 1, C1)
    p, i, x = D.mat_csr(C1)
    assert list(p) == SUM5_OPS[0][0] and list(i) == SUM5_OPS[0][1] and list(x) == SUM5_OPS[0][2]
    arr = D.ops_array(hs)
    C5 = L.sum_n_matrices_alloc(arr, 5)
    L.sum_n_matrices_fill_values(arr, 5, C5)
    p, i, x = D.mat_csr(C5)
    assert list(p) == [0, 3, 5, 5, 8] and list(i) == [0, 1, 5, 2, 4, 0, 1, 5]
    assert bits_equal(x, O.sum_n(4, 6, SUM5_OPS)[2])
    for M in (C1, C5, hs[0], hs[1], hs[3], hs[4]):
        L.free_matrix(M)


def test_sum_n_fill_add_and_overwrite():
    L = D.lib()
    A, B, Dm, E = s4_fixture()
    hs = handles(A, B, Dm, E)
    arr = D.ops_array(hs)
    C = L.sum_n_matrices_alloc(arr, 4)
    Cref = L.sum_n_matrices_alloc(arr, 4)
    L.sum_n_matrices_fill_values(arr, 4, Cref)
    L.sum_n_matrices_fill_zero(C)
    assert np.all(D.mat_x(C) == 0)
    for k in range(4):
        L.sum_n_matrices_fill_add(hs[k], k, C)
    assert bits_equal(D.mat_x(C), D.mat_x(Cref))
    # A overwritten after it was added does not leak in
    exp = D.mat_x(Cref)
    L.sum_n_matrices_fill_zero(C)
    L.sum_n_matrices_fill_add(hs[0], 0, C)
    D.set_x(hs[0], -D.mat_x(hs[0]))
    for k in range(1, 4):
        L.sum_n_matrices_fill_add(hs[k], k, C)
    assert bits_equal(D.mat_x(C), exp)
    for M in (C, Cref, *hs):
        L.free_matrix(M)


def test_sum_matrices_scaled_and_versions():
    """sum_matrices_alloc into a caller-made sparse C (capacity nnzA + nnzB and a
    too-small one), both fills against the oracle and the reference build, one
    values_version bump per fill (test_cache_coherence.h:544-556)."""
    L = D.lib()
    rng = np.random.default_rng(8)
    m, n = 41, 33
    A = random_csr(rng, m, n, 0.2, empty_row=3)
    B = random_csr(rng, m, n, 0.15, dense_col=5)
    hA, hB = D.sparse(*A, m, n), D.sparse(*B, m, n)
    d1, d2 = rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)
    for cap in (len(A[1]) + len(B[1]), 1):
        S = L.new_sparse_matrix_alloc(m, n, cap)
        L.sum_matrices_alloc(hA, hB, S)
        v = S.contents.values_version
        L.sum_matrices_fill_values(hA, hB, S)
        assert S.contents.values_version == v + 1
        exp = expect([A, B], m, n)
        assert_same(S, exp)
        v = S.contents.values_version
        L.sum_scaled_matrices_fill_values(hA, hB, S, D.dptr(d1), D.dptr(d2))
        assert S.contents.values_version == v + 1
        assert_same(S, expect([A, B], m, n, d=np.stack([d1, d2])))
        L.free_matrix(S)
    L.free_matrix(hA)
    L.free_matrix(hB)


def test_sum_n_random_with_repeated_entries():
    """Six 3000 x 2000 operands (one repeating some (row, col) pairs, one empty,
    one passed twice): fill, refill after a change, and the incremental fill
    through both add paths, bit for bit against the oracle."""
    L = D.lib()
    rng = np.random.default_rng(17)
    m, n = 3000, 2000
    ops = [random_csr(rng, m, n, 0.002 * (k + 1)) for k in range(4)]
    p, i, x = ops[1]
    # repeat every 5th entry of operand 1 in place (rows stay sorted)
    reps = np.zeros(len(i), int)
    reps[::5] = 1
    rows = np.repeat(np.arange(m), np.diff(p))
    cnt = np.bincount(rows, weights=1 + reps, minlength=m).astype(np.int64)
    ii = np.repeat(i, 1 + reps)
    xx = np.repeat(x, 1 + reps) * np.where(np.arange(len(ii)) % 2, 1.0, -0.75)
    pp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    ops[1] = (pp, ii.astype(np.int32), xx)
    ops.append((np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    ops.append(ops[0])
    hs = [D.sparse(*o, m, n) for o in ops[:5]]
    hs.append(hs[0])
    arr = D.ops_array(hs)
    C = L.sum_n_matrices_alloc(arr, len(hs))
    L.sum_n_matrices_fill_values(arr, len(hs), C)
    assert_same(C, expect(ops, m, n))
    x2 = rng.standard_normal(len(ops[2][1]))
    ops[2] = (ops[2][0], ops[2][1], x2)
    D.set_x(hs[2], x2)
    L.sum_n_matrices_fill_values(arr, len(hs), C)
    assert_same(C, expect(ops, m, n))
    L.sum_n_matrices_fill_zero(C)
    for k in range(len(hs)):
        L.sum_n_matrices_fill_add(hs[k], k, C)
    assert_same(C, expect(ops, m, n))
    L.free_matrix(C)
    for h in hs[:5]:
        L.free_matrix(h)


def test_sum_edge_cases():
    """All operands empty (sums allocate an empty union and every fill is a
    no-op), zero rows, and sum_matrices_alloc into a zero-capacity C."""
    L = D.lib()
    empt = ([0] * 6, [], [])
    hs = [D.sparse(*empt, 5, 4) for _ in range(3)]
    arr = D.ops_array(hs)
    C = L.sum_n_matrices_alloc(arr, 3)
    L.sum_n_matrices_fill_values(arr, 3, C)
    L.sum_n_matrices_fill_zero(C)
    for k in range(3):
        L.sum_n_matrices_fill_add(hs[k], k, C)
    p, i, x = D.mat_csr(C)
    assert list(p) == [0] * 6 and len(i) == 0 and C.contents.nnz == 0
    S = L.new_sparse_matrix_alloc(5, 4, 0)
    L.sum_matrices_alloc(hs[0], hs[1], S)
    d = np.ones(5)
    L.sum_scaled_matrices_fill_values(hs[0], hs[1], S, D.dptr(d), D.dptr(d))
    assert S.contents.nnz == 0
    z = [D.sparse([0], [], [], 0, 3) for _ in range(2)]
    Cz = L.sum_n_matrices_alloc(D.ops_array(z), 2)
    L.sum_n_matrices_fill_values(D.ops_array(z), 2, Cz)
    assert Cz.contents.m == 0 and Cz.contents.nnz == 0
    for M in (C, S, Cz, *hs, *z):
        L.free_matrix(M)
