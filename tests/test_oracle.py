"""CPU pinning of the oracle (oracle/sdl_oracle.c):

* every known-answer vector of the reference's own tests (tests/golden),
* bit-exact agreement with the reference's sparse sources built in
  oracle/_ref (same inputs, its Gustavson / transposition / row_reduce),
* closed forms on tiny inputs: a dense boolean product for the structure and
  a per-entry left fold in ascending k for the values (Python floats are
  individually rounded IEEE doubles, so the fold is exact), a dense
  scatter-add for coalescing, double transposition = identity.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import bits_equal, random_csr

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_literals.json")))


def _arr(v, dt):
    return np.array(v, dt)


# ------------------------------------------------------------ literals ----
@pytest.mark.parametrize("case", GOLD["spgemm"], ids=lambda c: c["name"])
def test_spgemm_literals(case):
    L, A = case["L"], case["A"]
    Cp, Ci, Cx, _ = O.spgemm(L["m"], L["p"], L["i"], L["x"], case["blocks"], case["d"],
                             A["m"], A["n"], A["p"], A["i"], A["x"])
    assert list(Cp) == case["C"]["p"]
    assert list(Ci) == case["C"]["i"]
    if case["C"]["x"] is not None:
        assert list(Cx) == case["C"]["x"]


@pytest.mark.parametrize("case", GOLD["ata"], ids=lambda c: c["name"])
def test_ata_literals(case):
    Ac = case["A_csc"]
    m, n = Ac["m"], Ac["n"]
    # A's CSR from its CSC (transpose of the n x m compressed columns)
    Ap, Ai, Ax = O.transpose(n, m, Ac["p"], Ac["i"], np.array(Ac["x"], float))
    Cp, Ci, Cx, _ = O.spgemm(n, Ac["p"], Ac["i"], np.array(Ac["x"], float), 1, case["d"], m, n,
                             Ap, Ai, Ax)
    assert list(Cp) == case["C"]["p"]
    assert list(Ci) == case["C"]["i"]
    if case["C"]["x"] is not None:
        exp = np.array(case["C"]["x"])
        if case["tol"] == 0:
            assert list(Cx) == list(exp)
        else:  # reference's is_equal_double (tests/test_helpers.c:10-18)
            assert np.all(np.abs(Cx - exp) <= np.maximum(case["tol"],
                                                         case["tol"] * np.maximum(abs(Cx), abs(exp))))


def test_btda_literal():
    case = GOLD["btda"][0]
    A, B = case["A"], case["B"]
    Cp, Ci, Cx, _ = O.btda(A["m"], A["n"], B["n"], A["p"], A["i"], A["x"], B["p"], B["i"],
                           B["x"], case["d"])
    assert list(Cp) == case["C"]["p"] and list(Ci) == case["C"]["i"]
    assert list(Cx) == case["C"]["x"]


@pytest.mark.parametrize("case", GOLD["transpose"], ids=lambda c: c["name"])
def test_transpose_literals(case):
    A = case["A"]
    Tp, Ti, Tx = O.transpose(A["m"], A["n"], A["p"], A["i"], np.array(A["x"], float))
    if case["AT"] is not None:
        assert list(Tp) == case["AT"]["p"] and list(Ti) == case["AT"]["i"]
        assert list(Tx) == case["AT"]["x"]
    # round trip (test_csr_csc_conversion.h:127-171)
    Rp, Ri, Rx = O.transpose(A["n"], A["m"], Tp, Ti, Tx)
    assert list(Rp) == A["p"] and list(Ri) == A["i"] and list(Rx) == list(map(float, A["x"]))


def test_coo_literals():
    c0 = GOLD["coo"][0]
    A = c0["A"]
    # new_COO_matrix is an expansion: rows from p, cols/x copied
    rows = np.repeat(np.arange(A["m"]), np.diff(A["p"]))
    assert list(rows) == c0["rows"]
    c1 = GOLD["coo"][1]
    A = c1["A"]
    r, c, x, vm = O.coo_lower(A["m"], A["p"], A["i"], np.array(A["x"], float))
    low = c1["lower"]
    assert list(r) == low["rows"] and list(c) == low["cols"]
    assert list(x) == list(map(float, low["x"])) and list(vm) == low["value_map"]
    assert list(np.array(c1["refresh_vals"], float)[vm]) == list(map(float, c1["refresh_x"]))


def test_row_reduce_literal():
    case = GOLD["row_reduce"][0]
    A = case["A"]
    x = np.array(A["x"], float)
    Cp, Ci, Cx, mp = O.row_reduce(A["m"], A["n"], A["p"], A["i"], x, case["group"],
                                  case["m_out"])
    assert list(Cp) == case["C"]["p"] and list(Ci) == case["C"]["i"]
    assert list(Cx) == case["C"]["x"]
    _, _, Cx2, _ = O.row_reduce(A["m"], A["n"], A["p"], A["i"], x * case["refill_scale"],
                                case["group"], case["m_out"])
    assert list(Cx2) == [v * case["refill_scale"] for v in case["C"]["x"]]


# ------------------------------------------------- closed forms (tiny) ----
def dense_fold(Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax):
    """C[i, j] = left fold over ascending k of (L[i,k] d[k]) A[k,j], built from
    the definition with a per-entry list; structure = boolean product."""
    nL = Am // blocks
    out = {}
    for b in range(blocks):
        for r in range(Lm):
            row = b * Lm + r
            terms = {}
            for q in range(Lp[r], Lp[r + 1]):
                k = Li[q] + b * nL
                v = float(Lx[q]) if d is None else float(Lx[q]) * float(d[k])
                for t in range(Ap[k], Ap[k + 1]):
                    terms.setdefault(int(Ai[t]), []).append((k, v * float(Ax[t])))
            for j, lst in terms.items():
                acc = 0.0
                for _, p in sorted(lst):
                    acc = acc + p
                out[(row, j)] = acc
    return out


@pytest.mark.parametrize("seed", range(6))
def test_spgemm_closed_form(seed):
    rng = np.random.default_rng(seed)
    blocks = 1 + seed % 3
    Lm, nL, An = 7 + seed, 6 + seed % 4, 9
    Lp, Li, Lx = random_csr(rng, Lm, nL, 0.4, dense_col=1, empty_row=2)
    Ap, Ai, Ax = random_csr(rng, nL * blocks, An, 0.35, empty_col=3)
    d = rng.uniform(0.1, 10.0, nL * blocks) if seed % 2 else None
    Cp, Ci, Cx, _ = O.spgemm(Lm, Lp, Li, Lx, blocks, d, nL * blocks, An, Ap, Ai, Ax)
    ref = dense_fold(Lm, Lp, Li, Lx, blocks, d, nL * blocks, An, Ap, Ai, Ax)
    got = {}
    for row in range(Lm * blocks):
        cols = list(Ci[Cp[row]:Cp[row + 1]])
        assert cols == sorted(set(cols))
        for e in range(Cp[row], Cp[row + 1]):
            got[(row, int(Ci[e]))] = float(Cx[e])
    assert set(got) == set(ref)
    for key, v in ref.items():
        assert np.float64(got[key]).view(np.int64) == np.float64(v).view(np.int64), key


def test_plan_and_accumulator_bit_identical():
    # gustavson.h:60-61 / test_linalg_sparse_matmuls.h:663-755
    rng = np.random.default_rng(3)
    Lp, Li, Lx = random_csr(rng, 30, 25, 0.3, dense_col=0)
    Ap, Ai, Ax = random_csr(rng, 25, 40, 0.2, dense_col=5)
    d = rng.uniform(0.1, 10, 25)
    _, _, X1, _ = O.spgemm(30, Lp, Li, Lx, 1, d, 25, 40, Ap, Ai, Ax, planned=True)
    _, _, X2, _ = O.spgemm(30, Lp, Li, Lx, 1, d, 25, 40, Ap, Ai, Ax, planned=False)
    assert bits_equal(X1, X2)


@pytest.mark.parametrize("seed", range(4))
def test_coo_coalesce_closed_form(seed):
    rng = np.random.default_rng(seed)
    m, n, nnz = 13, 11, 200
    rows = rng.integers(0, m, nnz).astype(np.int32)
    cols = rng.integers(0, n, nnz).astype(np.int32)
    x = rng.standard_normal(nnz)
    Cp, Ci, Cx = O.coo_coalesce(rows, cols, x, m, n)
    sums = {}
    for t in range(nnz):  # input order, starting from +0.0
        key = (int(rows[t]), int(cols[t]))
        sums[key] = sums.get(key, 0.0) + float(x[t])
    got = {(r, int(Ci[e])): float(Cx[e]) for r in range(m) for e in range(Cp[r], Cp[r + 1])}
    assert got == sums
    for r in range(m):
        seg = list(Ci[Cp[r]:Cp[r + 1]])
        assert seg == sorted(set(seg))
    # equals row_reduce with every triplet a one-entry row grouped by its row
    Ap = np.arange(nnz + 1, dtype=np.int32)
    Rp, Ri, Rx, _ = O.row_reduce(nnz, n, Ap, cols, x, rows, m)
    assert list(Rp) == list(Cp) and list(Ri) == list(Ci) and bits_equal(Rx, Cx)


def test_kron_spmv_closed_form():
    rng = np.random.default_rng(5)
    m, n, p = 9, 7, 3
    Ap, Ai, Ax = random_csr(rng, m, n, 0.4)
    x = rng.standard_normal(n * p)
    y = O.kron_spmv(m, n, Ap, Ai, Ax, x, p)
    for b in range(p):
        for i in range(m):
            s = 0.0
            for t in range(Ap[i], Ap[i + 1]):
                s = s + float(Ax[t]) * float(x[b * n + Ai[t]])
            assert y[b * m + i] == s


def test_da_and_gather_closed_form():
    rng = np.random.default_rng(8)
    Ap, Ai, Ax = random_csr(rng, 10, 6, 0.5)
    d = rng.uniform(0.1, 10, 10)
    Cx = O.da(10, Ap, Ax, d)
    rows = np.repeat(np.arange(10), np.diff(Ap))
    assert bits_equal(Cx, Ax * d[rows])
    rmap = np.array([3, 3, 0, 9, 5], np.int32)
    Gp, Gi, Gx = O.row_gather(Ap, Ai, Ax, rmap)
    ref_i = np.concatenate([Ai[Ap[r]:Ap[r + 1]] for r in rmap])
    ref_x = np.concatenate([Ax[Ap[r]:Ap[r + 1]] for r in rmap])
    assert list(Gi) == list(ref_i) and bits_equal(Gx, ref_x)


# ------------------------------------------------------------- sums ----
# test_matmul_dispatchers.h:3001-3012 (test_sum_n_matrices_n1_n5): five 4 x 6
# operands, M0 twice, one empty, a row empty in all of them.
SUM5_OPS = [([0, 2, 3, 3, 4], [1, 5, 2, 0], [1, 2, 3, 4]),
            ([0, 1, 3, 3, 3], [0, 2, 4], [5, 6, 7]),
            ([0, 2, 3, 3, 4], [1, 5, 2, 0], [1, 2, 3, 4]),
            ([0, 0, 0, 0, 0], [], []),
            ([0, 1, 1, 1, 3], [5, 1, 5], [8, 9, 10])]


def _dense(m, n, p, i, x):
    out = np.zeros((m, n))
    for r in range(m):
        out[r, np.asarray(i[p[r]:p[r + 1]], int)] = x[p[r]:p[r + 1]]
    return out


def test_sum_reference_literal_n5():
    Cp, Ci, Cx, _ = O.sum_n(4, 6, SUM5_OPS)
    assert list(Cp) == [0, 3, 5, 5, 8]
    assert list(Ci) == [0, 1, 5, 2, 4, 0, 1, 5]
    ref = sum(_dense(4, 6, *o) for o in SUM5_OPS)
    assert np.array_equal(_dense(4, 6, Cp, Ci, Cx), ref)
    # n = 1 copies the operand
    Cp, Ci, Cx, mp = O.sum_n(4, 6, SUM5_OPS[:1])
    assert list(Cp) == SUM5_OPS[0][0] and list(Ci) == SUM5_OPS[0][1]
    assert list(Cx) == SUM5_OPS[0][2] and list(mp) == [0, 1, 2, 3]


@pytest.mark.parametrize("nops", [1, 2, 3, 6])
def test_sum_closed_form(nops):
    """Structure: the dense union of the patterns.  Values: the dense sum taken
    operand by operand from zeros (x + 0.0 == x, so a missing entry changes
    nothing and the dense sum is the fold itself), bit for bit; the map sends
    every operand entry to its (row, col)."""
    rng = np.random.default_rng(40 + nops)
    m, n = 23, 17
    ops = [random_csr(rng, m, n, 0.05 + 0.05 * k, empty_row=k % 3) for k in range(nops)]
    Cp, Ci, Cx, mp = O.sum_n(m, n, ops)
    pat = np.zeros((m, n), bool)
    acc = np.zeros((m, n))
    for p, i, x in ops:
        pat |= _dense(m, n, p, i, np.ones(len(i))) != 0
        acc = acc + _dense(m, n, p, i, x)
    rows = np.repeat(np.arange(m), np.diff(Cp))
    assert np.array_equal(np.argwhere(pat), np.stack([rows, Ci], 1))
    assert bits_equal(Cx, acc[rows, Ci])
    o = 0
    for p, i, x in ops:
        r = np.repeat(np.arange(m), np.diff(p))
        assert np.array_equal(rows[mp[o:o + len(i)]], r)
        assert np.array_equal(Ci[mp[o:o + len(i)]], i)
        o += len(i)
    # scaled: each term the rounded product d[row] * x, summed in operand order
    if nops == 2:
        d = rng.uniform(-3, 3, (2, m))
        _, _, Sx, _ = O.sum_n(m, n, ops, d=d)
        sacc = d[0][:, None] * _dense(m, n, *ops[0]) + d[1][:, None] * _dense(m, n, *ops[1])
        assert bits_equal(Sx, sacc[rows, Ci])


def test_sum_repeated_entries_fold_in_order():
    """An operand repeating (row, col) contributes each copy in storage order."""
    ops = [([0, 3], [2, 2, 5], [1e16, 1.0, 2.0]), ([0, 1], [2], [-1e16])]
    Cp, Ci, Cx, mp = O.sum_n(1, 6, ops)
    assert list(Ci) == [2, 5] and list(mp) == [0, 0, 1, 0]
    assert Cx[0] == ((0.0 + 1e16) + 1.0) + -1e16 and Cx[1] == 2.0


# -------------------------------------------------- vs oracle/_ref ----
needs_ref = pytest.mark.skipif(not O.ref_available(),
                               reason="oracle/_ref not built (reference tree absent)")


@needs_ref
@pytest.mark.parametrize("seed", range(5))
def test_btda_matches_reference_build(seed):
    rng = np.random.default_rng(100 + seed)
    m, nA, nB = 60 + 20 * seed, 40, 35
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.08 + 0.04 * seed, dense_col=2, empty_row=1)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.1, empty_col=4)
    d = rng.uniform(0.1, 10, m)
    Cp, Ci, Cx, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    Rp, Ri, Rx = O.ref_btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    assert list(Cp) == list(Rp) and list(Ci) == list(Ri) and bits_equal(Cx, Rx)


@needs_ref
def test_btda_20000x5000_matches_reference_build():
    from sparsediff_linalg_amd import workloads
    w = workloads.btda_20000x5000(1)
    Cp, Ci, Cx, terms = O.btda(w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"], w["Bp"],
                               w["Bi"], w["Bx"], w["d"])
    Rp, Ri, Rx = O.ref_btda(w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"], w["Bp"],
                            w["Bi"], w["Bx"], w["d"])
    assert terms == 720000 and len(Ci) == 709850
    assert np.array_equal(Cp, Rp) and np.array_equal(Ci, Ri) and bits_equal(Cx, Rx)


# ------------------------------------------------ the product library ----
def _declared_functions():
    import re
    inc = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include")
    names = set()
    for root, _, files in os.walk(inc):
        for f in files:
            if not f.endswith(".h"):
                continue
            txt = open(os.path.join(root, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for mt in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
                name = mt.group(1)
                if name in ("if", "return", "sizeof") or txt[mt.start():].startswith("static"):
                    continue
                if txt[mt.start():mt.start() + 7] == "typedef":
                    continue
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    from sparsediff_linalg_amd import LIB_PATH
    lib = C.CDLL(LIB_PATH, mode=os.RTLD_LOCAL)
    missing = [n for n in sorted(_declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_spgemm_rows_equal_full_product():
    """The sampled-row oracle gives exactly the rows of the full product."""
    import numpy as _np
    from oracle import oracle as _O
    from tests.helpers import random_csr as _rc
    rng = _np.random.default_rng(21)
    m, nA, nB = 300, 70, 50
    Ap, Ai, Ax = _rc(rng, m, nA, 0.08)
    Bp, Bi, Bx = _rc(rng, m, nB, 0.1)
    d = rng.uniform(0.1, 10, m)
    Cp, Ci, Cx, _ = _O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    Lp, Li, Lx = _O.transpose(m, nB, Bp, Bi, Bx)
    rows = [0, 7, 13, nB - 1, 25]
    for r, (cols, vals) in zip(rows, _O.spgemm_rows(rows, nB, Lp, Li, Lx, 1, d, m, nA, Ap, Ai, Ax)):
        assert _np.array_equal(cols, Ci[Cp[r]:Cp[r + 1]])
        assert _np.array_equal(vals.view(_np.int64), Cx[Cp[r]:Cp[r + 1]].view(_np.int64))


@needs_ref
@pytest.mark.parametrize("nops", [1, 2, 3, 4, 5])
def test_sum_matches_reference_build(nops):
    """sum_csr_alloc / sum_csr_fill_values (2), sum_4_csr_alloc + accumulator (4)
    and sum_n_csr_alloc + accumulator (others), bit for bit."""
    rng = np.random.default_rng(300 + nops)
    m, n = 80, 60
    ops = [random_csr(rng, m, n, 0.04 * (k + 1), empty_row=k) for k in range(nops)]
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops)
    Rp, Ri, Rx = O.ref_sum(m, n, ops)
    assert list(Cp) == list(Rp) and list(Ci) == list(Ri) and bits_equal(Cx, Rx)
    if nops == 2:
        d1, d2 = rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)
        _, _, Sx, _ = O.sum_n(m, n, ops, d=np.stack([d1, d2]))
        _, _, Rs = O.ref_sum(m, n, ops, d=(d1, d2))
        assert bits_equal(Sx, Rs)


@needs_ref
def test_sum_reference_literal_through_reference_build():
    Rp, Ri, Rx = O.ref_sum(4, 6, SUM5_OPS)
    Cp, Ci, Cx, _ = O.sum_n(4, 6, SUM5_OPS)
    assert list(Rp) == list(Cp) and list(Ri) == list(Ci) and bits_equal(Rx, Cx)


# ------------------------------------------------------- dense helpers ----
def _kron_case(rng, m, k, n, integer=False):
    if integer:
        return (rng.integers(-4, 5, (k, n)).astype(float), rng.integers(-4, 5, (m, k)).astype(float),
                rng.integers(-4, 5, m * n).astype(float))
    return rng.standard_normal((k, n)), rng.standard_normal((m, k)), rng.standard_normal(m * n)


@pytest.mark.parametrize("shape", [(3, 2, 4), (7, 5, 1), (1, 1, 9), (40, 13, 37)])
def test_dense_kron_closed_form(shape):
    """(Y kron I_m) w and (I_n kron X^T) w against numpy's kron on column-major
    reshapes: exact on small integers, 1e-13 relative on normal values."""
    m, k, n = shape
    rng = np.random.default_rng(sum(shape))
    for integer in (True, False):
        Y, X, w = _kron_case(rng, m, k, n, integer)
        v1 = O.y_kron_i_vec(m, k, n, Y.reshape(-1, order="F"), w)
        r1 = np.kron(Y, np.eye(m)) @ w
        v2 = O.i_kron_xt_vec(m, k, n, X.reshape(-1, order="F"), w)
        r2 = np.kron(np.eye(n), X.T) @ w
        if integer:
            assert np.array_equal(v1, r1) and np.array_equal(v2, r2)
        else:
            assert np.allclose(v1, r1, rtol=1e-13, atol=1e-13 * np.abs(r1).max())
            assert np.allclose(v2, r2, rtol=1e-13, atol=1e-13 * np.abs(r2).max())


@needs_ref
@pytest.mark.parametrize("shape", [(3, 2, 4), (37, 19, 70), (64, 1, 5), (5, 3, 0)])
def test_dense_kron_matches_reference_build(shape):
    m, k, n = shape
    rng = np.random.default_rng(7 * m + k)
    Y, X, w = _kron_case(rng, m, k, n)
    r = O.ref_lib()
    Yf, Xf = Y.reshape(-1, order="F"), X.reshape(-1, order="F")
    assert bits_equal(O.y_kron_i_vec(m, k, n, Yf, w), O.y_kron_i_vec(m, k, n, Yf, w, r.Y_kron_I_vec))
    assert bits_equal(O.i_kron_xt_vec(m, k, n, Xf, w),
                      O.i_kron_xt_vec(m, k, n, Xf, w, r.I_kron_XT_vec))


def test_faithfulness_record():
    """The committed record of oracle/faithfulness.py: on the 20000 x 5000
    product the C restatement's structure and values were bit-identical to the
    reference build's, and its single-thread alloc and fill times were within
    +-20 % of the reference's (the restatement is the CPU baseline wherever the
    reference build is absent)."""
    rec = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "oracle_faithfulness.json")))
    assert rec["bit_identical"] is True and rec["nnz_out"] == 709850
    assert 0.8 <= rec["ratio_alloc"] <= 1.2 and 0.8 <= rec["ratio_fill"] <= 1.2
    assert rec["port"]["fill_runs"] >= 5 and rec["reference"]["fill_runs"] >= 5
    assert rec["host"]["nproc"] >= 1 and rec["host"]["cpu_model"]


@pytest.mark.parametrize("blocks", [1, 3])
def test_spgemm_range_equals_whole_product(blocks):
    """orc_spgemm_range (row blocks, for checks of products too large to
    form whole) gives the whole product's rows, bit for bit, and refuses a
    capacity below the range's term bound."""
    rng = np.random.default_rng(40 + blocks)
    Lm, nL, An = 37, 29, 41
    Lp, Li, Lx = random_csr(rng, Lm, nL, 0.2, empty_row=4)
    Ap, Ai, Ax = random_csr(rng, nL * blocks, An, 0.15, dense_col=7)
    d = rng.uniform(0.1, 10, nL * blocks)
    Cp, Ci, Cx, terms = O.spgemm(Lm, Lp, Li, Lx, blocks, d, nL * blocks, An, Ap, Ai, Ax)
    args = [np.ascontiguousarray(v, t) for v, t in ((Lp, np.int32), (Li, np.int32),
                                                    (Lx, np.float64))]
    Ac = [np.ascontiguousarray(v, t) for v, t in ((Ap, np.int32), (Ai, np.int32),
                                                  (Ax, np.float64))]
    cuts = [0, 5, 6, 20, Lm * blocks]
    for r0, r1 in zip(cuts, cuts[1:]):
        got = O.spgemm_range(r0, r1, Lm, *args, blocks, d, nL * blocks, An, *Ac, cap=terms)
        gp, gi, gx = got
        a, b = int(Cp[r0]), int(Cp[r1])
        assert np.array_equal(gp, np.asarray(Cp[r0:r1 + 1], np.int64) - a)
        assert np.array_equal(gi, Ci[a:b]) and bits_equal(gx, Cx[a:b])
    assert O.spgemm_range(0, Lm * blocks, Lm, *args, blocks, d, nL * blocks, An, *Ac,
                          cap=max(len(Ci) // 2, 1)) is None


def test_tracked_alloc_header_and_counters(tmp_path):
    """include/sparsediff_linalg/tracked_alloc.h as a C caller uses it
    (reference tracked_alloc.h:40-136): with -DSP_TRACK_MEMORY the inline
    sp_* wrappers and the library's own host allocations (a CSR struct's
    arrays) move the live / peak counters the library exports, freeing brings
    the live count back, sp_memory_reset_peak resets the mark; without the
    define the wrappers are libc's.  Host allocations only: no GPU call."""
    import subprocess
    from sparsediff_linalg_amd import LIB_PATH
    root = os.path.dirname(os.path.dirname(__file__))
    src = tmp_path / "t.c"
    src.write_text(r'''
#include <stdio.h>
#include "sparsediff_linalg/tracked_alloc.h"
#include "sparsediff_linalg/CSR_matrix.h"
int main(void)
{
#ifdef SP_TRACK_MEMORY
    size_t l0, p0, l1, p1, l2, p2, l3, p3;
    sp_memory_stats(&l0, &p0);
    char *a = (char *) sp_malloc(1000);
    sp_memory_stats(&l1, &p1);
    CSR_matrix *C = new_CSR_matrix(100, 50, 400);
    sp_memory_stats(&l2, &p2);
    free_CSR_matrix(C);
    a = (char *) sp_realloc(a, 5000);
    sp_free(a);
    sp_memory_stats(&l3, &p3);
    sp_memory_reset_peak();
    size_t l4, p4;
    sp_memory_stats(&l4, &p4);
    int ok = l1 >= l0 + 1000 && p1 >= l1 && l2 >= l1 + 101 * 4 + 400 * 12 && p2 >= l2
             && l3 == l0 && p3 >= l2 && p4 == l4 && l4 == l0 && sp_calloc(0, 0) != (void *) 1;
    printf("%s %zu %zu %zu %zu\n", ok ? "OK" : "BAD", l0, l1, l2, l3);
    return ok ? 0 : 1;
#else
    void *a = sp_calloc(10, 10);
    a = sp_realloc(a, 500);
    sp_free(a);
    printf("OK\n");
    return 0;
#endif
}
''')
    libdir = os.path.dirname(LIB_PATH)
    for define in (["-DSP_TRACK_MEMORY"], []):
        exe = tmp_path / ("t_track" if define else "t_plain")
        subprocess.run(["gcc", "-std=c99", "-O1", *define, "-I", os.path.join(root, "include"),
                        str(src), "-o", str(exe), "-L", libdir, "-l:libsdl_amd.so",
                        f"-Wl,-rpath,{libdir}"], check=True)
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout.startswith("OK"), (define, r.stdout, r.stderr)
    lib = C.CDLL(LIB_PATH, mode=os.RTLD_LOCAL)
    assert hasattr(lib, "g_allocated_bytes") and hasattr(lib, "g_peak_bytes")
