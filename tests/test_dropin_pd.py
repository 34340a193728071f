"""permuted_dense / stacked_pd kinds through the drop-in C API on the GPU:
every dispatcher kind pair of B^T D A, A^T D A and B A, DA, transpose,
row_gather / row_reduce, vstack of mixed kinds, and the Kronecker products.
Structure is checked exactly against the reference's rules (oracle/pd_oracle.py);
values against dense FP64 products within 1e-13 * (|B|^T |D| |A|) per entry
(the reference's CBLAS summation order is unspecified)."""
import ctypes as C
import json
import os
import zlib

import numpy as np
import pytest

from oracle import pd_oracle as PO
from tests import dropin as D
from tests.helpers import random_csr

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_literals.json")))


# ------------------------------------------------------------- builders ----
def rand_pd(rng, m, n, m0, n0, full=False):
    rp = list(range(m)) if full else sorted(rng.choice(m, m0, replace=False).tolist())
    cp = list(range(n)) if full else sorted(rng.choice(n, n0, replace=False).tolist())
    return PO.pd(m, n, rp, cp, rng.standard_normal((len(rp), len(cp))))


def rand_spd(rng, m, n, nb, max_rows=6, max_cols=7, n0_zero=False):
    rows = rng.permutation(m)[: min(m, nb * max_rows)]
    parts = np.array_split(rows, nb)
    bl = []
    for k, part in enumerate(parts):
        if len(part) == 0:
            continue
        n0 = 0 if (n0_zero and k == 1) else int(rng.integers(1, max_cols + 1))
        cp = sorted(rng.choice(n, min(n, n0), replace=False).tolist())
        rp = sorted(part.tolist())
        bl.append(PO.pd(m, n, rp, cp, rng.standard_normal((len(rp), len(cp)))))
    return PO.spd(m, n, bl)


def make(M):
    """GPU handle of an oracle PD / SPD."""
    if "blocks" not in M:
        return D.pd(M["m"], M["n"], M["rp"], M["cp"], M["X"])
    hs = [D.pd(b["m"], b["n"], b["rp"], b["cp"], b["X"]) for b in M["blocks"]]
    return D.spd(M["m"], M["n"], hs)


def rand_sparse(rng, m, n, dens):
    p, i, x = random_csr(rng, m, n, dens)
    return (p, i, x), D.sparse(p, i, x, m, n)


def dense_csr(p, i, x, m, n):
    out = np.zeros((m, n))
    for r in range(m):
        out[r, i[p[r]:p[r + 1]]] = x[p[r]:p[r + 1]]
    return out


def check(Cm, S, ref, bound):
    """Structure of handle Cm == oracle structure S; values within bound."""
    p, i, x = D.mat_csr(Cm)
    ep, ei = PO.csr_pattern(S)
    assert list(p) == ep, "row pointers differ"
    assert list(i) == ei, "column indices differ"
    got = dense_csr(p, i, x, S["m"], S["n"])
    mask = PO.dense(PO.with_values(S, np.ones((S["m"], S["n"])))) != 0
    err = np.abs(got - ref)[mask]
    assert np.all(err <= bound[mask]), f"max err {err.max() if err.size else 0}"


def kinds(rng, m, n, which):
    if which == "pd":
        M = rand_pd(rng, m, n, max(1, m // 2), max(1, n // 2))
        return M, make(M), PO.dense(M), None
    if which == "spd":
        M = rand_spd(rng, m, n, 4)
        return M, make(M), PO.dense(M), None
    (p, i, x), h = rand_sparse(rng, m, n, 0.25)
    return None, h, dense_csr(p, i, x, m, n), (p, i, n)


PAIRS = [("pd", "pd"), ("pd", "spd"), ("pd", "sparse"), ("spd", "pd"), ("spd", "spd"),
         ("spd", "sparse"), ("sparse", "pd"), ("sparse", "spd")]


@pytest.mark.parametrize("bk,ak", PAIRS)
def test_btda_kind_pairs(bk, ak):
    L = D.lib()
    rng = np.random.default_rng(zlib.crc32(f'{bk}{ak}'.encode()))
    m, nA, nB = 23, 13, 11
    Bo, Bh, Bd, Bcsr = kinds(rng, m, nB, bk)
    Ao, Ah, Ad, Acsr = kinds(rng, m, nA, ak)
    Cm = L.BTA_matrices_alloc(Ah, Bh)
    S = PO.bta(Ao if Ao else {"blocks": []}, Bo if Bo else None, A_csr=Acsr,
               B_csr=(Bcsr[0], Bcsr[1], nB) if Bcsr else None)
    kind_exp = 1 if bk == "pd" else 2 if bk == "spd" else (1 if ak == "pd" else 2)
    assert Cm.contents.kind == kind_exp
    d = rng.uniform(0.1, 10, m)
    v0 = Cm.contents.values_version
    L.BTDA_matrices_fill_values(Ah, D.dptr(d), Bh, Cm)
    assert Cm.contents.values_version == v0 + 1
    ref = Bd.T @ (d[:, None] * Ad)
    check(Cm, S, ref, PO.tol_bound(np.abs(Bd), d, np.abs(Ad)))
    L.BTA_matrices_fill_values(Ah, Bh, Cm)
    check(Cm, S, Bd.T @ Ad, PO.tol_bound(np.abs(Bd), np.ones(m), np.abs(Ad)))
    for M in (Cm, Ah, Bh):
        L.free_matrix(M)


@pytest.mark.parametrize("k", ["pd", "spd"])
def test_atda(k):
    L = D.lib()
    rng = np.random.default_rng(5 if k == "pd" else 6)
    m, n = 30, 17
    Ao, Ah, Ad, _ = kinds(rng, m, n, k)
    Cm = L.ATA_matrices_alloc(Ah)
    d = rng.uniform(0.1, 10, m)
    L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)
    check(Cm, PO.ata(Ao), Ad.T @ (d[:, None] * Ad), PO.tol_bound(np.abs(Ad), d, np.abs(Ad)))
    # values change, version bump, refill
    if k == "pd":
        D.set_x(Ah, D.mat_x(Ah) * 2.0)
        L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)
        Ad2 = Ad * 2.0
        check(Cm, PO.ata(Ao), Ad2.T @ (d[:, None] * Ad2), PO.tol_bound(np.abs(Ad2), d, np.abs(Ad2)))
    L.free_matrix(Cm)
    L.free_matrix(Ah)


@pytest.mark.parametrize("bk,ak", [("pd", "pd"), ("pd", "spd"), ("pd", "sparse"),
                                   ("sparse", "pd"), ("sparse", "spd")])
def test_ba_kind_pairs(bk, ak):
    L = D.lib()
    rng = np.random.default_rng(zlib.crc32(f'ba{bk}{ak}'.encode()))
    mB, k, nA = 9, 21, 12
    Bo, Bh, Bd, Bcsr = kinds(rng, mB, k, bk)
    Ao, Ah, Ad, Acsr = kinds(rng, k, nA, ak)
    Cm = L.BA_matrices_alloc(Bh, Ah)
    L.BA_matrices_fill_values(Bh, Ah, Cm)
    if bk == "pd":
        if ak == "pd":
            S = PO.ba_pd_pd(Bo, Ao)
        elif ak == "spd":
            S = PO.ba_pd_spd(Bo, Ao)
        else:
            S = PO.ba_pd_csr(Bo, Acsr[0], Acsr[1], nA)
    else:
        if ak == "pd":
            S = PO.ba_csc_pd(Bcsr[0], Bcsr[1], mB, Ao)
        else:
            raw = [PO.ba_csc_pd(Bcsr[0], Bcsr[1], mB, a) for a in PO.blocks(Ao)]
            S = PO.spd(mB, nA, PO.coalesce(mB, nA, raw))
    check(Cm, S, Bd @ Ad, 1e-13 * (np.abs(Bd) @ np.abs(Ad)) + 1e-300)
    for M in (Cm, Ah, Bh):
        L.free_matrix(M)


@pytest.mark.parametrize("k", ["pd", "spd"])
def test_da_transpose_gather_reduce(k):
    L = D.lib()
    rng = np.random.default_rng(11 if k == "pd" else 12)
    m, n = 20, 9
    Ao, Ah, Ad, _ = kinds(rng, m, n, k)
    # DA: one product per value
    Cm = L.copy_sparsity_matrices_alloc(Ah)
    d = rng.uniform(0.1, 10, m)
    allocs = L.sdl_device_alloc_count()
    L.DA_matrices_fill_values(D.dptr(d), Ah, Cm)
    assert L.sdl_device_alloc_count() == allocs
    p, i, x = D.mat_csr(Cm)
    assert (list(p), list(i)) == tuple(map(list, PO.csr_pattern(Ao)))
    ap, ai, ax = D.mat_csr(Ah)
    rows = np.repeat(np.arange(m), np.diff(ap))
    assert np.array_equal(x, ax * d[rows])
    L.free_matrix(Cm)
    # transpose (plain copies: bit-exact)
    T = L.transpose_matrices_alloc(Ah)
    L.transpose_matrices_fill_values(Ah, T)
    check(T, PO.transpose(Ao), Ad.T, np.zeros((n, m)))
    L.free_matrix(T)
    # row_gather (copies)
    mp = np.array([3, 0, 3, 7, 19, 11, 0], np.int32)
    G = Ah.contents.row_gather_alloc(Ah, D.iptr(mp), len(mp))
    G.contents.row_gather_fill_values(Ah, G)
    gp, gi, gx = D.mat_csr(G)
    assert np.array_equal(dense_csr(gp, gi, gx, len(mp), n), Ad[mp])
    L.free_matrix(G)
    # row_reduce (sums in source order)
    grp = rng.integers(0, 6, m).astype(np.int32)
    R = Ah.contents.row_reduce_alloc(Ah, D.iptr(grp), 6)
    R.contents.row_reduce_fill_values(Ah, R)
    rp, ri, rx = D.mat_csr(R)
    ref = np.zeros((6, n))
    for r in range(m):
        ref[grp[r]] += Ad[r]
    assert np.allclose(dense_csr(rp, ri, rx, 6, n), ref, rtol=1e-14, atol=1e-14)
    L.free_matrix(R)
    L.free_matrix(Ah)


def test_vstack_mixed_kinds_literal():
    L = D.lib()
    g = GOLD["vstack"][0]
    ops = []
    for o in g["ops"]:
        if o["kind"] == "sparse":
            ops.append(D.sparse(o["p"], o["i"], o["x"], o["m"], o["n"]))
        elif o["kind"] == "pd":
            ops.append(D.pd(o["m"], o["n"], o["row_perm"], o["col_perm"], o["X"]))
        else:
            bl = [D.pd(b["m"], b["n"], b["row_perm"], b["col_perm"], b["X"]) for b in o["blocks"]]
            ops.append(D.spd(o["m"], o["n"], bl))
    arr = D.ops_array(ops)
    Cm = L.vstack_matrices_alloc(arr, len(ops))
    L.vstack_matrices_fill_values(arr, len(ops), Cm)
    p, i, x = D.mat_csr(Cm)
    assert list(p) == g["C"]["p"] and list(i) == g["C"]["i"]
    assert list(x) == list(map(float, g["C"]["x"]))
    L.free_matrix(Cm)
    for M in ops:
        L.free_matrix(M)


@pytest.mark.parametrize("jk", ["sparse", "pd", "spd"])
@pytest.mark.parametrize("layout", ["I_A", "A_I"])
def test_dense_kron(jk, layout):
    L = D.lib()
    rng = np.random.default_rng(zlib.crc32(f'{jk}{layout}'.encode()))
    m, n, p, nJ = 3, 4, 3, 10
    A = rng.standard_normal((m, n))
    Ah = L.new_permuted_dense_full(m, n, D.dptr(A))
    Jo, Jh, Jd, _ = kinds(rng, n * p, nJ, jk)
    if layout == "I_A":
        Cm = L.BA_kron_matrices_alloc(Ah, p, Jh)
        L.BA_kron_matrices_fill_values(Ah, p, Jh, Cm)
        K = np.kron(np.eye(p), A)
    else:
        Cm = L.BA_dense_kron_I_matrices_alloc(Ah, p, Jh)
        L.BA_dense_kron_I_matrices_fill_values(Ah, p, Jh, Cm)
        K = np.kron(A, np.eye(p))
    assert Cm.contents.kind == 2
    ref = K @ Jd
    cp, ci, cx = D.mat_csr(Cm)
    got = dense_csr(cp, ci, cx, m * p, nJ)
    assert np.all(np.abs(got - ref) <= 1e-13 * (np.abs(K) @ np.abs(Jd)) + 1e-300)
    L.free_matrix(Cm)
    L.free_matrix(Jh)
    # kron_vec with the full PD
    x = rng.standard_normal(n * p)
    y = np.zeros(m * p)
    L.BA_kron_vec_matrices(Ah, D.dptr(x), y.ctypes.data_as(C.POINTER(C.c_double)), p)
    ref = (np.kron(np.eye(p), A) @ x)
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-14)
    L.free_matrix(Ah)


def test_stacked_blocks_scale_atda_spd():
    """256 blocks with sizes log-uniform 64..512 (the config-3 shape, smaller
    column space so blocks overlap): A^T D A of the stack against dense."""
    L = D.lib()
    rng = np.random.default_rng(3)
    nb, N = 24, 600
    sizes = np.exp(rng.uniform(np.log(64), np.log(256), (nb, 2))).astype(int)
    m = int(sizes[:, 0].sum())
    rows = rng.permutation(m)
    bl, start = [], 0
    for b in range(nb):
        rp = sorted(rows[start:start + sizes[b, 0]].tolist())
        start += sizes[b, 0]
        cp = sorted(rng.choice(N, sizes[b, 1], replace=False).tolist())
        bl.append(PO.pd(m, N, rp, cp, rng.standard_normal((len(rp), len(cp)))))
    Ao = PO.spd(m, N, bl)
    Ah = make(Ao)
    Ad = PO.dense(Ao)
    Cm = L.ATA_matrices_alloc(Ah)
    d = rng.uniform(0.1, 10, m)
    L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)
    check(Cm, PO.ata(Ao), Ad.T @ (d[:, None] * Ad), PO.tol_bound(np.abs(Ad), d, np.abs(Ad)))
    L.free_matrix(Cm)
    L.free_matrix(Ah)


# ------------------------------------------------- reference known answers ----
PD_CASES = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "pd_literals.json")))["cases"]


def fixture_handle(M):
    if M["kind"] == "sparse":
        return D.sparse(M["p"], M["i"], M["x"], M["m"], M["n"])
    if M["kind"] == "pd":
        return D.pd(M["m"], M["n"], M["row_perm"], M["col_perm"], M["X"])
    hs = [D.pd(b["m"], b["n"], b["row_perm"], b["col_perm"], b["X"]) for b in M["blocks"]]
    return D.spd(M["m"], M["n"], hs)


def expected_csr(C):
    """CSR of row-disjoint expected blocks (rows in order, each row's block
    columns ascending)."""
    rows = {}
    for b in C["blocks"]:
        n0 = len(b["col_perm"])
        for ii, r in enumerate(b["row_perm"]):
            rows[r] = (b["col_perm"], b["X"][ii * n0:(ii + 1) * n0])
    p, i, x = [0], [], []
    for r in range(C["m"]):
        if r in rows:
            i.extend(rows[r][0])
            x.extend(rows[r][1])
        p.append(len(i))
    return p, i, x


@pytest.mark.parametrize("c", PD_CASES, ids=[c["name"] for c in PD_CASES])
def test_pd_reference_literal(c):
    L = D.lib()
    A = fixture_handle(c["A"])
    B = fixture_handle(c["B"]) if "B" in c else None
    d = np.asarray(c["d"], float) if c.get("d") is not None else None
    op = c["op"]
    if op == "DA":
        Cm = L.copy_sparsity_matrices_alloc(A)
        L.DA_matrices_fill_values(D.dptr(d), A, Cm)
    elif op == "ATDA":
        Cm = L.ATA_matrices_alloc(A)
        L.ATDA_matrices_fill_values(A, D.dptr(d), Cm)
    elif op == "transpose":
        Cm = L.transpose_matrices_alloc(A)
        L.transpose_matrices_fill_values(A, Cm)
    elif op == "BTA":
        Cm = L.BTA_matrices_alloc(A, B)
        L.BTA_matrices_fill_values(A, B, Cm)
    elif op == "BTDA":
        Cm = L.BTA_matrices_alloc(A, B)
        L.BTDA_matrices_fill_values(A, D.dptr(d), B, Cm)
    else:
        Cm = L.BA_matrices_alloc(B, A)
        L.BA_matrices_fill_values(B, A, Cm)
    exp = c["C"]
    assert Cm.contents.kind == (1 if exp["kind"] == "pd" else 2)
    assert (Cm.contents.m, Cm.contents.n) == (exp["m"], exp["n"])
    p, i, x = D.mat_csr(Cm)
    ep, ei, ex = expected_csr(exp)
    assert list(p) == ep
    assert list(i) == ei
    assert list(x) == [float(v) for v in ex]
    for M in (Cm, A, B):
        if M is not None:
            L.free_matrix(M)


@pytest.mark.parametrize("case", ["atda", "btda"])
def test_stacked_blocks_full_size(case):
    """The full stacked permuted-dense workload (256 blocks, m0/n0 log-uniform
    64..512, col_perms in [0, 4096)): output structure exactly the oracle's,
    values on 2000 sampled cells within 1e-13 * sum |terms|."""
    from sparsediff_linalg_amd import workloads
    L = D.lib()
    w = workloads.stacked_pd_blocks(3)
    m, N, d = w["m"], w["n"], w["d"]
    Ao = PO.spd(m, N, [PO.pd(m, N, rp.tolist(), cp.tolist(), X) for rp, cp, X in w["A"]])
    Ah = make(Ao)
    if case == "btda":
        Bo = PO.spd(m, N, [PO.pd(m, N, rp.tolist(), cp.tolist(), X) for rp, cp, X in w["B"]])
        Bh = make(Bo)
        Cm = L.BTA_matrices_alloc(Ah, Bh)
        L.BTDA_matrices_fill_values(Ah, D.dptr(d), Bh, Cm)
        S = PO.bta(Ao, Bo)
        Lb = w["B"]
    else:
        Bh = None
        Cm = L.ATA_matrices_alloc(Ah)
        L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)
        S = PO.ata(Ao)
        Lb = w["A"]
    p, i, x = D.mat_csr(Cm)
    ep, ei = PO.csr_pattern(S)
    assert np.array_equal(p, np.asarray(ep)), "row pointers differ"
    assert np.array_equal(i, np.asarray(ei)), "column indices differ"
    rows_of = np.repeat(np.arange(len(p) - 1), np.diff(p))
    rng = np.random.default_rng(1)
    invs = []
    for (_, cpL, _), (_, cpA, _) in zip(Lb, w["A"]):
        iL = np.full(N, -1)
        iL[cpL] = np.arange(len(cpL))
        iA = np.full(N, -1)
        iA[cpA] = np.arange(len(cpA))
        invs.append((iL, iA))
    for q in rng.choice(len(x), 2000, replace=False):
        r, c = rows_of[q], i[q]
        ref = mag = 0.0
        for k, ((_, _, XL), (rpA, _, XA)) in enumerate(zip(Lb, w["A"])):
            iL, iA = invs[k]
            if iL[r] >= 0 and iA[c] >= 0:
                t = XL[:, iL[r]] * d[rpA] * XA[:, iA[c]]
                ref += t.sum()
                mag += np.abs(t).sum()
        assert abs(x[q] - ref) <= 1e-13 * mag + 1e-300, (r, c, x[q], ref)
    for M in (Cm, Ah, Bh):
        if M is not None:
            L.free_matrix(M)


def test_spd_disjoint_columns_single_source_outputs():
    """A^T D A of a stack whose blocks have disjoint col_perms: every output
    cell has one source, so the ordered coalesce copies runs of consecutive
    cells (0 + x, as the ordered sum).  Structure exactly the oracle's; every
    value against the dense product within 1e-13 * sum |terms|."""
    L = D.lib()
    rng = np.random.default_rng(12)
    nb = 24
    n0s = rng.integers(64, 257, nb)
    m0s = rng.integers(32, 129, nb)
    N = int(n0s.sum()) + 500
    m = int(m0s.sum())
    cols = rng.permutation(N)
    rows = rng.permutation(m)
    blocks, c0, r0 = [], 0, 0
    for k in range(nb):
        cp = sorted(cols[c0:c0 + n0s[k]].tolist())
        rp = sorted(rows[r0:r0 + m0s[k]].tolist())
        c0 += int(n0s[k])
        r0 += int(m0s[k])
        blocks.append(PO.pd(m, N, rp, cp, rng.standard_normal((len(rp), len(cp)))))
    Ao = PO.spd(m, N, blocks)
    d = rng.uniform(0.1, 10.0, m)
    Ah = make(Ao)
    Cm = L.ATA_matrices_alloc(Ah)
    L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)
    p, i, x = D.mat_csr(Cm)
    ep, ei = PO.csr_pattern(PO.ata(Ao))
    assert np.array_equal(p, np.asarray(ep)) and np.array_equal(i, np.asarray(ei))
    Ad = np.zeros((m, N))
    for b in blocks:
        Ad[np.ix_(b["rp"], b["cp"])] = b["X"]
    ref = Ad.T @ (d[:, None] * Ad)
    mag = np.abs(Ad).T @ (d[:, None] * np.abs(Ad))
    r = np.repeat(np.arange(N), np.diff(p))
    assert np.all(np.abs(x - ref[r, i]) <= 1e-13 * mag[r, i] + 1e-300)
    for M in (Cm, Ah):
        L.free_matrix(M)


def test_block_refill_split_equals_refill():
    """sdl_block_refill_device_split (events around the block products and
    the whole refill) gives the same values as sdl_block_refill_device, and
    its events are ordered (products end before the refill does)."""
    import ctypes as C
    lib = D.lib()
    rng = np.random.default_rng(8)
    nb, N = 12, 400
    sizes = np.exp(rng.uniform(np.log(64), np.log(200), (nb, 2))).astype(int)
    m = int(sizes[:, 0].sum())
    rows = rng.permutation(m)
    bl, start = [], 0
    for b in range(nb):
        rp = sorted(rows[start:start + sizes[b, 0]].tolist())
        start += sizes[b, 0]
        cp = sorted(rng.choice(N, sizes[b, 1], replace=False).tolist())
        bl.append(PO.pd(m, N, rp, cp, rng.standard_normal((len(rp), len(cp)))))
    Ah = make(PO.spd(m, N, bl))
    Cm = lib.ATA_matrices_alloc(Ah)
    d = rng.uniform(0.1, 10, m)
    lib.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm)  # binds the operand mirrors
    from sparsediff_linalg_amd import DeviceArray
    dd = DeviceArray.from_numpy(d)
    nval = int(Cm.contents.nnz)
    assert lib.sdl_block_refill_device(Cm, Ah, Ah, dd.ptr, None, None) == 0
    lib.sdl_synchronize()
    first = np.empty(nval)
    lib.sdl_memcpy_dtoh(first.ctypes.data, lib.sdl_block_device_values(Cm), first.nbytes)
    e0, eg, e1 = lib.sdl_event_create(), lib.sdl_event_create(), lib.sdl_event_create()
    assert lib.sdl_block_refill_device_split(Cm, Ah, Ah, dd.ptr, e0, eg, e1) == 0
    lib.sdl_synchronize()
    second = np.empty(nval)
    lib.sdl_memcpy_dtoh(second.ctypes.data, lib.sdl_block_device_values(Cm), second.nbytes)
    assert np.array_equal(first.view(np.int64), second.view(np.int64))
    t_products, t_all = lib.sdl_event_elapsed_ms(e0, eg), lib.sdl_event_elapsed_ms(e0, e1)
    assert 0.0 <= t_products <= t_all
    lib.free_matrix(Cm)
    lib.free_matrix(Ah)
