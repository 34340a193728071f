"""GPU: parity with the oracle at the benchmark workloads' full sizes.

* 4096 B^T D A instances (m log-uniform in [100, 50000], n = m/4, 4-12
  nnz/row): every instance's row lengths against the oracle's symbolic
  counts, and 128 instances (the largest and the smallest among them) whole,
  structure and values, after one refill round of the one-plan-per-group
  product.
* R-MAT scale 18 (a, b, c = 0.57, 0.19, 0.19; 16 edges per node) A^T A:
  every row of the exact plan and of the predicted plan, structure and
  values, against the oracle (nnz(C) = 1.27e9 < INT_MAX; scale 20 passes
  INT_MAX, so scale 18 is the whole-product case).
* Banded A^T A, 2^20 rows, bandwidth 64: 10,000+ rows (100 blocks of 100
  plus both ends) of the exact and the predicted plan.
* 50M COO triplets (35M distinct keys in 1M x 1M): the whole coalesced CSR.
* A^T A of the 2,000,000 x 500,000 rank-Zipf matrix (nnz(C) beyond INT_MAX):
  10,000+ rows (100 blocks of 100 plus the longest, the shortest and both
  ends); and of the 200,000 x 50,000 matrix of the same law, every row.
All bit for bit: structure exact, every value the oracle's fold."""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.multi import MultiProduct
from sparsediff_linalg_amd.predictor import Plan, Predictor
from tests.helpers import bits_equal, check_rows_parallel, row_blocks, row_term_bounds

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[full-size] {msg}", flush=True)


def dtoh(lib, dptr, count, dtype):
    out = np.empty(count, dtype)
    if count:
        check(lib.sdl_memcpy_dtoh(out.ctypes.data, dptr, out.nbytes), "dtoh")
    return out


def plan_fetcher(lib, plans):
    """fetch(r0, r1) over device plans [(handle, device Cx pointer)]."""
    def fetch(r0, r1):
        out = []
        for h, cx in plans:
            rp = lib.sdl_btda_plan_rowptr(h)
            cp = dtoh(lib, rp + 8 * r0, r1 - r0 + 1, np.int64)
            a, b = int(cp[0]), int(cp[-1])
            ci = dtoh(lib, lib.sdl_btda_plan_colind(h) + 4 * a, b - a, np.int32)
            x = dtoh(lib, cx + 8 * a, b - a, np.float64)
            out.append((cp - a, ci, x))
        return out
    return fetch


def ata_check(lib, Ap, Ai, Ax, n, d, plans, ranges):
    """Rows `ranges` of A^T diag(d) A from every device plan vs the oracle."""
    m = len(Ap) - 1
    Lp, Li, Lx = O.transpose(m, n, Ap, Ai, Ax)
    return check_rows_parallel(plan_fetcher(lib, plans), n, Lp, Li, Lx, 1, d, m, n, Ap, Ai, Ax,
                               ranges, threads=THREADS)


def refill(lib, plan, dAx, dd):
    Cx = DeviceArray(max(plan.nnz(), 1), np.float64)
    check(lib.sdl_btda_refill(plan.h, dAx.ptr, dAx.ptr, dd.ptr, Cx.ptr, None), "refill")
    return Cx


def sampled_blocks(n, rng, nblocks=100, size=100, extra=()):
    """nblocks random blocks of `size` rows, both ends, and single extra rows."""
    starts = rng.choice(max(n - size, 1), nblocks, replace=False)
    rr = {(int(s), int(min(n, s + size))) for s in starts}
    rr |= {(0, min(n, size)), (max(0, n - size), n)}
    rr |= {(int(r), int(r) + 1) for r in extra}
    return sorted(rr)


# ------------------------------------------------------ many instances ----
def test_multi_4096_instances():
    lib = load()
    say("generating 4096 instances")
    w = W.multi_instances(4)
    mp = MultiProduct(w, K=2)
    try:
        M = mp.M
        d = W.uniform(4242, M, 0.1, 10.0)
        D = DeviceArray.from_numpy(d)
        mp.refill_round(D, 0)
        lib.sdl_synchronize()
        ninst = len(w["m"])
        say(f"{len(mp.g)} groups, {mp.nnz} entries, {mp.terms} terms; row lengths")
        # every instance's output row lengths vs the oracle's symbolic count
        dev_len = {}
        for g in mp.g:
            cp = dtoh(lib, lib.sdl_btda_plan_rowptr(g.h), g.n + 1, np.int64)
            for k in range(g.i0, g.i1):
                c0, c1 = int(w["coff"][k]) - g.c0, int(w["coff"][k + 1]) - g.c0
                dev_len[k] = np.diff(cp[c0:c1 + 1])

        def operands(k):
            r0, r1 = int(w["roff"][k]), int(w["roff"][k + 1])
            c0 = int(w["coff"][k])
            ops = []
            for tag in ("A", "B"):
                p, i, x = w[tag]
                a, b = int(p[r0]), int(p[r1])
                ops.append(((p[r0:r1 + 1] - a).astype(np.int32),
                            (i[a:b] - c0).astype(np.int32), x[a:b]))
            return r1 - r0, int(w["n"][k]), ops

        def sym_len(k):
            m, n, ((Ap, Ai, _), (Bp, Bi, _)) = operands(k)
            Lp, Li, _ = O.transpose(m, n, Bp, Bi, None)
            Cp = np.zeros(n + 1, np.int64)
            O.lib().orc_spgemm_symbolic(n, O.I(np.ascontiguousarray(Lp, np.int32)),
                                        O.I(np.ascontiguousarray(Li, np.int32)), 1, m, n,
                                        O.I(Ap), O.I(Ai), O.L(Cp), None)
            return k, np.diff(Cp)

        with ThreadPoolExecutor(THREADS) as pool:
            for k, ref in pool.map(sym_len, range(ninst)):
                assert np.array_equal(dev_len[k], ref), k
        # 128 instances whole: the largest, the smallest and 126 more
        rng = np.random.default_rng(128)
        pick = {int(np.argmax(w["m"])), int(np.argmin(w["m"]))}
        pick |= set(int(v) for v in rng.choice(ninst, 126, replace=False))
        pick = sorted(pick)[:128] if len(pick) > 128 else sorted(pick)
        say(f"{len(pick)} instances whole")
        gof = {}
        for gi, g in enumerate(mp.g):
            for k in range(g.i0, g.i1):
                gof[k] = gi
        for k in pick:
            g = mp.g[gof[k]]
            m, n, ((Ap, Ai, Ax), (Bp, Bi, Bx)) = operands(k)
            r0 = int(w["roff"][k])
            Cp, Ci, Cx, _ = O.btda(m, n, n, Ap, Ai, Ax, Bp, Bi, Bx, d[r0:r0 + m])
            e0, e1 = g.ranges[k - g.i0]
            ci = dtoh(lib, lib.sdl_btda_plan_colind(g.h) + 4 * e0, e1 - e0, np.int32)
            cx = dtoh(lib, g.C.ptr + 8 * e0, e1 - e0, np.float64)
            assert e1 - e0 == len(Ci), k
            assert np.array_equal(ci - (int(w["coff"][k]) - g.c0), Ci), k
            assert bits_equal(cx, Cx), k
    finally:
        mp.close()


# ------------------------------------------------------- R-MAT / banded ----
def test_rmat18_ata_whole_exact_and_predicted():
    lib = load()
    say("R-MAT scale 18")
    Ap, Ai, Ax = W.rmat_csr(18)
    n = 1 << 18
    exact = Plan((Ap, Ai), None, n, n)
    pred = Predictor.from_file()
    predicted = Plan((Ap, Ai), None, n, n, mode=1, mlp=pred)
    try:
        assert exact.nnz() == predicted.nnz() == 1_274_921_546
        d = W.uniform(518, n, 0.1, 10.0)
        dAx, dd = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(d)
        Cx_e, Cx_p = refill(lib, exact, dAx, dd), refill(lib, predicted, dAx, dd)
        lib.sdl_synchronize()
        stats = predicted.pred_stats()
        say(f"predicted-plan stats {stats}; checking every row")
        Lp, Li, _ = O.transpose(n, n, Ap, Ai, None)
        ranges = row_blocks(n, row_term_bounds(n, Lp, Li, 1, n, Ap), 4_000_000)
        cnt = ata_check(lib, Ap, Ai, Ax, n, d,
                        [(exact.h, Cx_e.ptr), (predicted.h, Cx_p.ptr)], ranges)
        assert cnt == 1_274_921_546
    finally:
        exact.close()
        predicted.close()
        pred.close()


def test_banded64_2e20_sampled_exact_and_predicted():
    lib = load()
    n = 1 << 20
    say("banded 2^20, bandwidth 64")
    Ap, Ai, Ax = W.banded_csr(n, 64)
    exact = Plan((Ap, Ai), None, n, n)
    pred = Predictor.from_file()
    predicted = Plan((Ap, Ai), None, n, n, mode=1, mlp=pred)
    try:
        assert exact.nnz() == predicted.nnz()
        d = W.uniform(64, n, 0.1, 10.0)
        dAx, dd = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(d)
        Cx_e, Cx_p = refill(lib, exact, dAx, dd), refill(lib, predicted, dAx, dd)
        lib.sdl_synchronize()
        ranges = sampled_blocks(n, np.random.default_rng(64))
        cnt = ata_check(lib, Ap, Ai, Ax, n, d,
                        [(exact.h, Cx_e.ptr), (predicted.h, Cx_p.ptr)], ranges)
        rows = sum(b - a for a, b in ranges)
        say(f"{rows} rows, {cnt} entries")
        assert rows >= 10_000
    finally:
        exact.close()
        predicted.close()
        pred.close()


# ------------------------------------------------------------ config 2 ----
def test_coo_50m_triplets_whole():
    lib = load()
    m = n = 1_000_000
    say("50M COO triplets")
    rows, cols, x = W.coo_triplets(2, m=m, n=n)
    dr, dc, dx = (DeviceArray.from_numpy(a) for a in (rows, cols, x))
    h = C.c_void_p()
    check(lib.sdl_coo_plan_create(len(rows), dr.ptr, dc.ptr, m, n, C.byref(h)))
    try:
        nnz = lib.sdl_coo_plan_nnz(h)
        out = DeviceArray(nnz, np.float64)
        check(lib.sdl_coo_fill(h, dx.ptr, out.ptr, None))
        lib.sdl_synchronize()
        Cp, Ci, Cx = O.coo_coalesce(rows, cols, x, m, n)
        assert nnz == len(Ci) == 35_000_000
        assert np.array_equal(dtoh(lib, lib.sdl_coo_plan_rowptr(h), m + 1, np.int64), Cp)
        assert np.array_equal(dtoh(lib, lib.sdl_coo_plan_colind(h), nnz, np.int32), Ci)
        assert bits_equal(out.numpy(), Cx)
    finally:
        lib.sdl_coo_plan_destroy(h)


def test_ata_2m_x_500k_10k_rows():
    lib = load()
    m, n = 2_000_000, 500_000
    say("A^T A, A 2,000,000 x 500,000")
    Ap, Ai, Ax = W.zipf_csr(2, m, n)
    pl = Plan((Ap, Ai), None, n, n)
    try:
        nnz = pl.nnz()
        assert nnz > 2 ** 31
        d = W.uniform(77, m, 0.1, 10.0)
        dAx, dd = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(d)
        Cx = refill(lib, pl, dAx, dd)
        lib.sdl_synchronize()
        cp = dtoh(lib, lib.sdl_btda_plan_rowptr(pl.h), n + 1, np.int64)
        lens = np.diff(cp)
        ranges = sampled_blocks(n, np.random.default_rng(5),
                                extra=(int(np.argmax(lens)), int(np.argmin(lens))))
        cnt = ata_check(lib, Ap, Ai, Ax, n, d, [(pl.h, Cx.ptr)], ranges)
        rows = sum(b - a for a, b in ranges)
        say(f"{rows} rows, {cnt} entries")
        assert rows >= 10_000
    finally:
        pl.close()


def test_ata_200k_x_50k_whole():
    lib = load()
    m, n = 200_000, 50_000
    say("A^T A, A 200,000 x 50,000")
    Ap, Ai, Ax = W.zipf_csr(21, m, n)
    pl = Plan((Ap, Ai), None, n, n)
    try:
        d = W.uniform(78, m, 0.1, 10.0)
        dAx, dd = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(d)
        Cx = refill(lib, pl, dAx, dd)
        lib.sdl_synchronize()
        Lp, Li, _ = O.transpose(m, n, Ap, Ai, None)
        ranges = row_blocks(n, row_term_bounds(n, Lp, Li, 1, m, Ap), 4_000_000)
        cnt = ata_check(lib, Ap, Ai, Ax, n, d, [(pl.h, Cx.ptr)], ranges)
        assert cnt == pl.nnz()
        say(f"{cnt} entries")
    finally:
        pl.close()
