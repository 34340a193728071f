"""GPU: the row-accumulator refills, whole products bit for bit against the
oracle: output rows longer than 1024 entries, and short rows built from A
rows longer than a workgroup.
Rows of up to 65536 entries with n <= 524,288 columns go by rank (column
bitmap + ranks in LDS); rows whose multi-term list overflows, and every long
row when n is larger, take the dense per-workgroup accumulator."""
import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan

pytestmark = pytest.mark.gpu


def hub_csr(seed, m, n, length, hubs, hubw):
    """m x n: the first `hubs` rows hold columns 0 .. hubw-1 (every output row
    below hubw then sums many terms per entry: long multi-term lists), the
    others `length` distinct uniform columns."""
    rng = np.random.default_rng(seed)
    rows = [np.arange(hubw) if i < hubs else np.sort(rng.choice(n, length, replace=False))
            for i in range(m)]
    p = np.zeros(m + 1, np.int64)
    p[1:] = np.cumsum([len(r) for r in rows])
    i = np.concatenate(rows).astype(np.int32)
    return p.astype(np.int32), i, rng.standard_normal(len(i))


def refill_and_compare(m, n, Ap, Ai, Ax, seed):
    lib = load()
    pl = Plan((Ap, Ai), None, n, n)
    nnz = pl.nnz()
    assert not lib.sdl_btda_plan_is_pull(pl.h)
    d = W.uniform(seed, m, 0.1, 10.0)
    dAx, dd = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(d)
    Cx = DeviceArray(max(nnz, 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr, Cx.ptr, None))
    lib.sdl_synchronize()
    cp = np.empty(n + 1, np.int64)
    ci = np.empty(max(nnz, 1), np.int32)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), cp.nbytes)
    lib.sdl_memcpy_dtoh(ci.ctypes.data, lib.sdl_btda_plan_colind(pl.h), ci.nbytes)
    Cp, Ci, Cv, _ = O.btda(m, n, n, Ap, Ai, Ax, Ap, Ai, Ax, d)
    assert np.array_equal(cp, np.asarray(Cp, np.int64))
    assert np.array_equal(ci[:nnz], Ci)
    assert np.array_equal(Cx.numpy()[:nnz].view(np.int64), np.asarray(Cv).view(np.int64))
    pl.close()
    return np.diff(cp)


def test_rank_rows_with_spills_match_oracle():
    # 4000 output rows, 88 % with 1025..3100 entries (rank bin); the 2000 below the hub
    # width take ~60k terms each, so their lists overflow and spill
    m, n = 600, 4000
    Ap, Ai, Ax = hub_csr(3, m, n, 100, 30, 2000)
    lens = refill_and_compare(m, n, Ap, Ai, Ax, 31)
    assert (lens > 1024).mean() > 0.85


def test_rank_rows_many_l_chunks_match_oracle():
    # hub columns appear in more than 1024 rows: several term chunks per row
    m, n = 1500, 3000
    Ap, Ai, Ax = hub_csr(4, m, n, 50, 1100, 1200)
    lens = refill_and_compare(m, n, Ap, Ai, Ax, 32)
    assert lens.max() > 1024


def test_dense_rows_beyond_rank_width_match_oracle():
    # n = 600,000 > 524,288 columns: long rows take the dense accumulators
    # (12 rows over the same 1500 columns: 1500 output rows of 1500 entries,
    # 12 terms each)
    m, n = 12, 600_000
    Ap, Ai, Ax = hub_csr(5, m, n, 0, 12, 1500)
    lens = refill_and_compare(m, n, Ap, Ai, Ax, 33)
    assert (lens > 1024).sum() == 1500


def test_global_bitmap_symbolic_matches_oracle():
    # n = 1,300,000 columns: a column bitmap no longer fits LDS, and 8 rows over
    # the same 5000 columns give every output row 40,000 terms, beyond the
    # largest hash table, so the symbolic passes use bitmaps in global memory
    m, n = 8, 1_300_000
    Ap, Ai, Ax = hub_csr(6, m, n, 0, 8, 5000)
    lens = refill_and_compare(m, n, Ap, Ai, Ax, 34)
    assert (lens == 5000).sum() == 5000


def test_short_rows_with_long_a_rows_match_oracle():
    """Output rows of at most 1000 entries (the LDS row accumulator) built from
    A rows of 300 entries, longer than the workgroup: each k's first 256
    terms come from the pipelined loads, the rest from the tail loop."""
    m, n = 400, 1000
    Ap, Ai, Ax = hub_csr(8, m, n, 300, 0, 0)
    lens = refill_and_compare(m, n, Ap, Ai, Ax, 81)
    assert lens.max() <= 1024 and lens.min() > 256
