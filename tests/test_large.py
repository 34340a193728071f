"""GPU: the large single product and coalescing.  The 2,000,000 x 500,000
A^T A (rank-Zipf(1.2) row lengths in [1, 512], mean 16; nnz(C) beyond INT_MAX,
so the int64 device plan) checked on sampled rows against the oracle, bit for
bit; device COO coalescing checked whole against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan

pytestmark = pytest.mark.gpu


def test_coo_coalesce_device_matches_oracle():
    lib = load()
    m = n = 200_000
    rows, cols, x = W.coo_triplets(9, n_items=3_000_000, unique=2_100_000, m=m, n=n)
    dr, dc, dx = (DeviceArray.from_numpy(a) for a in (rows, cols, x))
    h = C.c_void_p()
    check(lib.sdl_coo_plan_create(len(rows), dr.ptr, dc.ptr, m, n, C.byref(h)))
    nnz = lib.sdl_coo_plan_nnz(h)
    out = DeviceArray(nnz, np.float64)
    check(lib.sdl_coo_fill(h, dx.ptr, out.ptr, None))
    lib.sdl_synchronize()
    cp = np.empty(m + 1, np.int64)
    ci = np.empty(nnz, np.int32)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_coo_plan_rowptr(h), cp.nbytes)
    lib.sdl_memcpy_dtoh(ci.ctypes.data, lib.sdl_coo_plan_colind(h), ci.nbytes)
    Cp, Ci, Cx = O.coo_coalesce(rows, cols, x, m, n)
    assert nnz == len(Ci) == 2_100_000
    assert np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
    assert np.array_equal(out.numpy().view(np.int64), Cx.view(np.int64))
    lib.sdl_coo_plan_destroy(h)


def test_ata_large_sampled_rows():
    lib = load()
    m, n = 2_000_000, 500_000
    Ap, Ai, Ax = W.zipf_csr(2, m, n)
    pl = Plan((Ap, Ai), None, n, n)  # A^T A: B = A
    nnz = pl.nnz()
    assert nnz > 2 ** 31  # beyond the int32 ABI: device plan only
    assert not lib.sdl_btda_plan_is_pull(pl.h)  # terms beyond int32: row accumulators
    dAx = DeviceArray.from_numpy(Ax)
    d = W.uniform(77, m, 0.1, 10.0)
    dd = DeviceArray.from_numpy(d)
    Cx = DeviceArray(nnz, np.float64)
    check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr, Cx.ptr, None))
    lib.sdl_synchronize()
    cp = np.empty(n + 1, np.int64)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), cp.nbytes)
    assert cp[-1] == nnz
    Lp, Li, Lx = O.transpose(m, n, Ap, Ai, Ax)
    rng = np.random.default_rng(5)
    rows = np.concatenate([rng.choice(n, 44, replace=False),
                           [0, n - 1, int(np.argmax(np.diff(cp))), int(np.argmin(np.diff(cp)))]])
    ci_dev = lib.sdl_btda_plan_colind(pl.h)
    for r, (cols, vals) in zip(rows, O.spgemm_rows(rows, n, Lp, Li, Lx, 1, d, m, n, Ap, Ai, Ax)):
        a, b = int(cp[r]), int(cp[r + 1])
        gi = np.empty(b - a, np.int32)
        gx = np.empty(b - a, np.float64)
        if b > a:
            lib.sdl_memcpy_dtoh(gi.ctypes.data, ci_dev + 4 * a, gi.nbytes)
            lib.sdl_memcpy_dtoh(gx.ctypes.data, Cx.offset(a), gx.nbytes)
        assert np.array_equal(gi, cols), r
        assert np.array_equal(gx.view(np.int64), vals.view(np.int64)), r
    pl.close()
