"""GPU: the device-resident lower-triangular COO hand-off (sdl_lower_coo_*)
against the oracle's export (COO_matrix.c:48-98): rows, cols, value_map and
refreshed values of a product plan's output, bit for bit, with the refresh
repeated after a refill with new d and captured in a HIP graph."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan
from tests.helpers import bits_equal, random_csr

pytestmark = pytest.mark.gpu
vp = C.c_void_p


def plan_output(lib, pl, n):
    cp = np.empty(n + 1, np.int64)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), cp.nbytes)
    nnz = int(cp[-1])
    ci = np.empty(nnz, np.int32)
    lib.sdl_memcpy_dtoh(ci.ctypes.data, lib.sdl_btda_plan_colind(pl.h), ci.nbytes)
    return cp, ci


def read(lib, ptr, n, dt):
    a = np.empty(n, dt)
    if n:
        lib.sdl_memcpy_dtoh(a.ctypes.data, ptr, a.nbytes)
    return a


@pytest.mark.parametrize("case", ["ata", "banded", "empty-rows"])
def test_lower_coo_of_product(case):
    lib = load()
    rng = np.random.default_rng(5)
    if case == "ata":
        m, n = 3000, 900
        Ap, Ai, Ax = random_csr(rng, m, n, 0.01)
    elif case == "banded":
        n = m = 4096
        Ap, Ai, Ax = W.banded_csr(m, 9)
    else:  # columns 0..199 of A empty: rows and columns of C empty
        m, n = 2000, 600
        Ap, Ai, Ax = random_csr(rng, m, n, 0.01)
        keep = Ai >= 200
        rows = np.repeat(np.arange(m), np.diff(Ap))[keep]
        Ai, Ax = Ai[keep], Ax[keep]
        Ap = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int32)
    pl = Plan((Ap, Ai), None, n, n)  # A^T A: B = A
    cp, ci = plan_output(lib, pl, n)
    dA = DeviceArray.from_numpy(np.ascontiguousarray(Ax))
    nnz = int(cp[-1])
    Cx = DeviceArray(max(nnz, 1), np.float64)
    h = vp()
    check(lib.sdl_lower_coo_create(n, lib.sdl_btda_plan_rowptr(pl.h),
                                   lib.sdl_btda_plan_colind(pl.h), C.byref(h)), "lower coo")
    cnt = lib.sdl_lower_coo_nnz(h)
    x = DeviceArray(max(cnt, 1), np.float64)
    stream = lib.sdl_library_stream()
    for seed in (1, 2):
        d = W.uniform(seed, m, 0.1, 10.0)
        dd = DeviceArray.from_numpy(d)
        check(lib.sdl_btda_refill(pl.h, dA.ptr, dA.ptr, dd.ptr, Cx.ptr, None), "refill")
        check(lib.sdl_lower_coo_refresh(h, Cx.ptr, x.ptr, stream), "refresh")
        lib.sdl_synchronize()
        cx = read(lib, Cx.ptr, nnz, np.float64)
        r, c, v, vm = O.coo_lower(n, cp.astype(np.int32), ci, cx)
        assert cnt == len(r)
        assert np.array_equal(read(lib, lib.sdl_lower_coo_rows(h), cnt, np.int32), r)
        assert np.array_equal(read(lib, lib.sdl_lower_coo_cols(h), cnt, np.int32), c)
        assert np.array_equal(read(lib, lib.sdl_lower_coo_value_map(h), cnt, np.int64), vm)
        assert bits_equal(read(lib, x.ptr, cnt, np.float64), v)
    # the refresh captured once and replayed after each new refill
    check(lib.sdl_graph_begin(stream), "capture")
    rc = lib.sdl_lower_coo_refresh(h, Cx.ptr, x.ptr, stream)
    g = vp()
    check(lib.sdl_graph_end(stream, C.byref(g)), "end capture")
    check(rc, "captured refresh")
    for seed in (3, 4):
        dd = DeviceArray.from_numpy(W.uniform(seed, m, 0.1, 10.0))
        check(lib.sdl_btda_refill(pl.h, dA.ptr, dA.ptr, dd.ptr, Cx.ptr, stream), "refill")
        lib.sdl_memset_d(x.ptr, 0xFF, max(cnt, 1) * 8)
        check(lib.sdl_graph_launch(g, stream), "replay")
        lib.sdl_synchronize()
        cx = read(lib, Cx.ptr, nnz, np.float64)
        assert bits_equal(read(lib, x.ptr, cnt, np.float64),
                          O.coo_lower(n, cp.astype(np.int32), ci, cx)[2]), seed
    lib.sdl_graph_destroy(g)
    lib.sdl_lower_coo_destroy(h)
    pl.close()


def test_lower_coo_zero_rows():
    lib = load()
    dp = DeviceArray.from_numpy(np.zeros(1, np.int64))
    h = vp()
    check(lib.sdl_lower_coo_create(0, dp.ptr, None, C.byref(h)), "lower coo")
    assert lib.sdl_lower_coo_nnz(h) == 0
    check(lib.sdl_lower_coo_refresh(h, None, None, None), "refresh")
    lib.sdl_lower_coo_destroy(h)
