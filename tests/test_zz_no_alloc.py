"""GPU: sdl_btda_refill_batched_static, the batched refill that allocates
nothing (its transposed d lives in the library's fixed device array), against
sdl_btda_refill_batched on the same plan and inputs, bit for bit, on two
calls in a row (the second on a smaller product and a smaller K, so it runs
over what the first left in the array), with no device allocation during
either call; plus the size limit's error and a check of one refill of each
call against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads
from tests.helpers import bits_equal, random_csr

pytestmark = pytest.mark.gpu


def _plan(lib, m, nA, nB, Ap, Ai, Bp, Bi):
    arrs = [DeviceArray.from_numpy(v) for v in (Ap, Ai, Bp, Bi)]
    h = C.c_void_p()
    check(lib.sdl_btda_plan_create(m, nA, nB, arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, arrs[3].ptr,
                                   0, C.byref(h)), "plan")
    return h, arrs


def _case(lib, m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, K, seed):
    h, keep = _plan(lib, m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        assert lib.sdl_btda_plan_is_pull(h) == 1
        nnz = lib.sdl_btda_plan_nnz(h)
        check(lib.sdl_btda_plan_reserve(h, K), "reserve")
        D = np.stack([workloads.uniform(seed + r, m, 0.1, 10.0) for r in range(K)])
        dAx, dBx, dD = (DeviceArray.from_numpy(v) for v in (Ax, Bx, D))
        ldc = (nnz + 63) // 64 * 64
        ref = DeviceArray((K, ldc), np.float64)
        out = DeviceArray((K, ldc), np.float64)
        check(lib.sdl_btda_refill_batched(h, dAx.ptr, dBx.ptr, dD.ptr, m, K, ref.ptr, ldc, None),
              "batched")
        lib.sdl_synchronize()
        a0 = lib.sdl_device_alloc_count()
        check(lib.sdl_btda_refill_batched_static(h, dAx.ptr, dBx.ptr, dD.ptr, m, K, out.ptr, ldc,
                                                 None), "static")
        lib.sdl_synchronize()
        assert lib.sdl_device_alloc_count() == a0
        X, R = out.numpy()[:, :nnz], ref.numpy()[:, :nnz]
        assert bits_equal(X, R)
        _, _, cx, _ = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, D[K - 1])
        assert bits_equal(X[K - 1], cx)
    finally:
        lib.sdl_btda_plan_destroy(h)


def test_static_refill_matches_batched_twice():
    lib = load()
    print("[no-alloc] 20000x5000 product, K = 64", flush=True)
    w = workloads.btda_20000x5000(1)
    _case(lib, w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"], w["Bp"], w["Bi"], w["Bx"],
          64, 700)
    print("[no-alloc] 3000 x (700, 500) random product, K = 13", flush=True)
    rng = np.random.default_rng(22)
    m, nA, nB = 3000, 700, 500
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.01)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.012)
    _case(lib, m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, 13, 900)


def test_static_refill_rejects_oversize():
    """m * K beyond the fixed array: an error return, nothing launched."""
    lib = load()
    rng = np.random.default_rng(23)
    m, nA, nB = 3000, 700, 500
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.01)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.012)
    h, keep = _plan(lib, m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        assert lib.sdl_btda_plan_is_pull(h) == 1
        nnz = lib.sdl_btda_plan_nnz(h)
        K = int(lib.sdl_static_scratch_bytes() // (8 * m)) + 64
        dAx, dBx = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(Bx)
        # full-size buffers, although the call must fail before any launch
        dD = DeviceArray((K, m), np.float64)
        dC = DeviceArray((K, nnz), np.float64)
        a0 = lib.sdl_device_alloc_count()
        rc = lib.sdl_btda_refill_batched_static(h, dAx.ptr, dBx.ptr, dD.ptr, m, K, dC.ptr, nnz,
                                                None)
        assert rc != 0 and b"static" in lib.sdl_last_error()
        assert lib.sdl_device_alloc_count() == a0
    finally:
        lib.sdl_btda_plan_destroy(h)
