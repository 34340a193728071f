"""Device-resident B^T diag(d) A plans (include/sdl_device.h) against the
oracle: structure bit-exact, values bit-exact (the reference's fold order)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads
from tests.helpers import bits_equal, max_ulp, random_csr

pytestmark = pytest.mark.gpu


class DevPlan:
    def __init__(self, m, nA, nB, Ap, Ai, Bp, Bi, mode=0):
        self.lib = load()
        self.dAp, self.dAi = DeviceArray.from_numpy(Ap), DeviceArray.from_numpy(Ai)
        self.dBp, self.dBi = DeviceArray.from_numpy(Bp), DeviceArray.from_numpy(Bi)
        h = C.c_void_p()
        check(self.lib.sdl_btda_plan_create(m, nA, nB, self.dAp.ptr, self.dAi.ptr, self.dBp.ptr,
                                            self.dBi.ptr, mode, C.byref(h)), "plan_create")
        self.h = h
        self.nB = nB
        self.nnz = self.lib.sdl_btda_plan_nnz(h)

    def structure(self):
        p = np.zeros(self.nB + 1, np.int32)
        i = np.zeros(max(self.nnz, 1), np.int32)
        check(self.lib.sdl_btda_plan_structure(self.h, p.ctypes.data_as(C.POINTER(C.c_int)),
                                               i.ctypes.data_as(C.POINTER(C.c_int))))
        return p, i[:self.nnz]

    def refill(self, Ax, Bx, d):
        dAx, dBx = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(Bx)
        dd = DeviceArray.from_numpy(d) if d is not None else None
        dC = DeviceArray(max(self.nnz, 1), np.float64)
        check(self.lib.sdl_btda_refill(self.h, dAx.ptr, dBx.ptr, dd.ptr if dd else None, dC.ptr,
                                       None))
        self.lib.sdl_synchronize()
        return dC.numpy()[:self.nnz]

    def refill_batched(self, Ax, Bx, D):
        K, m = D.shape
        dAx, dBx = DeviceArray.from_numpy(Ax), DeviceArray.from_numpy(Bx)
        dD = DeviceArray.from_numpy(D)
        dC = DeviceArray((K, max(self.nnz, 1)), np.float64)
        check(self.lib.sdl_btda_refill_batched(self.h, dAx.ptr, dBx.ptr, dD.ptr, m, K, dC.ptr,
                                               max(self.nnz, 1), None))
        self.lib.sdl_synchronize()
        return dC.numpy()[:, :self.nnz]

    def close(self):
        self.lib.sdl_btda_plan_destroy(self.h)


def _check_case(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d):
    Cp, Ci, Cx, terms = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    P = DevPlan(m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        assert P.nnz == len(Ci)
        p, i = P.structure()
        assert np.array_equal(p, Cp.astype(np.int32))
        assert np.array_equal(i, Ci)
        x = P.refill(Ax, Bx, d)
        assert bits_equal(x, Cx), f"max ulp {max_ulp(x, Cx)}"
    finally:
        P.close()


def test_btda_literal_reference_case():
    # test_csc_matrix.h:143-228 (BTA_alloc + BTDA_fill): C = {37, 47, 108}
    Ap = np.array([0, 2, 3, 5, 6], np.int32)
    Ai = np.array([0, 2, 1, 0, 2, 1], np.int32)
    Ax = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    Bp = np.array([0, 1, 2, 3, 4], np.int32)
    Bi = np.array([0, 1, 0, 1], np.int32)
    Bx = np.array([1.0, 2.0, 3.0, 4.0])
    d = np.array([1.0, 2.0, 3.0, 4.0])
    P = DevPlan(4, 3, 2, Ap, Ai, Bp, Bi)
    p, i = P.structure()
    assert list(p) == [0, 2, 3] and list(i) == [0, 2, 1]
    assert list(P.refill(Ax, Bx, d)) == [37.0, 47.0, 108.0]
    P.close()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_btda_random_small(seed):
    rng = np.random.default_rng(seed)
    m, nA, nB = 40 + seed * 17, 30 + seed, 25 + 2 * seed
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.15, dense_col=3, empty_row=5, empty_col=7)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.2, dense_col=1, empty_row=6, empty_col=2)
    d = rng.uniform(0.1, 10.0, m)
    _check_case(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    _check_case(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, None)


def test_btda_dense_and_large_rows():
    # rows of C that land in the bitmap / large hash / global classes
    rng = np.random.default_rng(7)
    m, nA, nB = 3000, 20000, 40
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.002)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.3, dense_col=0)
    d = rng.uniform(0.1, 10.0, m)
    _check_case(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)


def test_btda_accumulator_path():
    # terms > 3 nnz(C): the row-accumulator schedule (reference: no plan)
    rng = np.random.default_rng(11)
    m, nA, nB = 400, 12, 10
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.6)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.6)
    d = rng.uniform(0.1, 10.0, m)
    Cp, Ci, Cx, terms = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, d)
    assert terms > 3 * len(Ci)
    P = DevPlan(m, nA, nB, Ap, Ai, Bp, Bi)
    assert not load().sdl_btda_plan_is_pull(P.h)
    assert bits_equal(P.refill(Ax, Bx, d), Cx)
    P.close()


def test_btda_empty_operands():
    m, nA, nB = 5, 4, 3
    Ap = np.zeros(m + 1, np.int32)
    Bp = np.zeros(m + 1, np.int32)
    P = DevPlan(m, nA, nB, Ap, np.zeros(1, np.int32), Bp, np.zeros(1, np.int32))
    assert P.nnz == 0
    p, _ = P.structure()
    assert list(p) == [0, 0, 0, 0]
    P.close()


def test_btda_20000x5000_bit_exact_single_and_batched():
    w = workloads.btda_20000x5000(1)
    m, nA, nB = w["m"], w["nA"], w["nB"]
    Cp, Ci, Cx, terms = O.btda(m, nA, nB, w["Ap"], w["Ai"], w["Ax"], w["Bp"], w["Bi"], w["Bx"],
                               w["d"])
    P = DevPlan(m, nA, nB, w["Ap"], w["Ai"], w["Bp"], w["Bi"])
    try:
        p, i = P.structure()
        assert np.array_equal(p, Cp.astype(np.int32)) and np.array_equal(i, Ci)
        assert load().sdl_btda_plan_is_pull(P.h) == 1
        x = P.refill(w["Ax"], w["Bx"], w["d"])
        assert bits_equal(x, Cx), f"max ulp {max_ulp(x, Cx)}"
        K = 8
        D = np.stack([workloads.uniform(100 + r, m, 0.1, 10.0) for r in range(K)])
        XB = P.refill_batched(w["Ax"], w["Bx"], D)
        for r in (0, 3, K - 1):
            _, _, cx, _ = O.btda(m, nA, nB, w["Ap"], w["Ai"], w["Ax"], w["Bp"], w["Bi"], w["Bx"],
                                 D[r])
            assert bits_equal(XB[r], cx)
    finally:
        P.close()


@pytest.mark.parametrize("K", [16, 13])
def test_every_refill_configuration_bit_exact(K):
    """Every candidate launch configuration of the batched refill (including
    the interleaved kernels that need the reservation) gives the oracle's
    values for every refill of a K = 16 batch, and of a K = 13 batch whose
    last chunk is partial."""
    from sparsediff_linalg_amd._lib import sdl_refill_cfg
    lib = load()
    rng = np.random.default_rng(21)
    m, nA, nB = 3000, 700, 500
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.01)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.012)
    D = np.stack([workloads.uniform(300 + r, m, 0.1, 10.0) for r in range(K)])
    ref = [O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx, D[r])[2] for r in range(K)]
    P = DevPlan(m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        check(lib.sdl_btda_plan_reserve(P.h, K))
        cands = (sdl_refill_cfg * 64)()
        nc = lib.sdl_refill_candidates(cands, 64)
        assert any(cands[c].variant == 6 for c in range(nc))
        dAx, dBx, dD = (DeviceArray.from_numpy(v) for v in (Ax, Bx, D))
        ldc = (P.nnz + 63) // 64 * 64
        dC = DeviceArray((K, ldc), np.float64)
        for c in range(nc):
            cfg = cands[c]
            check(lib.sdl_btda_refill_batched_cfg(P.h, dAx.ptr, dBx.ptr, dD.ptr, m, K, dC.ptr, ldc,
                                                  C.byref(cfg), None))
            lib.sdl_synchronize()
            X = dC.numpy()[:, :P.nnz]
            for r in range(K):
                assert bits_equal(X[r], ref[r]), (cfg.variant, cfg.rchunk, r)
    finally:
        P.close()


@pytest.mark.parametrize("case", ["uniform", "dense_col", "empty", "no_d"])
def test_single_refill_kernels_agree(case):
    """The tile refill (scale factors formed once per tile; tiles of more than
    1024 L entries gather per term) and the per-term-gather refill give the
    oracle's values bit for bit; the query and the argument checks."""
    lib = load()
    rng = np.random.default_rng(61)
    if case == "dense_col":   # L rows of ~1500 entries: tiles beyond the table
        m, nA, nB = 5000, 3000, 30
        Ap, Ai, Ax = random_csr(rng, m, nA, 0.001)
        Bp, Bi, Bx = random_csr(rng, m, nB, 0.3, dense_col=0)
    elif case == "empty":     # empty rows and columns on both sides
        m, nA, nB = 3000, 900, 800
        Ap, Ai, Ax = random_csr(rng, m, nA, 0.004, empty_row=4, empty_col=9)
        Bp, Bi, Bx = random_csr(rng, m, nB, 0.004, empty_row=7, empty_col=3)
    else:
        m, nA, nB = 8000, 2000, 1500
        Ap, Ai, Ax = random_csr(rng, m, nA, 0.003)
        Bp, Bi, Bx = random_csr(rng, m, nB, 0.003)
    d = None if case == "no_d" else rng.uniform(0.1, 10.0, m)
    Cp, Ci, Cx, terms = O.btda(m, nA, nB, Ap, Ai, Ax, Bp, Bi, Bx,
                               d if d is not None else np.ones(m))
    P = DevPlan(m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        assert lib.sdl_btda_plan_is_pull(P.h) == 1
        assert lib.sdl_btda_plan_single_refill(P.h, -1) == 1   # the default
        tile = P.refill(Ax, Bx, d)
        assert lib.sdl_btda_plan_single_refill(P.h, 0) == 1
        gath = P.refill(Ax, Bx, d)
        assert lib.sdl_btda_plan_single_refill(P.h, -1) == 0
        if d is not None:
            assert bits_equal(tile, Cx), f"max ulp {max_ulp(tile, Cx)}"
        assert bits_equal(tile, gath)
        assert lib.sdl_btda_plan_single_refill(P.h, 5) == -1
        assert lib.sdl_btda_plan_single_refill(P.h, 1) == 0
    finally:
        P.close()


def test_single_refill_kind_needs_pull():
    lib = load()
    rng = np.random.default_rng(11)
    m, nA, nB = 400, 12, 10
    Ap, Ai, Ax = random_csr(rng, m, nA, 0.6)
    Bp, Bi, Bx = random_csr(rng, m, nB, 0.6)
    P = DevPlan(m, nA, nB, Ap, Ai, Bp, Bi)
    try:
        assert not lib.sdl_btda_plan_is_pull(P.h)
        assert lib.sdl_btda_plan_single_refill(P.h, -1) == 0
        assert lib.sdl_btda_plan_single_refill(P.h, 1) == -1
    finally:
        P.close()
