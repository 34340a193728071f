"""GPU: the drop-in dense helpers (include/sparsediff_linalg/dense.h) and their
device forms.  repeat / tile_double / scaled_ones / A_transpose are copies
(checked against numpy exactly; A_transpose on the reference's 70 x 45 case,
test_dense.h:8-33); the Kronecker-vector products bit for bit against the
oracle (pinned on the reference's own mini_numpy.c build)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load
from tests import dropin as D
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def test_repeat_tile_fill():
    L = D.lib()
    a = np.array([1.5, -2.0, 3.25, 0.0, -0.0])
    out = np.full(15, np.nan)
    L.repeat(D.dptr(out), D.dptr(a), 5, 3)
    assert bits_equal(out, np.repeat(a, 3))
    L.tile_double(D.dptr(out), D.dptr(a), 5, 3)
    assert bits_equal(out, np.tile(a, 3))
    big = np.random.default_rng(1).standard_normal(100_000)
    o2 = np.empty(700_000)
    L.repeat(D.dptr(o2), D.dptr(big), len(big), 7)
    assert bits_equal(o2, np.repeat(big, 7))
    L.tile_double(D.dptr(o2), D.dptr(big), len(big), 7)
    assert bits_equal(o2, np.tile(big, 7))
    o3 = np.zeros(1001)
    L.scaled_ones(D.dptr(o3), 1001, -2.5)
    assert np.all(o3 == -2.5)
    L.repeat(D.dptr(out), D.dptr(a), 0, 3)  # empty: nothing written
    L.scaled_ones(D.dptr(o3), 0, 1.0)
    assert np.all(o3 == -2.5)


@pytest.mark.parametrize("mn", [(70, 45), (1, 1), (300, 257), (33, 1), (1, 65)])
def test_a_transpose(mn):
    L = D.lib()
    m, n = mn
    A = np.arange(m * n, dtype=np.float64)
    AT = np.full(m * n, np.nan)
    L.A_transpose(D.dptr(AT), D.dptr(A), m, n)
    assert np.array_equal(AT, A.reshape(m, n).T.reshape(-1))


@pytest.mark.parametrize("shape", [(3, 2, 4), (37, 19, 70), (1, 1, 1), (64, 5, 33), (5, 3, 0),
                                   (130, 40, 129), (512, 256, 3), (2048, 1024, 2),
                                   (3, 600, 900), (20000, 16, 64), (1024, 512, 2)])
def test_kron_vec_products(shape):
    L = D.lib()
    m, k, n = shape
    rng = np.random.default_rng(m * 131 + k * 7 + n)
    Y = rng.standard_normal(k * n)
    X = rng.standard_normal(m * k)
    w = rng.standard_normal(m * n)
    v1 = np.full(max(m * k, 1), np.nan)
    L.Y_kron_I_vec(m, k, n, D.dptr(Y), D.dptr(w), D.dptr(v1))
    assert bits_equal(v1[:m * k], O.y_kron_i_vec(m, k, n, Y, w))
    v2 = np.full(max(k * n, 1), np.nan)
    L.I_kron_XT_vec(m, k, n, D.dptr(X), D.dptr(w), D.dptr(v2))
    if k * n:
        assert bits_equal(v2[:k * n], O.i_kron_xt_vec(m, k, n, X, w))


def test_kron_vec_device_forms():
    lib = load()
    rng = np.random.default_rng(9)
    m, k, n = 1000, 64, 700
    Y, X, w = rng.standard_normal(k * n), rng.standard_normal(m * k), rng.standard_normal(m * n)
    dY, dX, dw = (DeviceArray.from_numpy(a) for a in (Y, X, w))
    dv1, dv2 = DeviceArray(m * k, np.float64), DeviceArray(k * n, np.float64)
    check(lib.sdl_y_kron_i_vec(m, k, n, dY.ptr, dw.ptr, dv1.ptr, None), "y_kron_i")
    check(lib.sdl_i_kron_xt_vec(m, k, n, dX.ptr, dw.ptr, dv2.ptr, None), "i_kron_xt")
    lib.sdl_synchronize()
    assert bits_equal(dv1.numpy(), O.y_kron_i_vec(m, k, n, Y, w))
    assert bits_equal(dv2.numpy(), O.i_kron_xt_vec(m, k, n, X, w))
