"""GPU: Hessian assembly across expression nodes (sdl_hessian_*): the sum of
node products A_k^T diag(d_k) A_k and its lower-triangular export, evaluated
call by call and replayed from one captured HIP graph, bit for bit against the
oracle's per-node Gustavson, ordered sum and lower export; the graph replayed
several times with new d and A values written in place between replays and
eager evaluations in between."""
import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, load, workloads as W
from sparsediff_linalg_amd.hessian import Hessian
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def oracle_hessian(nodes, n, ds):
    ops = []
    for (p, i, x), d in zip(nodes, ds):
        m = len(p) - 1
        Lp, Li, Lx = O.transpose(m, n, p, i, x)
        Cp, Ci, Cx, _ = O.spgemm(n, Lp, Li, Lx, 1, d, m, n, p, i, x)
        ops.append((Cp.astype(np.int32), Ci, Cx))
    Hp, Hi, Hx, _ = O.sum_n(n, n, ops)
    r, c, lx, _ = O.coo_lower(n, Hp, Hi, Hx)
    return Hp, Hi, Hx, r, c, lx


def test_hessian_eval_and_graph_replays():
    lib = load()
    nodes, n = W.lagrangian_nodes(3, n=30_000, nnodes=7, m_lo=2_000, m_hi=9_000, window=9_000)
    H = Hessian([(p, i) for p, i, _ in nodes], n)
    try:
        rng = np.random.default_rng(11)
        ds = [rng.uniform(0.1, 10, len(p) - 1) for p, _, _ in nodes]
        Ax = [DeviceArray.from_numpy(x) for _, _, x in nodes]
        D = [DeviceArray.from_numpy(d) for d in ds]
        Hx = DeviceArray(max(H.nnz, 1), np.float64)
        Lx = DeviceArray(max(H.lower_nnz, 1), np.float64)
        H.eval(Ax, D, Hx, Lx)
        lib.sdl_synchronize()
        Hp, Hi, Ox, r, c, lx = oracle_hessian(nodes, n, ds)
        hp, hi, lr, lc = H.structure()
        assert np.array_equal(hp, Hp.astype(np.int64)) and np.array_equal(hi, Hi)
        assert np.array_equal(lr, r) and np.array_equal(lc, c)
        assert bits_equal(Hx.numpy()[:H.nnz], Ox) and bits_equal(Lx.numpy()[:H.lower_nnz], lx)
        # one captured graph, replayed with new values written in place
        H.capture(Ax, D, Hx, Lx)
        Hx2 = DeviceArray(max(H.nnz, 1), np.float64)
        for rep in range(3):
            ds = [rng.uniform(0.1, 10, len(p) - 1) for p, _, _ in nodes]
            xs = [rng.standard_normal(len(x)) for _, _, x in nodes]
            for k in range(len(nodes)):
                D[k].copy_from(ds[k])
                Ax[k].copy_from(xs[k])
            H.replay()
            lib.sdl_synchronize()
            cur = [(p, i, x) for (p, i, _), x in zip(nodes, xs)]
            _, _, Ox, _, _, lx = oracle_hessian(cur, n, ds)
            assert bits_equal(Hx.numpy()[:H.nnz], Ox), rep
            assert bits_equal(Lx.numpy()[:H.lower_nnz], lx), rep
            # eager work between replays, into another buffer
            H.eval(Ax, D, Hx2)
            lib.sdl_synchronize()
            assert bits_equal(Hx2.numpy()[:H.nnz], Ox), rep
    finally:
        H.close()


@pytest.mark.parametrize("streams", [0, 1, 2])
def test_hessian_node_streams(streams):
    """The node refills in one launch over every node's tiles, in order on one
    stream, or forked over two, give the oracle's Hessian bit for bit, eagerly
    and replayed; an out-of-range stream count is rejected."""
    lib = load()
    nodes, n = W.lagrangian_nodes(4, n=20_000, nnodes=5, m_lo=1_000, m_hi=5_000, window=6_000)
    H = Hessian([(p, i) for p, i, _ in nodes], n)
    try:
        with pytest.raises(RuntimeError):
            H.set_streams(-1)
        with pytest.raises(RuntimeError):
            H.set_streams(5)
        H.set_streams(streams)
        rng = np.random.default_rng(13)
        ds = [rng.uniform(0.1, 10, len(p) - 1) for p, _, _ in nodes]
        Ax = [DeviceArray.from_numpy(x) for _, _, x in nodes]
        D = [DeviceArray.from_numpy(d) for d in ds]
        Hx = DeviceArray(max(H.nnz, 1), np.float64)
        Lx = DeviceArray(max(H.lower_nnz, 1), np.float64)
        _, _, Ox, _, _, lx = oracle_hessian(nodes, n, ds)
        H.eval(Ax, D, Hx, Lx)
        lib.sdl_synchronize()
        assert bits_equal(Hx.numpy()[:H.nnz], Ox) and bits_equal(Lx.numpy()[:H.lower_nnz], lx)
        Hx.copy_from(np.zeros(max(H.nnz, 1)))
        H.capture(Ax, D, Hx, Lx)
        for _ in range(2):
            H.replay()
            lib.sdl_synchronize()
            assert bits_equal(Hx.numpy()[:H.nnz], Ox)
    finally:
        H.close()


def test_hessian_many_nodes_one_launch_each_16():
    """20 nodes: the default path launches the node refills in groups of 16
    tiles' descriptors; the Hessian equals the oracle's bit for bit."""
    lib = load()
    nodes, n = W.lagrangian_nodes(5, n=12_000, nnodes=20, m_lo=300, m_hi=2_500, window=4_000)
    H = Hessian([(p, i) for p, i, _ in nodes], n)
    try:
        rng = np.random.default_rng(17)
        ds = [rng.uniform(0.1, 10, len(p) - 1) for p, _, _ in nodes]
        Ax = [DeviceArray.from_numpy(x) for _, _, x in nodes]
        D = [DeviceArray.from_numpy(d) for d in ds]
        Hx = DeviceArray(max(H.nnz, 1), np.float64)
        Lx = DeviceArray(max(H.lower_nnz, 1), np.float64)
        H.eval(Ax, D, Hx, Lx)
        lib.sdl_synchronize()
        _, _, Ox, _, _, lx = oracle_hessian(nodes, n, ds)
        assert bits_equal(Hx.numpy()[:H.nnz], Ox) and bits_equal(Lx.numpy()[:H.lower_nnz], lx)
    finally:
        H.close()
