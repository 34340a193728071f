"""GPU: the device-resident sum plan (sdl_sum_*) against the oracle, bit for
bit: structure, the operand-entry map, the plain and row-scaled fills, the
incremental add (both add paths), more operands than one launch carries, and
the fill captured in a HIP graph and replayed with operand values changed
between replays."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load
from tests.helpers import bits_equal, random_csr

pytestmark = pytest.mark.gpu
vp = C.c_void_p


class DevSum:
    def __init__(self, m, n, ops, keep_rows=1):
        self.lib = load()
        self.m, self.n, self.ops = m, n, ops
        self.dev = [tuple(DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in o) for o in ops]
        k = len(ops)
        self.dX = (vp * k)(*[d[2].ptr for d in self.dev])
        dP = (vp * k)(*[d[0].ptr for d in self.dev])
        dI = (vp * k)(*[d[1].ptr for d in self.dev])
        nz = (C.c_longlong * k)(*[len(o[1]) for o in ops])
        self.h = vp()
        check(self.lib.sdl_sum_plan_create(k, m, n, dP, dI, nz, keep_rows, C.byref(self.h)),
              "sum plan")
        self.nnz = self.lib.sdl_sum_plan_nnz(self.h)
        self.out = DeviceArray(max(self.nnz, 1), np.float64)

    def structure(self):
        cp = np.empty(self.m + 1, np.int64)
        ci = np.empty(self.nnz, np.int32)
        self.lib.sdl_memcpy_dtoh(cp.ctypes.data, self.lib.sdl_sum_plan_rowptr(self.h), cp.nbytes)
        if self.nnz:
            self.lib.sdl_memcpy_dtoh(ci.ctypes.data, self.lib.sdl_sum_plan_colind(self.h),
                                     ci.nbytes)
        return cp, ci

    def map(self):
        tot = sum(len(o[1]) for o in self.ops)
        mp = np.empty(tot, np.int32)
        if tot:
            self.lib.sdl_memcpy_dtoh(mp.ctypes.data, self.lib.sdl_sum_plan_map(self.h), mp.nbytes)
        return mp

    def values(self):
        self.lib.sdl_synchronize()
        return self.out.numpy()[:self.nnz]

    def close(self):
        self.lib.sdl_sum_plan_destroy(self.h)


def make_ops(rng, m, n, nops, dup_op=None, empty_op=None):
    ops = [random_csr(rng, m, n, 0.01 + 0.01 * (k % 5)) for k in range(nops)]
    if empty_op is not None:
        ops[empty_op] = (np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    if dup_op is not None:  # repeat every 3rd entry of one operand in place
        p, i, x = ops[dup_op]
        reps = (np.arange(len(i)) % 3 == 0).astype(int)
        rows = np.repeat(np.arange(m), np.diff(p))
        cnt = np.bincount(rows, weights=1 + reps, minlength=m).astype(np.int64)
        pp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
        ops[dup_op] = (pp, np.repeat(i, 1 + reps).astype(np.int32),
                       np.repeat(x, 1 + reps) * np.where(np.arange(len(i) + reps.sum()) % 2,
                                                          1.0, 1.0 / 3.0))
    return ops


def test_sum_plan_fill_scaled_and_add():
    lib = load()
    rng = np.random.default_rng(31)
    m, n = 1500, 1200
    ops = make_ops(rng, m, n, 5, dup_op=2, empty_op=3)
    S = DevSum(m, n, ops)
    assert not lib.sdl_sum_plan_is_masked(S.h)  # operand 2 repeats entries
    Cp, Ci, Cx, mp = O.sum_n(m, n, ops)
    cp, ci = S.structure()
    assert np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
    assert np.array_equal(S.map(), mp)
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), Cx)
    d = rng.uniform(-2, 2, (5, m))
    dd = DeviceArray.from_numpy(d)
    check(lib.sdl_sum_fill(S.h, S.dX, dd.ptr, S.out.ptr, None), "scaled fill")
    assert bits_equal(S.values(), O.sum_n(m, n, ops, d=d)[2])
    lib.sdl_memset_d(S.out.ptr, 0, S.nnz * 8)
    for k in range(5):
        check(lib.sdl_sum_add(S.h, k, S.dev[k][2].ptr, S.out.ptr, None), "add")
    assert bits_equal(S.values(), Cx)
    S.close()


@pytest.mark.parametrize("nops", [1, 3, 8, 9, 12, 16])
def test_sum_plan_operand_masks(nops):
    """Canonical operands (sorted unique columns), one empty: the fill streams
    values through per-entry operand masks (one byte up to 8 operands, two
    bytes up to 16); many 256- or 128-entry tiles."""
    lib = load()
    rng = np.random.default_rng(40 + nops)
    m, n = 5000, 3000
    ops = make_ops(rng, m, n, nops, empty_op=1 if nops > 2 else None)
    prev = lib.sdl_sum_scatter_ratio(0.0)  # these operands rarely overlap: force the masks
    try:
        S = DevSum(m, n, ops, keep_rows=1)
    finally:
        lib.sdl_sum_scatter_ratio(prev)
    assert lib.sdl_sum_plan_is_masked(S.h) and lib.sdl_sum_plan_fill_kind(S.h) == 1
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), O.sum_n(m, n, ops)[2])
    d = rng.uniform(-2, 2, (nops, m))
    dd = DeviceArray.from_numpy(d)
    check(lib.sdl_sum_fill(S.h, S.dX, dd.ptr, S.out.ptr, None), "scaled fill")
    assert bits_equal(S.values(), O.sum_n(m, n, ops, d=d)[2])
    S.close()


def test_sum_plan_17_operands_use_lists():
    """Beyond 16 operands the masks do not apply: the entry lists sum the same
    way."""
    lib = load()
    rng = np.random.default_rng(47)
    m, n = 900, 700
    ops = make_ops(rng, m, n, 17)
    prev = lib.sdl_sum_scatter_ratio(0.0)
    try:
        S = DevSum(m, n, ops, keep_rows=0)
    finally:
        lib.sdl_sum_scatter_ratio(prev)
    assert not lib.sdl_sum_plan_is_masked(S.h) and lib.sdl_sum_plan_fill_kind(S.h) == 0
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), O.sum_n(m, n, ops)[2])
    S.close()


@pytest.mark.parametrize("nops,dup", [(2, None), (12, 3), (30, 7)])
def test_sum_plan_scatter_fill(nops, dup):
    """Operands that rarely share an output entry and cover long runs of it
    (column bands, plus a few entries anywhere; for dup, an operand repeating
    every third entry): single-source entries written straight from their
    operand, the shared ones through the lists; plain and row-scaled fills bit
    for bit against the oracle.  Interleaved operands keep the other fills."""
    lib = load()
    rng = np.random.default_rng(50 + nops)
    m, w = 2000, 300
    n = w * nops
    ops = []
    for k in range(nops):  # operand k: a band of columns, plus ~200 entries anywhere
        band = rng.random((m, w)) < 0.2
        full = np.zeros((m, n), bool)
        full[:, k * w:(k + 1) * w] = band
        full |= rng.random((m, n)) < 0.1 / n
        p = np.concatenate([[0], np.cumsum(full.sum(1))]).astype(np.int32)
        i = np.nonzero(full)[1].astype(np.int32)
        ops.append((p, i, rng.standard_normal(len(i))))
    if dup is not None:  # repeat every third entry of one operand in place
        p, i, x = ops[dup]
        reps = (np.arange(len(i)) % 3 == 0).astype(int)
        rows = np.repeat(np.arange(m), np.diff(p))
        cnt = np.bincount(rows, weights=1 + reps, minlength=m).astype(np.int64)
        ops[dup] = (np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32),
                    np.repeat(i, 1 + reps).astype(np.int32), np.repeat(x, 1 + reps) * 0.75)
    S = DevSum(m, n, ops, keep_rows=1)
    assert lib.sdl_sum_plan_fill_kind(S.h) == 2
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops)
    cp, ci = S.structure()
    assert np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), Cx)
    d = rng.uniform(-2, 2, (nops, m))
    dd = DeviceArray.from_numpy(d)
    check(lib.sdl_sum_fill(S.h, S.dX, dd.ptr, S.out.ptr, None), "scaled fill")
    assert bits_equal(S.values(), O.sum_n(m, n, ops, d=d)[2])
    S.close()


def test_sum_plan_interleaved_large_keeps_lists_or_masks():
    """Operands that rarely share an entry but interleave within rows, into an
    output beyond 4,194,304 entries: no scatter fill (its writes would land in
    shared lines); the values bit for bit against the oracle."""
    lib = load()
    rng = np.random.default_rng(58)
    m, n = 20000, 2000
    ops = [random_csr(rng, m, n, 0.06) for _ in range(2)]
    S = DevSum(m, n, ops, keep_rows=0)
    try:
        assert S.nnz > 4 << 20
        assert lib.sdl_sum_plan_fill_kind(S.h) != 2
        check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
        assert bits_equal(S.values(), O.sum_n(m, n, ops)[2])
    finally:
        S.close()


def test_sum_plan_unsorted_columns():
    """An operand whose rows list columns out of order: same union and sums
    (operand entries in storage order), through the entry lists."""
    lib = load()
    rng = np.random.default_rng(44)
    m, n = 800, 600
    ops = make_ops(rng, m, n, 3)
    p, i, x = ops[1]
    i = i.copy()
    for r in range(0, m, 3):
        i[p[r]:p[r + 1]] = i[p[r]:p[r + 1]][::-1]
    ops[1] = (p, i, x)
    S = DevSum(m, n, ops, keep_rows=0)
    assert not lib.sdl_sum_plan_is_masked(S.h)
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops)
    cp, ci = S.structure()
    assert np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), Cx)
    S.close()


def test_sum_plan_many_operands():
    """40 operands: the launch takes its pointers from device memory."""
    lib = load()
    rng = np.random.default_rng(32)
    m, n = 400, 300
    ops = make_ops(rng, m, n, 40, dup_op=11, empty_op=39)
    S = DevSum(m, n, ops, keep_rows=0)
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops)
    cp, ci = S.structure()
    assert np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
    check(lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, None), "fill")
    assert bits_equal(S.values(), Cx)
    S.close()


def test_sum_fill_graph_replays():
    """The fill captured once, replayed three times; between replays an
    operand's device values are rewritten (eager copies) and every replay
    matches the oracle for the values it saw."""
    lib = load()
    rng = np.random.default_rng(33)
    m, n = 2000, 2000
    ops = make_ops(rng, m, n, 4)
    S = DevSum(m, n, ops, keep_rows=0)
    stream = lib.sdl_library_stream()
    lib.sdl_synchronize()
    check(lib.sdl_graph_begin(stream), "capture")
    rc = lib.sdl_sum_fill(S.h, S.dX, None, S.out.ptr, stream)
    g = vp()
    check(lib.sdl_graph_end(stream, C.byref(g)), "end capture")
    check(rc, "captured fill")
    for rep in range(3):
        x1 = rng.standard_normal(len(ops[1][1]))
        ops[1] = (ops[1][0], ops[1][1], x1)
        lib.sdl_memcpy_htod(S.dev[1][2].ptr, x1.ctypes.data, x1.nbytes)
        lib.sdl_memset_d(S.out.ptr, 0xFF, S.nnz * 8)
        check(lib.sdl_graph_launch(g, stream), "replay")
        assert bits_equal(S.values(), O.sum_n(m, n, ops)[2]), rep
    lib.sdl_graph_destroy(g)
    S.close()
