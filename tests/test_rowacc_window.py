"""Row-accumulator refills (plans without a pull schedule: terms > 3 nnz, the
reference's accumulator policy, gustavson.h:45-52) of rows whose columns span
at most 1024 columns run one wavefront per row with a dense LDS accumulator
over the span; the other rows keep the workgroup accumulator with a column
search.  Both fold each entry's terms from +0.0 in L storage order
(gustavson.c:117-141), so refills with the windows on and off must agree bit
for bit, and with the oracle's values.  Cases: a band (spans of 129), B^T A
with spans of 200..12000 columns (every window size, the search rows and
rows beyond the 1024-entry bin), and rows drawn from a few wide patterns
(many terms per entry, spans too wide for a window)."""
import ctypes as C

import ctypes as C

import numpy as np
import pytest

from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan
from tests.test_symbolic_window import oracle_btda, window_ladder

pytestmark = pytest.mark.gpu


def patterns(seed, m=3000, n=5000, npat=12, width=90):
    """m rows, each a copy of one of npat random patterns of `width` columns
    in [0, n): every output entry of A^T A collects ~m / npat terms."""
    rng = np.random.default_rng(seed)
    pats = [np.sort(rng.choice(n, width, replace=False)) for _ in range(npat)]
    which = rng.integers(0, npat, m)
    p = np.concatenate([[0], np.cumsum([width] * m)]).astype(np.int32)
    return p, np.concatenate([pats[w] for w in which]).astype(np.int32), n


def cases():
    p, i, _ = W.banded_csr(4096, 32)
    yield "band32", (p, i), None, 4096, 4096
    A, B, nA, nB = window_ladder(5, blocks=6)
    yield "window-ladder", A, B, nA, nB
    p, i, n = patterns(3)
    yield "patterns", (p, i), None, n, n
    p, i, n = wide_blocks()
    yield "wide-blocks", (p, i), None, n, n
    p, i, n = wide_blocks(5, nblocks=40, ncols=2048)
    yield "wide-blocks-2048", (p, i), None, n, n
    # output rows of 1500..9000 entries spanning as many columns (the 2048,
    # 4096 and 8192 windows, and beyond them the dense accumulator: n = 2^20
    # leaves no by-rank kernel), ~15 terms per entry
    A, B, nA, nB = window_ladder(7, spans=(1500, 3000, 5000, 7000, 8192, 9000), blocks=3)
    yield "wide-ladder", A, B, nA, nB


def refill(A, B, nA, nB, xA, xB, d, mode, wide=None):
    lib = load()
    prev = lib.sdl_rowacc_window(mode)
    try:
        pl = Plan(A, B, nA, nB)
    finally:
        lib.sdl_rowacc_window(prev)
    assert not lib.sdl_btda_plan_is_pull(pl.h)
    win = (C.c_longlong * 4)()
    check(lib.sdl_btda_plan_window_rows(pl.h, win), "window rows")
    if mode == 0:
        assert win[2] == 0 and win[3] == 0
    elif wide is not None:
        assert (win[3] > 0) == wide, list(win)
    dA, dB, dd = (DeviceArray.from_numpy(v) for v in (xA, xB, d))
    out = DeviceArray(max(pl.nnz(), 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dA.ptr, dB.ptr, dd.ptr, out.ptr, None), "refill")
    lib.sdl_synchronize()
    cp, ci = pl.structure()
    x = out.numpy()[:pl.nnz()].copy()
    pl.close()
    return cp, ci, x


def wide_blocks(seed=3, nblocks=48, ncols=4096):
    """Stacked dense blocks as CSR: output rows of 1025..4096 entries spanning
    at most 4096 columns, with hundreds of terms per entry (no recorded
    positions), so the rank bin's rows take the 2048- / 4096-column windows."""
    wb = W.stacked_pd_blocks(seed, nblocks=nblocks, lo=64, hi=400, ncols=ncols)
    p, i, _ = W.blocks_as_csr(wb)
    return p, i, wb["n"]


@pytest.mark.parametrize("name", ["band32", "window-ladder", "patterns", "wide-blocks",
                                  "wide-blocks-2048", "wide-ladder"])
def test_window_accumulator_matches_oracle_and_search(name):
    A, B, nA, nB = next((c[1], c[2], c[3], c[4]) for c in cases() if c[0] == name)
    m = len(A[0]) - 1
    xA, xB, d, Cp, Ci, Cx = oracle_btda(A, B, m, nA, nB)
    outs = [refill(A, B, nA, nB, xA, xB, d, mode, wide=name.startswith("wide") or None)
            for mode in (0, 1)]
    for cp, ci, x in outs:
        assert np.array_equal(cp, Cp.astype(np.int32)) and np.array_equal(ci, Ci), name
        assert np.array_equal(x.view(np.int64), Cx.view(np.int64)), name


def test_rowacc_window_knob():
    lib = load()
    prev = lib.sdl_rowacc_window(-1)
    assert prev in (0, 1)
    assert lib.sdl_rowacc_window(2) == -1 and lib.sdl_rowacc_window(-2) == -1
    assert lib.sdl_rowacc_window(prev) == prev


def test_kron_blocks_through_windows():
    """C = kron(I_p, A) J through the drop-in BA_kron product (L = A repeated
    over p blocks, block b's L entries reading J rows offset by b n): A banded
    (41 entries per row), J's row t the 40 columns from 4 t, so every output
    row spans ~200 columns with ~8 terms per entry -- span-window symbolic
    rows and window-accumulator refills in every block.  Windows on and off,
    structure and values bit-exact against the oracle."""
    import ctypes as C
    from oracle import oracle as O
    from tests import dropin as D
    lib = load()
    m = n = 300
    p = 4
    Ap, Ai, _ = W.banded_csr(n, 20)
    Ax = np.linspace(-1.0, 1.0, int(Ap[-1]))
    rows = n * p
    ncols = 4 * rows + 40
    Jp = np.arange(0, rows * 40 + 1, 40, dtype=np.int32)
    Ji = (4 * np.repeat(np.arange(rows), 40) + np.tile(np.arange(40), rows)).astype(np.int32)
    Jx = np.linspace(0.5, 2.0, len(Ji))
    Op, Oi, Ox, terms = O.spgemm(m, Ap, Ai, Ax, p, None, rows, ncols, Jp, Ji, Jx)
    assert terms > 3 * Op[-1]
    L = D.lib()
    for mode in (1, 0):
        prev = (lib.sdl_symbolic_window(mode), lib.sdl_rowacc_window(mode))
        try:
            A = D.sparse(Ap, Ai, Ax, m, n)
            J = D.sparse(Jp, Ji, Jx, rows, ncols)
            Cm = L.BA_kron_matrices_alloc(A, p, J)
            L.BA_kron_matrices_fill_values(A, p, J, Cm)
            q, i, xx = D.mat_csr(Cm)
            assert np.array_equal(np.asarray(q), Op) and np.array_equal(np.asarray(i), Oi), mode
            assert np.array_equal(np.asarray(xx).view(np.int64), Ox.view(np.int64)), mode
            L.free_matrix(Cm)
            L.free_matrix(J)
            L.free_matrix(A)
        finally:
            lib.sdl_symbolic_window(prev[0])
            lib.sdl_rowacc_window(prev[1])


@pytest.mark.parametrize("name", ["band32", "wide-ladder"])
def test_batched_refill_row_accumulators(name):
    """sdl_btda_refill_batched on a row-accumulator plan (windows on): K = 3
    d vectors in one call, each refill bit-exact against the oracle's values
    for its d."""
    A, B, nA, nB = next((c[1], c[2], c[3], c[4]) for c in cases() if c[0] == name)
    m = len(A[0]) - 1
    lib = load()
    pl = Plan(A, B, nA, nB)
    assert not lib.sdl_btda_plan_is_pull(pl.h)
    nnz = pl.nnz()
    K = 3
    ds = [np.linspace(0.5 + r, 3.0 + r, m) for r in range(K)]
    ref = [oracle_btda(A, B, m, nA, nB, d) for d in ds]
    xA, xB = ref[0][0], ref[0][1]
    dA, dB = DeviceArray.from_numpy(xA), DeviceArray.from_numpy(xB)
    dD = DeviceArray.from_numpy(np.concatenate(ds))
    ldc = (nnz + 63) // 64 * 64
    out = DeviceArray(K * ldc, np.float64)
    check(lib.sdl_btda_refill_batched(pl.h, dA.ptr, dB.ptr, dD.ptr, m, K, out.ptr, ldc, None),
          "batched refill")
    lib.sdl_synchronize()
    got = out.numpy()
    for r in range(K):
        Cx = ref[r][5]
        assert np.array_equal(got[r * ldc:r * ldc + nnz].view(np.int64), Cx.view(np.int64)), r
    pl.close()
