"""Span windows in the symbolic phase: rows whose columns all lie in a span of
at most 8192 columns are counted and filled in an LDS mark array over that
span (no hash table, no sort).  With windows off and on, at every
lanes-per-L-entry setting, and with the hash and bitmap kernels expanding
terms by lane groups or by binary search, plans must give the
oracle's structure (gustavson.c:43-68 count, :207-231 fill order: ascending
columns per row), and a refill over a window-built plan the oracle's values.
Cases: banded and strided bands (narrow spans), spans of every window size
and beyond it, products whose spans are wide or whose rows take the column
bitmap (no window row), shuffled A rows, and B^T A with B != A."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan
from tests.helpers import random_csr

pytestmark = pytest.mark.gpu


def window_rows(pl):
    """Rows counted and filled in span windows (the first two of the four
    window figures)."""
    o = (C.c_longlong * 4)()
    check(load().sdl_btda_plan_window_rows(pl.h, o), "window rows")
    return list(o)[:2]


def oracle_btda(A, B, m, nA, nB, d=None):
    """Structure and values of B^T diag(d) A by the oracle, with A's values
    linspace(-1, 1) and B's linspace(1, 2) (B is A when None)."""
    Ap, Ai = A
    Bp, Bi = B if B is not None else A
    xA = np.linspace(-1.0, 1.0, int(Ap[-1]))
    xB = xA if B is None else np.linspace(1.0, 2.0, int(Bp[-1]))
    if d is None:
        d = np.linspace(0.5, 3.0, m)
    Lp, Li, Lx = O.transpose(m, nB, Bp, Bi, xB)
    Cp, Ci, Cx, _ = O.spgemm(nB, Lp, Li, Lx, 1, d, m, nA, Ap, Ai, xA)
    return xA, xB, d, Cp, Ci, Cx


def window_ladder(seed, spans=(200, 400, 900, 1800, 3500, 7000, 8192, 12000), blocks=4,
                  R=16, S=8, n=1 << 20):
    """B^T A whose output rows have spans of every window size and beyond:
    output row j (column j of B) draws R consecutive rows of A, row t of them
    holding the consecutive columns [c0 + t S, c0 + t S + w) with w = span -
    (R - 1) S, so the row's columns fill [c0, c0 + span) densely; one B row in
    20 also feeds a random other output row (wider, irregular spans).  A is
    512 x 2^20, B 512 x 32."""
    rng = np.random.default_rng(seed)
    Ap, Ai, Bi = [0], [], []
    c0 = 0
    nb = len(spans) * blocks
    for s, span in enumerate(spans):
        w = span - (R - 1) * S
        for b in range(blocks):
            for t in range(R):
                Ai.append(np.arange(c0 + t * S, c0 + t * S + w))
                Ap.append(Ap[-1] + w)
                bc = {s * blocks + b}
                if rng.random() < 0.05:
                    bc.add(int(rng.integers(0, nb)))
                Bi.append(sorted(bc))
            c0 += span + 37
    Bp = np.concatenate([[0], np.cumsum([len(x) for x in Bi])])
    return ((np.asarray(Ap, np.int32), np.concatenate(Ai).astype(np.int32)),
            (Bp.astype(np.int32), np.concatenate(Bi).astype(np.int32)), n, nb)


def cases():
    p, i, _ = W.banded_csr(4096, 64)
    yield "banded64", (p, i), None, 4096, 4096, True
    p, i, _ = W.strided_band_csr(1 << 12, 1 << 20, 129, 4)
    yield "strided", (p, i), None, 1 << 20, 1 << 20, True
    A, B, nA, nB = window_ladder(5)
    yield "window-ladder", A, B, nA, nB, True
    rng = np.random.default_rng(8)
    Ap, Ai, _ = random_csr(rng, 3000, 800, 0.01)
    Bp, Bi, _ = random_csr(rng, 3000, 600, 0.012)
    yield "uniform", (Ap, Ai), (Bp, Bi), 800, 600, False
    p, i, _ = W.rmat_csr(12)
    yield "rmat12", (p, i), None, 1 << 12, 1 << 12, False
    # B^T A: A a band, B rows of a few columns near the same diagonal
    p, i, _ = W.banded_csr(5000, 40)
    bp, bi, _ = W.banded_csr(5000, 3)
    yield "band-btda", (p, i), (bp, bi), 5000, 5000, True


@pytest.mark.parametrize("mode,lanes,groups", [(0, 0, 0), (0, 0, 1), (0, 4, 1), (1, 0, 1),
                                               (1, 4, 1), (1, 16, 1), (1, 64, 1)])
def test_window_structure_matches_oracle(mode, lanes, groups):
    """Windows off / on, lanes per L entry automatic or forced, and the hash
    and bitmap kernels' term expansion by lane groups or by binary search."""
    lib = load()
    prev = lib.sdl_symbolic_window(mode)
    prev_lanes = lib.sdl_symbolic_window_lanes(lanes)
    prev_groups = lib.sdl_symbolic_lane_groups(groups)
    try:
        for name, A, B, nA, nB, narrow in cases():
            m = len(A[0]) - 1
            pl = Plan(A, B, nA, nB)
            cp, ci = pl.structure()
            win = window_rows(pl)
            _, _, _, Cp, Ci, _ = oracle_btda(A, B, m, nA, nB)
            assert np.array_equal(cp, Cp.astype(np.int32)), (name, mode)
            assert np.array_equal(ci, Ci), (name, mode)
            if mode == 0:
                assert win == [0, 0], (name, win)
            elif narrow:
                assert win[0] > 0 and win[1] > 0, (name, mode, win)
            pl.close()
    finally:
        lib.sdl_symbolic_window(prev)
        lib.sdl_symbolic_window_lanes(prev_lanes)
        lib.sdl_symbolic_lane_groups(prev_groups)


def test_window_knobs_checked():
    lib = load()
    prev = lib.sdl_symbolic_window(-1)
    assert prev in (0, 1)
    assert lib.sdl_symbolic_window(2) == -1
    assert lib.sdl_symbolic_window(-2) == -1
    assert lib.sdl_symbolic_window(-1) == prev
    assert lib.sdl_symbolic_window(prev) == prev
    lanes = lib.sdl_symbolic_window_lanes(-1)
    for bad in (1, 2, 3, 12, 128, -2):
        assert lib.sdl_symbolic_window_lanes(bad) == -1
    assert lib.sdl_symbolic_window_lanes(-1) == lanes
    groups = lib.sdl_symbolic_lane_groups(-1)
    assert groups in (0, 1)
    assert lib.sdl_symbolic_lane_groups(2) == -1 and lib.sdl_symbolic_lane_groups(-2) == -1
    assert lib.sdl_symbolic_lane_groups(-1) == groups


def test_window_plan_refill_and_shuffled_rows():
    """Shuffled A rows (the span takes each row's true smallest and largest
    column) through the default window mode: structure and refill values
    bit-exact against the oracle."""
    lib = load()
    n = 1 << 20
    p, i, _ = W.strided_band_csr(1 << 12, n, 129, 4)
    rng = np.random.default_rng(12)
    i = i.copy()
    for r in range(len(p) - 1):
        i[p[r]:p[r + 1]] = rng.permutation(i[p[r]:p[r + 1]])
    pl = Plan((p, i), None, n, n)
    assert lib.sdl_symbolic_window(-1) == 0 or window_rows(pl)[1] > 0
    xA, _, d, Cp, Ci, Cx = oracle_btda((p, i), None, len(p) - 1, n, n)
    cp, ci = pl.structure()
    assert np.array_equal(cp, Cp.astype(np.int32)) and np.array_equal(ci, Ci)
    dx, dd = DeviceArray.from_numpy(xA), DeviceArray.from_numpy(d)
    out = DeviceArray(max(len(Cx), 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dx.ptr, dx.ptr, dd.ptr, out.ptr, None), "refill")
    lib.sdl_synchronize()
    assert np.array_equal(out.numpy()[:len(Cx)].view(np.int64), Cx.view(np.int64))
    pl.close()
