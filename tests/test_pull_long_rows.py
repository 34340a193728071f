"""Pull schedules of output rows longer than 2048 entries (their term lists
are built with the row's columns in global memory, searched through a sample
held in LDS): B^T A with a dense column of B, so output row 0 collects every A
row (about 6,000 and 28,000 entries, and a row of 2^17 + 1 columns whose
sample stride is 64), planned with SDL_PLAN_PULL; structure and refill values
bit-exact against the oracle (the fold in L storage order,
gustavson.c:117-141, which the pull schedule keeps per entry)."""
import numpy as np
import pytest

from sparsediff_linalg_amd import DeviceArray, check, load
from sparsediff_linalg_amd.predictor import Plan
from tests.helpers import random_csr
from tests.test_symbolic_window import oracle_btda

pytestmark = pytest.mark.gpu

SDL_PLAN_PULL = 256


def rows_csr(rng, m, n, k):
    """m rows of k distinct uniform columns in [0, n), ascending."""
    cols = [np.sort(rng.choice(n, k, replace=False)) for _ in range(m)]
    p = np.arange(0, m * k + 1, k, dtype=np.int32)
    return p, np.concatenate(cols).astype(np.int32)


@pytest.mark.parametrize("m,nA,k", [(3000, 6000, 10), (4000, 30000, 20), (1100, 131073, 120)])
def test_pull_schedule_long_rows(m, nA, k):
    rng = np.random.default_rng(m + nA)
    Ap, Ai = rows_csr(rng, m, nA, k)
    nB = 40
    Bp, Bi, _ = random_csr(rng, m, nB, 0.05, dense_col=0)
    lib = load()
    pl = Plan((Ap, Ai), (Bp, Bi), nA, nB, mode=SDL_PLAN_PULL)
    assert lib.sdl_btda_plan_is_pull(pl.h)
    xA, xB, d, Cp, Ci, Cx = oracle_btda((Ap, Ai), (Bp, Bi), m, nA, nB)
    assert np.diff(Cp).max() > 2048
    cp, ci = pl.structure()
    assert np.array_equal(cp, Cp.astype(np.int32)) and np.array_equal(ci, Ci)
    dA, dB, dd = (DeviceArray.from_numpy(v) for v in (xA, xB, d))
    out = DeviceArray(max(len(Cx), 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dA.ptr, dB.ptr, dd.ptr, out.ptr, None), "refill")
    lib.sdl_synchronize()
    assert np.array_equal(out.numpy()[:len(Cx)].view(np.int64), Cx.view(np.int64))
    pl.close()
