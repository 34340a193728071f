"""GPU: the by-rank rows of a row-accumulator plan refilled from recorded term
positions (the plan stores each term's entry position within its row; the
entries of several terms become a sorted item list; rows whose list
outgrows LDS go to the dense accumulator), against the oracle bit for bit,
and against the per-refill ranking of the same plan (SDL_PLAN_RANK_REFILL).

Products: A^T A of A = [banded block; rank-Zipf(1.2) block] (the banded
rows -- 4000, bandwidth 100 or 150, their own columns -- put the product's
terms above 3x its entries, so the row-accumulator schedule; the Zipf rows --
row lengths in [1, 512], mean 16 -- give the long output rows, whose terms
are 1.02x and 1.14x their entries, the second with 2210 rows whose
multi-term list passes 1024 terms)."""
import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

RANK_REFILL = 512


def refill(pl, dAx, d, nnz):
    lib = load()
    dd = DeviceArray.from_numpy(d) if d is not None else None
    dC = DeviceArray(max(nnz, 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr if dd else None, dC.ptr, None),
          "refill")
    lib.sdl_synchronize()
    return dC.numpy()[:nnz]


def mixed(mz, nz, brows, bw, seed=3):
    """A = [banded brows x brows (bandwidth bw); rank-Zipf mz x nz in the
    next nz columns]."""
    bp, bi, bx = W.banded_csr(brows, bw)
    zp, zi, zx = W.zipf_csr(seed, mz, nz)
    p = np.concatenate([bp, zp[1:] + bp[-1]]).astype(np.int32)
    i = np.concatenate([bi, zi + brows]).astype(np.int32)
    x = np.concatenate([bx, zx])
    return p, i, x, brows + nz


@pytest.fixture(scope="module", params=[(20000, 50000, 4000, 100), (30000, 20000, 4000, 150)],
                ids=["zipf20000x50000", "zipf30000x20000"])
def product(request):
    p, i, x, n = mixed(*request.param)
    m = len(p) - 1
    d = np.random.default_rng(11).uniform(0.1, 10.0, m)
    Cp, Ci, Cx, _ = O.btda(m, n, n, p, i, x, p, i, x, d)
    _, _, Cx1, _ = O.btda(m, n, n, p, i, x, p, i, x, None)
    return dict(m=m, n=n, nz=request.param[1], p=p, i=i, x=x, d=d, Cp=Cp, Ci=Ci, Cx=Cx,
                Cx1=Cx1)


@pytest.mark.parametrize("setting", ["recorded", "recorded_cut640", "ranked_each_refill"])
def test_rank_bin_bit_exact(product, setting):
    P = product
    n = P["n"]
    kw = {}
    if setting == "ranked_each_refill":
        kw["mode"] = RANK_REFILL
    if setting == "recorded_cut640":
        kw["tune"] = {"rowacc_cut": 640}  # the banded rows (<= 601 entries) stay in LDS
    print(f"[rank] m={P['m']} n={n} {setting}", flush=True)
    pl = Plan((P["p"], P["i"]), None, n, n, **kw)
    try:
        sch = pl.schedule()
        print(f"[rank] schedule {sch}", flush=True)
        assert not sch["pull"] and sch["rank_rows"] > 0
        assert sch["rank_positions"] == (setting != "ranked_each_refill")
        if sch["rank_positions"] and P["nz"] == 20000:
            assert sch["dense_rows"] > 0 and sch["items"] > 0  # both side paths used
        sp, si = pl.structure()
        assert np.array_equal(sp, P["Cp"].astype(np.int32))
        assert np.array_equal(si, P["Ci"])
        dAx = DeviceArray.from_numpy(P["x"])
        assert bits_equal(refill(pl, dAx, P["d"], len(P["Ci"])), P["Cx"])
        assert bits_equal(refill(pl, dAx, None, len(P["Ci"])), P["Cx1"])
        # refills repeat: the recorded lists are read, never consumed
        assert bits_equal(refill(pl, dAx, P["d"], len(P["Ci"])), P["Cx"])
    finally:
        pl.close()


def test_negative_zero_single_terms():
    """A single-term entry is 0 + x: a product of -0.0 stores +0.0, as the
    accumulators' fold does."""
    p, i, x, n = mixed(20000, 50000, 4000, 100)
    m = len(p) - 1
    x = x.copy()
    x[::7] = -0.0
    Cp, Ci, Cx, _ = O.btda(m, n, n, p, i, x, p, i, x, None)
    pl = Plan((p, i), None, n, n)
    try:
        assert pl.schedule()["rank_positions"]
        got = refill(pl, DeviceArray.from_numpy(x), None, len(Ci))
        assert bits_equal(got, Cx)
        assert not np.any(np.signbit(got) & (got == 0.0))
    finally:
        pl.close()
