"""GPU: the learned row binning.  Features, logits and gradients of the HIP
kernels against the CPU FP64 oracle (oracle/mlp_oracle.py, itself pinned by
finite differences); training on the GPU lowers the loss; predicted plans give
the exact plan's structure bit for bit (trained, absent and adversarial
predictors, overflow fallback); parameter files round-trip and reject
corruption; the drop-in allocs use a bound predictor; the exact path sizes
its count pass from min(terms, column span), robust to unsorted A rows; and
what the network trained on configs 1-3 gains over that exact path on a
held-out product of config 3's law."""
import ctypes as C
import os
import time

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.predictor import Plan, Predictor, training_products, training_set
from tests.helpers import random_csr

pytestmark = pytest.mark.gpu

def csc_of(p, i, m, n):
    """CSC (column pointers, row indices ascending) of a CSR structure."""
    p = np.asarray(p)
    rows = np.repeat(np.arange(m), np.diff(p))
    order = np.lexsort((rows, np.asarray(i)))
    cp = np.zeros(n + 1, np.int64)
    np.add.at(cp, np.asarray(i) + 1, 1)
    return np.cumsum(cp), rows[order]


@pytest.fixture(scope="module")
def trained():
    X, Y, n = training_set(training_products(small=True))  # configs 1-3 only
    mlp = Predictor(seed=1)
    mlp.train(X, Y, n, epochs=100, batch=2048, lr=3e-3, seed=7)
    yield mlp
    mlp.close()


def test_features_match_oracle():
    rng = np.random.default_rng(1)
    m, nA, nB = 400, 90, 70
    Ap, Ai, _ = random_csr(rng, m, nA, 0.05)
    Bp, Bi, _ = random_csr(rng, m, nB, 0.08)
    pl = Plan((Ap, Ai), (Bp, Bi), nA, nB)
    F = pl.features().numpy()
    Lp, Li = csc_of(Bp, Bi, m, nB)
    ref = MO.row_features(Lp, Li, Ap, Ai, nA)
    assert np.allclose(F, ref, rtol=1e-6, atol=1e-5), np.abs(F - ref).max()


def test_forward_and_gradients_match_oracle():
    rng = np.random.default_rng(2)
    P = (rng.standard_normal(MO.NPARAM) * 0.2).astype(np.float32)
    n = 3000
    X = (rng.random((n, MO.NIN)) * 8).astype(np.float32)
    Y = rng.integers(0, MO.NOUT, n).astype(np.int32)
    mlp = Predictor()
    mlp.set_params(P)
    dX, dY = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y)
    Z = DeviceArray((n, MO.NOUT), np.float32)
    check(load().sdl_mlp_forward(mlp.handle, dX.ptr, n, Z.ptr))
    zref = MO.forward(P.astype(np.float64), X.astype(np.float64))
    assert np.allclose(Z.numpy(), zref, rtol=1e-5, atol=1e-4)
    G = DeviceArray(MO.NPARAM, np.float32)
    loss = C.c_float()
    check(load().sdl_mlp_gradients(mlp.handle, dX.ptr, dY.ptr, n, G.ptr, C.byref(loss)))
    lref, gref = MO.loss_and_grad(P.astype(np.float64), X.astype(np.float64), Y)
    g = G.numpy()
    assert abs(loss.value - lref) <= 1e-5 * abs(lref) + 1e-6
    assert np.allclose(g, gref, rtol=1e-3, atol=1e-6), np.abs(g - gref).max()
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)
    # the tau rule on the device matches the oracle's
    bins = mlp.predict(dX, n)
    assert np.mean(bins == MO.predict_bins(P.astype(np.float64), X.astype(np.float64), 0.95)) > 0.999
    mlp.close()


def test_training_lowers_loss_and_predicts_bins(trained):
    X, Y, n = training_set(training_products(small=True))  # configs 1-3 only
    fresh = Predictor(seed=1)
    G = DeviceArray(MO.NPARAM, np.float32)
    l0, l1 = C.c_float(), C.c_float()
    check(load().sdl_mlp_gradients(fresh.handle, X.ptr, Y.ptr, n, G.ptr, C.byref(l0)))
    check(load().sdl_mlp_gradients(trained.handle, X.ptr, Y.ptr, n, G.ptr, C.byref(l1)))
    assert l1.value < 0.5 * l0.value
    pred = trained.predict(X, n)
    y = Y.numpy()
    # the tau rule leans to the larger bin: most rows exact, few under-predicted
    # (an under-predicted row is counted twice)
    assert np.mean(pred == y) > 0.7
    assert np.mean(pred < y) < 0.05
    fresh.close()


def structures_equal(a, b):
    pa, ia = a.structure()
    pb, ib = b.structure()
    return np.array_equal(pa, pb) and np.array_equal(ia, ib)


def cases():
    rng = np.random.default_rng(4)
    Ap, Ai, _ = random_csr(rng, 3000, 800, 0.01)
    Bp, Bi, _ = random_csr(rng, 3000, 600, 0.012)
    yield "uniform", (Ap, Ai), (Bp, Bi), 800, 600
    p, i, _ = W.banded_csr(4096, 16)
    yield "banded16", (p, i), None, 4096, 4096
    p, i, _ = W.zipf_csr(11, 20000, 4000)
    yield "zipf", (p, i), None, 4000, 4000
    p, i, _ = W.rmat_csr(12)
    yield "rmat12", (p, i), None, 1 << 12, 1 << 12
    wb = W.stacked_pd_blocks(5, nblocks=12, lo=16, hi=64, ncols=512)
    p, i, _ = W.blocks_as_csr(wb)
    yield "blocks", (p, i), None, 512, 512
    # wide: the term bound asks for 128 KB column bitmaps (the span bound does not)
    p, i, _ = W.strided_band_csr(1 << 12, 1 << 20, 129, 4)
    yield "strided", (p, i), None, 1 << 20, 1 << 20
    # uniform columns in a wide space: rows a prediction may size (candidates)
    p, i, _ = W.csr_uniform(15, 1 << 18, 1 << 15, 8, 8, values=False)
    yield "uniform-wide", (p, i), None, 1 << 15, 1 << 15
    # B^T A with B a dense 16 x 64 pattern and A 16 rows of 300 columns in
    # 2^20: 64 output rows of 4800 terms and ~4800 entries whose term bound
    # takes a 128 KB column bitmap, where a certain small prediction would
    # count at a fifth of the cost (certain candidates; a small prediction for
    # them overflows)
    p, i, _ = W.csr_uniform(17, 16, 1 << 20, 300, 300, values=False)
    bp = np.arange(0, 16 * 64 + 1, 64, dtype=np.int32)
    bi = np.tile(np.arange(64, dtype=np.int32), 16)
    yield "wide-bitmap", (p, i), (bp, bi), 1 << 20, 64


@pytest.mark.parametrize("which", ["trained", "none", "adversarial"])
def test_predicted_structure_is_exact(which, trained):
    adv = None
    if which == "adversarial":
        adv = Predictor()
        P = np.zeros(MO.NPARAM, np.float32)
        P[-6] = 20.0  # every row predicted in the smallest bin: tables of 64 slots
        adv.set_params(P)
    mlp = {"trained": trained, "none": None, "adversarial": adv}[which]
    overflow = 0
    for name, A, B, nA, nB in cases():
        ex = Plan(A, B, nA, nB)
        pr = Plan(A, B, nA, nB, mode=1, mlp=mlp)
        assert pr.nnz() == ex.nnz(), name
        assert structures_equal(ex, pr), name
        st = pr.pred_stats()
        assert st["predicted"] + st["exact"] + st["overflow"] <= nB
        overflow += st["overflow"]
        ex.close()
        pr.close()
    if which == "adversarial":
        assert overflow > 0  # the fallback was exercised
        adv.close()


def test_parameter_file_roundtrip_and_checksum(tmp_path, trained):
    path = str(tmp_path / "p.bin")
    trained.save(path)
    other = Predictor(seed=9)
    check(load().sdl_mlp_load(other.handle, path.encode()))
    assert np.array_equal(other.params(), trained.params())
    raw = bytearray(open(path, "rb").read())
    raw[40] ^= 0x01
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(raw))
    assert load().sdl_mlp_load(other.handle, bad.encode()) != 0
    assert np.array_equal(other.params(), trained.params())  # unchanged on rejection
    other.close()


def test_shipped_parameters_load():
    from sparsediff_linalg_amd.predictor import DEFAULT_PATH
    if not os.path.exists(DEFAULT_PATH):
        pytest.skip("no shipped predictor file")
    p = Predictor.from_file(DEFAULT_PATH)
    assert np.all(np.isfinite(p.params()))
    p.close()


def test_dropin_alloc_uses_bound_predictor(trained):
    from tests import dropin as D
    L = D.lib()
    p, i, x = W.banded_csr(2048, 8)
    A = D.sparse(p, i, x, 2048, 2048)
    exact = L.ATA_matrices_alloc(A)
    lib = load()
    lib.sdl_set_row_predictor(trained.handle)
    try:
        pred = L.ATA_matrices_alloc(A)
    finally:
        lib.sdl_set_row_predictor(None)
    pe, ie, _ = D.mat_csr(exact)
    pp, ip, _ = D.mat_csr(pred)
    assert np.array_equal(pe, pp) and np.array_equal(ie, ip)
    d = np.linspace(0.5, 2.0, 2048)
    L.ATDA_matrices_fill_values(A, D.dptr(d), exact)
    L.ATDA_matrices_fill_values(A, D.dptr(d), pred)
    assert np.array_equal(D.mat_csr(exact)[2].view(np.int64), D.mat_csr(pred)[2].view(np.int64))
    for M in (exact, pred, A):
        L.free_matrix(M)


def plan_stats(pl):
    o = (C.c_longlong * 11)()
    check(load().sdl_btda_plan_stats(pl.h, o), "plan stats")
    return list(o)


def oracle_ata(p, i, n, d):
    from oracle import oracle as O
    m = len(p) - 1
    x = np.linspace(-1.0, 1.0, int(p[-1]))
    Lp, Li, Lx = O.transpose(m, n, p, i, x)
    return x, O.spgemm(n, Lp, Li, Lx, 1, d, m, n, p, i, x)


def refill_matches(pl, x, d, Cx):
    dx, dd = DeviceArray.from_numpy(x), DeviceArray.from_numpy(d)
    out = DeviceArray(max(len(Cx), 1), np.float64)
    check(load().sdl_btda_refill(pl.h, dx.ptr, dx.ptr, dd.ptr, out.ptr, None))
    load().sdl_synchronize()
    return np.array_equal(out.numpy()[:len(Cx)].view(np.int64), Cx.view(np.int64))


def test_span_bound_sizes_the_count_pass():
    """A^T A of a strided band (4096 rows of 129 consecutive columns starting
    4 apart, in 2^20 columns): each output row has ~4160 terms, so the term
    bound would count it in a 2^20-bit column bitmap, but its columns span
    ~253, so the exact path counts every row in an LDS hash table or a span
    window.  Structure and values against the oracle."""
    n = 1 << 20
    p, i, _ = W.strided_band_csr(1 << 12, n, 129, 4)
    pl = Plan((p, i), None, n, n)
    st = plan_stats(pl)
    win = (C.c_longlong * 4)()
    check(load().sdl_btda_plan_window_rows(pl.h, win), "window rows")
    cp, ci = pl.structure()
    nonempty = int(np.count_nonzero(np.diff(cp)))
    assert st[8] + win[0] == nonempty and st[9] == 0 and st[10] == 0, (st, list(win))
    d = np.linspace(0.5, 3.0, len(p) - 1)
    x, (Cp, Ci, Cx, _) = oracle_ata(p, i, n, d)
    assert np.array_equal(cp, Cp.astype(np.int32)) and np.array_equal(ci, Ci)
    assert refill_matches(pl, x, d, Cx)
    pl.close()


@pytest.mark.parametrize("with_net", [False, True])
def test_unsorted_a_rows_exact(with_net, trained):
    """The span bound takes each A row's smallest and largest column, not its
    first and last: A rows stored with shuffled columns (the reference's
    Gustavson accepts them) give the oracle's structure and values, exact and
    predicted."""
    n = 1 << 20
    p, i, _ = W.strided_band_csr(1 << 12, n, 129, 4)
    rng = np.random.default_rng(9)
    i = i.copy()
    for r in range(len(p) - 1):
        i[p[r]:p[r + 1]] = rng.permutation(i[p[r]:p[r + 1]])
    pl = Plan((p, i), None, n, n, mode=1 if with_net else 0, mlp=trained if with_net else None)
    d = np.linspace(0.5, 3.0, len(p) - 1)
    x, (Cp, Ci, Cx, _) = oracle_ata(p, i, n, d)
    cp, ci = pl.structure()
    assert np.array_equal(cp, Cp.astype(np.int32)) and np.array_equal(ci, Ci)
    assert refill_matches(pl, x, d, Cx)
    pl.close()


def test_trained_predictor_on_held_out_blocks(trained):
    """What the network trained on configs 1-3 gains over the exact path
    (count pass sized from min(terms, span)) on a held-out product of
    config 3's law: 256 permuted-dense blocks (m0, n0 log-uniform in
    [64, 512]) read as CSR, with column permutations drawn from 2^20 columns
    instead of 4096, A^T A.  Most output rows then come from one block: ~215
    columns against ~4.6e4 terms and a span near 2^20, so only a prediction
    can size them small.  Prints both symbolic times and the statistics;
    requires the exact structure, recounts of at most 5 % of the lowered rows,
    and no loss beyond 25 % (the timing is noisy; the gain is reported)."""
    wb = W.stacked_pd_blocks(13, nblocks=256, ncols=1 << 20)
    p, i, _ = W.blocks_as_csr(wb)
    n = wb["n"]
    lib = load()

    def timed(**kw):
        Plan((p, i), None, n, n, **kw).close()  # warm-up (kernels, allocator)
        best = None
        for _ in range(3):
            lib.sdl_synchronize()
            t0 = time.perf_counter()
            pl = Plan((p, i), None, n, n, **kw)
            lib.sdl_synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[1]:
                if best:
                    best[0].close()
                best = (pl, dt)
            else:
                pl.close()
        return best

    ex, t_exact = timed()
    pr, t_pred = timed(mode=1, mlp=trained)
    assert structures_equal(ex, pr)
    st = pr.pred_stats()
    lowered = st["predicted"] + st["overflow"]
    print(f"\nheld-out blocks A^T A symbolic: exact {t_exact * 1e3:.1f} ms, predicted "
          f"{t_pred * 1e3:.1f} ms (gain {t_exact / t_pred:.2f}); {st}")
    assert st["candidates"] > 0
    assert st["overflow"] <= 0.05 * max(lowered, 1)
    assert t_pred <= 1.25 * t_exact
    for pl in (ex, pr):
        pl.close()
