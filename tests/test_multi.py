"""GPU: many B^T D A instances as one block-diagonal product (reduced scale of
the 4096-instance workload).  Per instance, structure and values equal the
oracle's product of that instance alone, bit for bit; the per-call path (one
launch per instance and d) equals the batched rounds; the captured graph of
all rounds replays to the eager result every time, with eager work between
the replays; the autotuned launch shape gives the same values."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, workloads as W
from sparsediff_linalg_amd.multi import MultiProduct, kmajor_ld

pytestmark = pytest.mark.gpu

ROUNDS = 3


@pytest.fixture(scope="module")
def setup():
    w = W.multi_instances(6, ninst=40, lo=100, hi=3000)
    mp = MultiProduct(w, max_terms=150_000, K=2)  # several groups, 2 rounds per launch
    M = mp.M
    rounds = [W.uniform(700 + r, M, 0.1, 10.0) for r in range(ROUNDS)]
    D = DeviceArray.from_numpy(np.concatenate(rounds))            # round-major
    ldt = kmajor_ld(ROUNDS)
    dt = np.zeros((M, ldt))
    dt[:, :ROUNDS] = np.stack(rounds, axis=1)
    DT = DeviceArray.from_numpy(dt)                               # k-major
    eager = []
    for r in range(ROUNDS):
        mp.refill_round(D, r)
        load().sdl_synchronize()
        eager.append(mp.values())
    yield w, mp, D, (DT, ldt), eager
    mp.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def test_groups_cover_instances(setup):
    w, mp, _, _, _ = setup
    assert len(mp.g) >= 2
    assert mp.g[0].i0 == 0 and mp.g[-1].i1 == len(w["m"])
    for a, b in zip(mp.g, mp.g[1:]):
        assert a.i1 == b.i0
    assert all(g.pull for g in mp.g)


@pytest.mark.parametrize("r", [0, ROUNDS - 1])
def test_instances_match_oracle(setup, r):
    w, mp, D, _, eager = setup
    d_all = D.numpy()[r * mp.M:(r + 1) * mp.M]
    lib = load()
    for gi, g in enumerate(mp.g):
        hp = np.empty(g.n + 1, np.int32)
        hi = np.empty(max(g.nnz, 1), np.int32)
        check(lib.sdl_btda_plan_structure(g.h, hp.ctypes.data_as(C.POINTER(C.c_int)),
                                          hi.ctypes.data_as(C.POINTER(C.c_int))))
        for k in range(g.i0, g.i1):
            r0, r1 = int(w["roff"][k]), int(w["roff"][k + 1])
            c0, c1 = int(w["coff"][k]), int(w["coff"][k + 1])
            ops = []
            for tag in ("A", "B"):
                p, i, x = w[tag]
                a, b = int(p[r0]), int(p[r1])
                ops.append(((p[r0:r1 + 1] - a).astype(np.int32), (i[a:b] - c0).astype(np.int32),
                            x[a:b]))
            (Ap, Ai, Ax), (Bp, Bi, Bx) = ops
            Cp, Ci, Cx, _ = O.btda(r1 - r0, c1 - c0, c1 - c0, Ap, Ai, Ax, Bp, Bi, Bx, d_all[r0:r1])
            e0, e1 = g.ranges[k - g.i0]
            lc0, lc1 = c0 - g.c0, c1 - g.c0
            assert np.array_equal(hp[lc0:lc1 + 1] - hp[lc0], Cp), k
            assert np.array_equal(hi[e0:e1] - lc0, Ci), k
            assert np.array_equal(bits(eager[r][gi][e0:e1]), bits(Cx)), k


def test_per_call_path_equals_batched(setup):
    _, mp, D, _, eager = setup
    mp.plain_round(D, 1)
    load().sdl_synchronize()
    for a, b in zip(mp.values(), eager[1]):
        assert np.array_equal(bits(a), bits(b))


def test_kmajor_rounds_match_eager(setup):
    _, mp, _, (DT, ldt), eager = setup
    mp.refill_rounds(DT, ldt, 0, 2)
    load().sdl_synchronize()
    for slot in (0, 1):
        for a, b in zip(mp.values(slot), eager[slot]):
            assert np.array_equal(bits(a), bits(b)), slot


def test_graph_replays_match_eager(setup):
    _, mp, D, (DT, ldt), eager = setup
    lib = load()
    g = mp.capture(DT, ldt, ROUNDS)  # launches: rounds 0-1, then round 2
    try:
        for rep in range(3):
            check(lib.sdl_graph_launch(g, None))
            lib.sdl_synchronize()
            for a, b in zip(mp.values(0), eager[2]):
                assert np.array_equal(bits(a), bits(b)), rep
            for a, b in zip(mp.values(1), eager[1]):
                assert np.array_equal(bits(a), bits(b)), rep
            # eager work between replays
            mp.refill_round(D, 0)
            lib.sdl_synchronize()
            for a, b in zip(mp.values(), eager[0]):
                assert np.array_equal(bits(a), bits(b))
    finally:
        lib.sdl_graph_destroy(g)


def test_autotuned_shape_same_values(setup):
    _, mp, D, _, eager = setup
    best = mp.autotune(D, reps=2)
    assert len(best) == len(mp.g)
    mp.refill_round(D, 2)
    load().sdl_synchronize()
    for a, b in zip(mp.values(), eager[2]):
        assert np.array_equal(bits(a), bits(b))


def test_tile_layout_matches_pairs(setup):
    """Values bound in tile layout (sdl_btda_plan_reserve_tile): single
    rounds, k-major rounds and graph replays equal the operand-pair path,
    which test_instances_match_oracle pins to the oracle."""
    w, _, D, (DT, ldt), eager = setup
    lib = load()
    mt = MultiProduct(w, max_terms=150_000, K=2, tile=True)
    try:
        for r in range(ROUNDS):
            mt.refill_round(D, r)
            lib.sdl_synchronize()
            for a, b in zip(mt.values(), eager[r]):
                assert np.array_equal(bits(a), bits(b)), r
        mt.refill_rounds(DT, ldt, 0, 2)
        lib.sdl_synchronize()
        for slot in (0, 1):
            for a, b in zip(mt.values(slot), eager[slot]):
                assert np.array_equal(bits(a), bits(b)), slot
        g = mt.capture(DT, ldt, ROUNDS)
        try:
            for rep in range(2):
                check(lib.sdl_graph_launch(g, None))
                lib.sdl_synchronize()
                for a, b in zip(mt.values(0), eager[2]):
                    assert np.array_equal(bits(a), bits(b)), rep
        finally:
            lib.sdl_graph_destroy(g)
    finally:
        mt.close()
