"""The plans' tuning table (include/sdl_device.h, sdl_tune_* / sdl_btda_tune).

CPU: the saved file's format, whose checksum is recomputed here with an
independent FNV-1a-64; corrupt, out-of-range and truncated files are rejected
whole and leave the table as it was.
GPU: every valid threshold setting gives the oracle's structure and values bit
for bit (symbolic hash / bitmap / global classes on a 2^20-column product,
LDS / rank / dense row accumulators on a variable-bandwidth product); the
tuner records its winner under the product's signature, later plans of that
signature take it, and it survives save / clear / load."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sparsediff_linalg_amd import DeviceArray, check, load, tuning, workloads as W
from sparsediff_linalg_amd.predictor import Plan
from tests.helpers import bits_equal

HEADER = "sdl-tune-table 1\n"


def fnv1a64(data: bytes) -> int:
    h = 0xcbf29ce484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def table_text(lines):
    body = HEADER + "".join(l + "\n" for l in lines)
    return body + f"checksum {fnv1a64(body.encode()):016x}\n"


@pytest.fixture
def empty_table():
    tuning.clear()
    yield
    tuning.clear()


def test_fnv1a64_known_answers():
    # published FNV-1a-64 test vectors
    assert fnv1a64(b"") == 0xcbf29ce484222325
    assert fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert fnv1a64(b"foobar") == 0x85944171f73967e8


def test_shipped_predictor_checksum_is_fnv1a64():
    """The predictor file closes with the same checksum: FNV-1a-64 of the
    parameter bytes (magic 8 bytes, count int32, tau float32, parameters)."""
    import os
    import struct
    from sparsediff_linalg_amd.predictor import DEFAULT_PATH
    raw = open(DEFAULT_PATH, "rb").read()
    assert raw[:8] == b"SDLMLP1\0"
    n = struct.unpack("<i", raw[8:12])[0]
    assert len(raw) == 16 + 4 * n + 8
    assert struct.unpack("<Q", raw[-8:])[0] == fnv1a64(raw[16:16 + 4 * n])


def test_table_file_round_trip(tmp_path, empty_table):
    text = table_text(["gfx950 20 20 5 2048 4 512 16384 3 8",
                       "gfx942 18 19 3 8192 8 1024 65536 -1 8"])
    src = tmp_path / "t.txt"
    src.write_text(text)
    tuning.load_table(str(src))
    assert tuning.size() == 2
    out = tmp_path / "u.txt"
    tuning.save(str(out))
    assert out.read_text() == text
    # loading the same entries again replaces them, adds nothing
    tuning.load_table(str(out))
    assert tuning.size() == 2


@pytest.mark.parametrize("bad", [
    "flip", "range_hash", "range_dens", "range_cut", "refill", "header", "truncated", "trailing",
    "short_line"])
def test_table_file_rejected_whole(tmp_path, empty_table, bad):
    good = ["gfx950 20 20 5 2048 4 512 16384 3 8"]
    src = tmp_path / "ok.txt"
    src.write_text(table_text(good))
    tuning.load_table(str(src))
    assert tuning.size() == 1
    entries = {
        "range_hash": ["gfx950 10 10 1 3000 4 512 16384 -1 8"],   # not a power of two
        "range_dens": ["gfx950 10 10 1 2048 0 512 16384 -1 8"],
        "range_cut": ["gfx950 10 10 1 2048 4 512 512 -1 8"],       # rank_cut <= rowacc_cut
        "refill": ["gfx950 10 10 1 2048 4 512 16384 3 12"],        # no such launch shape
        "short_line": ["gfx950 10 10 1 2048 4 512 16384 3"],
    }
    if bad in entries:
        text = table_text(good + entries[bad])
    else:
        text = table_text(good + ["gfx950 10 10 1 2048 4 512 16384 -1 8"])
        if bad == "flip":
            text = text.replace("gfx950 10 10 1 2048", "gfx950 10 10 1 4096")
        elif bad == "header":
            text = "sdl-tune-table 2\n" + text[len(HEADER):]
        elif bad == "truncated":
            text = text[:text.rfind("checksum")]
        elif bad == "trailing":
            text = text + "x"
    f = tmp_path / "bad.txt"
    f.write_text(text)
    with pytest.raises(RuntimeError):
        tuning.load_table(str(f))
    assert tuning.size() == 1


def test_defaults_and_unknown_key():
    assert tuning.default() == {"hash_max": 8192, "dens": 8, "rowacc_cut": 1024,
                                "rank_cut": 65536}
    with pytest.raises(KeyError):
        tuning.sym_tune({"hashmax": 1024})


# ---------------------------------------------------------------- GPU
def variable_band(seg=600, bws=(8, 40, 100, 200), seed=0):
    """Rows r of segment s hold the columns r - bw_s .. r + bw_s: A^T A rows of
    about 33 to 800 entries, far more terms than entries (the row-accumulator
    refill)."""
    rows = []
    n = seg * len(bws)
    for s, bw in enumerate(bws):
        for r in range(s * seg, (s + 1) * seg):
            rows.append(np.arange(max(0, r - bw), min(n - 1, r + bw) + 1))
    lens = np.array([len(r) for r in rows])
    p = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    i = np.concatenate(rows).astype(np.int32)
    x = np.random.default_rng(seed).standard_normal(len(i))
    return p, i, x, n


def refill(pl, dAx, d, nnz):
    lib = load()
    dd = DeviceArray.from_numpy(d)
    dC = DeviceArray(max(nnz, 1), np.float64)
    check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr, dC.ptr, None), "refill")
    lib.sdl_synchronize()
    return dC.numpy()[:nnz]


SETTINGS = [
    {},
    {"hash_max": 1024, "dens": 2},
    {"hash_max": 2048, "dens": 32},
    {"hash_max": 4096, "dens": 1},
    {"rowacc_cut": 256, "rank_cut": 600},
    {"rowacc_cut": 64, "rank_cut": 65},
    {"hash_max": 1024, "dens": 64, "rowacc_cut": 512, "rank_cut": 8192},
]


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["wide_zipf", "variable_band"])
def test_every_setting_bit_exact_vs_oracle(which, empty_table):
    if which == "wide_zipf":
        p, i, x = W.zipf_csr(3, 6000, 1 << 20)
        n = 1 << 20
    else:
        p, i, x, n = variable_band()
    m = len(p) - 1
    d = np.random.default_rng(5).uniform(0.1, 10.0, m)
    Cp, Ci, Cx, terms = O.btda(m, n, n, p, i, x, p, i, x, d)
    dAx = DeviceArray.from_numpy(x)
    for s in SETTINGS:
        print(f"[tuning] {which} {s}", flush=True)
        pl = Plan((p, i), None, n, n, tune=s)
        try:
            want = dict(tuning.default(), **s)
            assert pl.tune() == want
            if which == "variable_band":
                assert not load().sdl_btda_plan_is_pull(pl.h)
            sp, si = pl.structure()
            assert np.array_equal(sp, Cp.astype(np.int32)), s
            assert np.array_equal(si, Ci), s
            assert bits_equal(refill(pl, dAx, d, len(Ci)), Cx), s
        finally:
            pl.close()


@pytest.mark.gpu
def test_out_of_range_setting_is_an_error(empty_table):
    p, i, x, n = variable_band(seg=50)
    for s in ({"hash_max": 3000}, {"hash_max": 16384}, {"rowacc_cut": 2048},
              {"rank_cut": 1 << 17}, {"dens": 0}):
        with pytest.raises(RuntimeError):
            Plan((p, i), None, n, n, tune=s)


@pytest.mark.gpu
def test_tuner_records_and_table_round_trips(tmp_path, empty_table):
    p, i, x, n = variable_band()
    dAx = DeviceArray.from_numpy(x)
    best, t_best, t_def = tuning.tune((p, i), None, n, n, values=(dAx, dAx), reps=1)
    print(f"[tuning] best {best} {t_best:.3f} ms vs default {t_def:.3f} ms", flush=True)
    assert t_best <= t_def
    assert tuning.size() == 1
    pl = Plan((p, i), None, n, n)
    assert pl.tune() == best
    pl.close()
    path = str(tmp_path / "tune.txt")
    tuning.save(path)
    tuning.clear()
    pl = Plan((p, i), None, n, n)
    assert pl.tune() == tuning.default()
    pl.close()
    tuning.load_table(path)
    pl = Plan((p, i), None, n, n)
    assert pl.tune() == best
    # a product of another signature keeps the built-in values
    q, j, _, n2 = variable_band(seg=40)
    other = Plan((q, j), None, n2, n2)
    assert other.tune() == tuning.default()
    other.close()
    # the tuned plan still refills bit-exactly
    d = np.random.default_rng(9).uniform(0.1, 10.0, len(p) - 1)
    Cp, Ci, Cx, _ = O.btda(len(p) - 1, n, n, p, i, x, p, i, x, d)
    assert bits_equal(refill(pl, dAx, d, len(Ci)), Cx)
    pl.close()


@pytest.mark.gpu
def test_autotune_records_refill_shape(empty_table):
    w = W.btda_20000x5000(1)
    lib = load()
    m, nA, nB = w["m"], w["nA"], w["nB"]
    pl = Plan((w["Ap"], w["Ai"]), (w["Bp"], w["Bi"]), nA, nB)
    K = 16
    nnz = pl.nnz()
    dAx, dBx = DeviceArray.from_numpy(w["Ax"]), DeviceArray.from_numpy(w["Bx"])
    D = DeviceArray.from_numpy(np.random.default_rng(3).uniform(0.1, 10.0, (K, m)))
    dC = DeviceArray((K, nnz), np.float64)
    check(lib.sdl_btda_plan_reserve(pl.h, K), "reserve")
    from sparsediff_linalg_amd._lib import sdl_refill_cfg
    best, us = sdl_refill_cfg(), C.c_float()
    check(lib.sdl_btda_autotune(pl.h, dAx.ptr, dBx.ptr, D.ptr, m, K, dC.ptr, nnz, 3,
                                C.byref(best), C.byref(us)), "autotune")
    assert tuning.size() == 1
    pl.close()
