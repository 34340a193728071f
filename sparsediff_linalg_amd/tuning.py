"""Host side of the plans' tuning table: the symbolic accumulator thresholds
(largest LDS hash table, bitmap density factor) and the row-accumulator
refill's row cuts, per (GPU architecture, shape signature), tuned on a
product by timing candidates and saved to a checksummed text file.

    from sparsediff_linalg_amd import tuning
    best, t_best, t_def = tuning.tune((Ap, Ai), None, n, n, values=(dAx, dAx))
    tuning.save("sdl_tune.txt")      # later: tuning.load_table("sdl_tune.txt")

Every plan created afterwards for a product of the same signature uses the
recorded thresholds.  The thresholds only choose between kernels; structures
and values are identical under every valid setting.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceArray, check, load, sdl_sym_tune

FIELDS = ("hash_max", "dens", "rowacc_cut", "rank_cut")


def sym_tune(d) -> sdl_sym_tune:
    """A C threshold struct from a dict (missing keys: the built-in values)."""
    t = sdl_sym_tune()
    check(load().sdl_sym_tune_default(C.byref(t)), "default tune")
    for k, v in dict(d).items():
        if k not in FIELDS:
            raise KeyError(f"unknown threshold {k!r}")
        setattr(t, k, int(v))
    return t


def default() -> dict:
    return sym_tune({}).as_dict()


def tune(A, B=None, nA=None, nB=None, mode=0, values=None, reps=3):
    """Times plan creation over the symbolic candidates (and, with values =
    (device A values, device B values), refills over the row cuts), records
    the winner for this product's signature and returns (thresholds, tuned
    ms, default ms)."""
    lib = load()
    Ap, Ai = A[0], A[1]
    Bp, Bi = (B[0], B[1]) if B is not None else (Ap, Ai)
    m = len(Ap) - 1
    nB = nA if nB is None else nB
    arrs = [DeviceArray.from_numpy(np.ascontiguousarray(a, np.int32)) for a in (Ap, Ai, Bp, Bi)]
    best = sdl_sym_tune()
    tb, td = C.c_float(), C.c_float()
    ax, bx = (values[0].ptr, values[1].ptr) if values is not None else (None, None)
    check(lib.sdl_btda_tune(m, int(nA), int(nB), *(a.ptr for a in arrs), mode, ax, bx, reps,
                            C.byref(best), C.byref(tb), C.byref(td)), "tune")
    return best.as_dict(), tb.value, td.value


def size() -> int:
    return load().sdl_tune_table_size()


def clear() -> None:
    load().sdl_tune_table_clear()


def save(path: str) -> None:
    check(load().sdl_tune_table_save(path.encode()), "tune table save")


def load_table(path: str) -> None:
    check(load().sdl_tune_table_load(path.encode()), "tune table load")
