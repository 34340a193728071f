"""Synthetic workloads of the benchmark configurations, generated by the
library's deterministic host generators (splitmix64-seeded xoshiro256**)."""
from __future__ import annotations

import numpy as np

from ._lib import load, ptr


def csr_uniform(seed: int, m: int, n: int, lo: int, hi: int, values: bool = True):
    """m x n CSR, row lengths uniform in [lo, hi], distinct uniform sorted
    columns, values N(0,1).  Returns (p, i, x) numpy arrays (x None if not values)."""
    lib = load()
    p = np.empty(m + 1, np.int32)
    nnz = lib.sdl_gen_row_lengths_uniform(seed, m, lo, hi, ptr(p))
    i = np.empty(nnz, np.int32)
    x = np.empty(nnz, np.float64) if values else None
    lib.sdl_gen_fill_rows(seed, m, n, ptr(p), ptr(i), ptr(x) if values else None)
    return p, i, x


def uniform(seed: int, count: int, lo: float, hi: float):
    out = np.empty(count, np.float64)
    load().sdl_gen_uniform(seed, count, lo, hi, ptr(out))
    return out


def normal(seed: int, count: int):
    out = np.empty(count, np.float64)
    load().sdl_gen_normal(seed, count, ptr(out))
    return out


def btda_config1(seed: int = 1, m: int = 20000, n: int = 5000, nnz_row: int = 6):
    """A and B independent m x n CSR with exactly nnz_row distinct uniform columns
    per row (values N(0,1)); d uniform [0.1, 10]."""
    Ap, Ai, Ax = csr_uniform(seed * 1000 + 1, m, n, nnz_row, nnz_row)
    Bp, Bi, Bx = csr_uniform(seed * 1000 + 2, m, n, nnz_row, nnz_row)
    d = uniform(seed * 1000 + 3, m, 0.1, 10.0)
    return dict(m=m, nA=n, nB=n, Ap=Ap, Ai=Ai, Ax=Ax, Bp=Bp, Bi=Bi, Bx=Bx, d=d)


def blocks_config3(seed: int = 3, nblocks: int = 256, lo: int = 64, hi: int = 512,
                   ncols: int = 4096):
    """Stacked permuted-dense operands: nblocks dense FP64 N(0,1) blocks, m0 and
    n0 independently log-uniform on [lo, hi] (rounded); the rows are a random
    partition of [0, sum m0) into the blocks' sorted row_perms; each col_perm
    is a sorted random n0-subset of [0, ncols).  B shares A's row partition
    with independent col_perms and values.  d uniform [0.1, 10]."""
    g = np.random.Generator(np.random.PCG64(seed))
    m0 = np.rint(np.exp(g.uniform(np.log(lo), np.log(hi), nblocks))).astype(int)
    n0A = np.rint(np.exp(g.uniform(np.log(lo), np.log(hi), nblocks))).astype(int)
    n0B = np.rint(np.exp(g.uniform(np.log(lo), np.log(hi), nblocks))).astype(int)
    m = int(m0.sum())
    rows = g.permutation(m)
    A, B = [], []
    start = 0
    for k in range(nblocks):
        rp = np.sort(rows[start:start + m0[k]]).astype(np.int32)
        start += m0[k]
        cpA = np.sort(g.choice(ncols, n0A[k], replace=False)).astype(np.int32)
        cpB = np.sort(g.choice(ncols, n0B[k], replace=False)).astype(np.int32)
        A.append((rp, cpA, g.standard_normal((m0[k], n0A[k]))))
        B.append((rp, cpB, g.standard_normal((m0[k], n0B[k]))))
    d = g.uniform(0.1, 10.0, m)
    return dict(m=m, n=ncols, A=A, B=B, d=d)
