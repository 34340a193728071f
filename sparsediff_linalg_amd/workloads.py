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


def btda_20000x5000(seed: int = 1, m: int = 20000, n: int = 5000, nnz_row: int = 6):
    """A and B independent m x n CSR with exactly nnz_row distinct uniform columns
    per row (values N(0,1)); d uniform [0.1, 10]."""
    Ap, Ai, Ax = csr_uniform(seed * 1000 + 1, m, n, nnz_row, nnz_row)
    Bp, Bi, Bx = csr_uniform(seed * 1000 + 2, m, n, nnz_row, nnz_row)
    d = uniform(seed * 1000 + 3, m, 0.1, 10.0)
    return dict(m=m, nA=n, nB=n, Ap=Ap, Ai=Ai, Ax=Ax, Bp=Bp, Bi=Bi, Bx=Bx, d=d)


def stacked_pd_blocks(seed: int = 3, nblocks: int = 256, lo: int = 64, hi: int = 512,
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


def csr_from_lengths(seed: int, lengths, n: int, values: bool = True):
    """CSR with the given row lengths, distinct uniform sorted columns per row
    (values N(0,1)), by the library's generator."""
    lib = load()
    lengths = np.asarray(lengths, np.int64)
    m = len(lengths)
    p = np.zeros(m + 1, np.int64)
    np.cumsum(lengths, out=p[1:])
    assert p[-1] < 2 ** 31, "nnz exceeds int32"
    p = p.astype(np.int32)
    i = np.empty(int(p[-1]), np.int32)
    x = np.empty(int(p[-1]), np.float64) if values else None
    lib.sdl_gen_fill_rows(seed, m, n, ptr(p), ptr(i), ptr(x) if values else None)
    return p, i, x


def zipf_lengths(seed: int, m: int, s: float = 1.2, lo: int = 1, hi: int = 512,
                 mean: float = 16.0):
    """Row lengths by rank-Zipf: length(rank r) = clip(round(c r^-s), lo, hi)
    with c solved (bisection) so the mean is `mean`; ranks randomly permuted
    over the rows."""
    r = np.arange(1, m + 1, dtype=np.float64)
    base = r ** -s

    def mean_of(c):
        return np.clip(np.rint(c * base), lo, hi).mean()

    a, b = 1.0, 1e12
    for _ in range(200):
        c = np.sqrt(a * b)
        if mean_of(c) < mean:
            a = c
        else:
            b = c
    lengths = np.clip(np.rint(b * base), lo, hi).astype(np.int64)
    g = np.random.Generator(np.random.PCG64(seed))
    return lengths[g.permutation(m)]


def zipf_csr(seed: int, m: int, n: int, mean: float = 16.0):
    """The A of the large A^T A product: m x n, rank-Zipf(1.2) row lengths in
    [1, 512] with mean `mean`, distinct uniform columns, FP64 N(0,1)."""
    return csr_from_lengths(seed, zipf_lengths(seed, m, mean=mean), n)


def strided_band_csr(m: int, n: int, width: int, stride: int, seed: int = 5):
    """m x n: row i holds the `width` consecutive columns starting at
    (stride * i) mod (n - width + 1); FP64 N(0,1) values.  Each column of A^T A
    collects about width / stride rows whose columns span about
    width + width - 1 around it, so the product rows have width * width /
    stride terms but only ~2 * width - 1 distinct columns."""
    start = (stride * np.arange(m, dtype=np.int64)) % (n - width + 1)
    p = np.arange(m + 1, dtype=np.int64) * width
    i = (start[:, None] + np.arange(width, dtype=np.int64)[None, :]).reshape(-1).astype(np.int32)
    x = normal(seed, m * width)
    return p.astype(np.int32), i, x


def banded_csr(m: int, bw: int, seed: int = 5):
    """m x m banded: row i holds columns max(0, i-bw) .. min(m-1, i+bw); FP64
    N(0,1) values."""
    lo = np.maximum(0, np.arange(m) - bw)
    hi = np.minimum(m - 1, np.arange(m) + bw)
    lengths = (hi - lo + 1).astype(np.int64)
    p = np.zeros(m + 1, np.int64)
    np.cumsum(lengths, out=p[1:])
    row = np.repeat(np.arange(m), lengths)
    i = (lo[row] + (np.arange(p[-1]) - p[row])).astype(np.int32)
    x = normal(seed, int(p[-1]))
    return p.astype(np.int32), i, x


def rmat_csr(scale: int, edges_per_node: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 5):
    """R-MAT graph on 2^scale nodes with edges_per_node * 2^scale generated
    edges (quadrant probabilities a, b, c, 1-a-b-c per level), duplicates
    coalesced (one entry per (row, col), values summed); FP64 N(0,1) values."""
    n = 1 << scale
    E = edges_per_node * n
    g = np.random.Generator(np.random.PCG64(seed))
    rows = np.zeros(E, np.int64)
    cols = np.zeros(E, np.int64)
    cum = np.array([a, a + b, a + b + c])
    for lvl in range(scale):
        u = g.random(E)
        q = np.searchsorted(cum, u, side="right")  # 0..3
        rows |= ((q >> 1) & 1).astype(np.int64) << lvl
        cols |= (q & 1).astype(np.int64) << lvl
    vals = normal(seed * 1000 + 7, E)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    uniq, start = np.unique(key, return_index=True)
    x = np.add.reduceat(vals, start) if len(start) else vals[:0]
    r = (uniq // n).astype(np.int64)
    i = (uniq % n).astype(np.int32)
    p = np.zeros(n + 1, np.int64)
    np.add.at(p, r + 1, 1)
    np.cumsum(p, out=p)
    return p.astype(np.int32), i, x


def blocks_as_csr(w):
    """CSR (m x n) of a stacked permuted-dense operand from stacked_pd_blocks."""
    m, n = w["m"], w["n"]
    rows = [None] * m
    for rp, cp, X in w["A"]:
        for ii, r in enumerate(rp):
            rows[r] = (cp, X[ii])
    lengths = np.array([len(rows[r][0]) if rows[r] is not None else 0 for r in range(m)])
    p = np.zeros(m + 1, np.int64)
    np.cumsum(lengths, out=p[1:])
    i = np.empty(p[-1], np.int32)
    x = np.empty(p[-1], np.float64)
    for r in range(m):
        if rows[r] is not None:
            i[p[r]:p[r + 1]] = rows[r][0]
            x[p[r]:p[r + 1]] = rows[r][1]
    return p.astype(np.int32), i, x


def multi_instances(seed: int = 4, ninst: int = 4096, lo: int = 100, hi: int = 50000,
                    nnz_lo: int = 4, nnz_hi: int = 12):
    """Independent B^T D A instances as block-diagonal operands: instance i has
    m_i log-uniform on [lo, hi] rows and n_i = m_i // 4 columns; A_i and B_i
    have row lengths uniform in [nnz_lo, min(nnz_hi, n_i)] with distinct
    uniform columns, FP64 N(0,1).  Returns the block-diagonal CSR of A and B
    (global column indices) with the instances' row and column offsets."""
    g = np.random.Generator(np.random.PCG64(seed))
    m = np.rint(np.exp(g.uniform(np.log(lo), np.log(hi), ninst))).astype(np.int64)
    n = m // 4
    roff = np.zeros(ninst + 1, np.int64)
    np.cumsum(m, out=roff[1:])
    coff = np.zeros(ninst + 1, np.int64)
    np.cumsum(n, out=coff[1:])
    M = int(roff[-1])
    assert M < 2 ** 31 and coff[-1] < 2 ** 31
    lib = load()
    out = {}
    for tag, salt in (("A", 1), ("B", 2)):
        lengths = np.empty(M, np.int64)
        tmp = np.empty(int(m.max()) + 1, np.int32)
        for i in range(ninst):
            s = seed * 10_000_000 + i * 10 + salt
            hi_i = int(min(nnz_hi, n[i]))
            lib.sdl_gen_row_lengths_uniform(s, int(m[i]), nnz_lo, hi_i, ptr(tmp[:m[i] + 1]))
            lengths[roff[i]:roff[i + 1]] = np.diff(tmp[:m[i] + 1])
        p = np.zeros(M + 1, np.int64)
        np.cumsum(lengths, out=p[1:])
        assert p[-1] < 2 ** 31
        i_all = np.empty(int(p[-1]), np.int32)
        x_all = np.empty(int(p[-1]), np.float64)
        for i in range(ninst):
            s = seed * 10_000_000 + i * 10 + salt
            a, b = int(p[roff[i]]), int(p[roff[i + 1]])
            lp = (p[roff[i]:roff[i + 1] + 1] - a).astype(np.int32)
            lib.sdl_gen_fill_rows(s, int(m[i]), int(n[i]), ptr(lp), ptr(i_all[a:b]),
                                  ptr(x_all[a:b]))
            i_all[a:b] += np.int32(coff[i])
        out[tag] = (p.astype(np.int32), i_all, x_all)
    return dict(m=m, n=n, roff=roff, coff=coff, A=out["A"], B=out["B"])


def coo_triplets(seed: int = 2, n_items: int = 50_000_000, unique: int = 35_000_000,
                 m: int = 1_000_000, n: int = 1_000_000):
    """COO triplets for coalescing: `unique` distinct (row, col) keys drawn
    uniformly from the m x n space, then n_items - unique triplets re-using
    uniformly chosen earlier keys, the whole list shuffled; values N(0,1).
    Returns (rows int32, cols int32, x float64)."""
    g = np.random.Generator(np.random.PCG64(seed))
    space = np.int64(m) * np.int64(n)
    keys = np.unique(g.integers(0, space, int(unique * 1.001) + 64, dtype=np.int64))
    while len(keys) < unique:
        keys = np.unique(np.concatenate([keys, g.integers(0, space, unique, dtype=np.int64)]))
    keys = g.permutation(keys)[:unique]
    rep = keys[g.integers(0, unique, n_items - unique)]
    allk = g.permutation(np.concatenate([keys, rep]))
    rows = (allk // n).astype(np.int32)
    cols = (allk % n).astype(np.int32)
    x = normal(seed * 1000 + 11, n_items)
    return rows, cols, x


def sum_operands(seed: int = 6, m: int = 1_000_000, n: int = 1_000_000, nops: int = 4,
                 per_row: int = 16, shared: float = 0.5):
    """Operands of a Hessian-assembly sum: nops m x n CSR matrices with about
    per_row entries per row.  Operand 0 draws its columns uniformly; every
    later operand keeps each of operand 0's draws with probability `shared`
    and draws the others afresh (repeated draws in a row collapse).  Values
    N(0,1).  Returns [(p int32, i int32, x float64), ...]."""
    g = np.random.Generator(np.random.PCG64(seed))
    rows = np.repeat(np.arange(m, dtype=np.int64), per_row)
    base = rows * n + g.integers(0, n, m * per_row, dtype=np.int64)
    ops = []
    for k in range(nops):
        if k == 0:
            keys = base
        else:
            keep = g.random(m * per_row) < shared
            keys = np.where(keep, base, rows * n + g.integers(0, n, m * per_row, dtype=np.int64))
        keys = np.unique(keys)
        p = np.zeros(m + 1, np.int64)
        np.cumsum(np.bincount(keys // n, minlength=m), out=p[1:])
        i = (keys % n).astype(np.int32)
        ops.append((p.astype(np.int32), i, normal(seed * 100 + k, len(i))))
    return ops


def lagrangian_nodes(seed: int = 7, n: int = 200_000, nnodes: int = 12, m_lo: int = 20_000,
                     m_hi: int = 120_000, nnz_lo: int = 3, nnz_hi: int = 9,
                     window: int = 50_000):
    """Chain-rule nodes of a synthetic Lagrangian Hessian over n variables:
    node k is an m_k x n Jacobian-like CSR (m_k log-uniform in [m_lo, m_hi],
    row lengths uniform in [nnz_lo, nnz_hi], distinct columns uniform in a
    window of `window` variables starting at a node-specific offset, so the
    nodes' Hessian blocks overlap), FP64 N(0,1) values.  Returns a list of
    (p, i, x) and n."""
    g = np.random.Generator(np.random.PCG64(seed))
    nodes = []
    for k in range(nnodes):
        m = int(np.rint(np.exp(g.uniform(np.log(m_lo), np.log(m_hi)))))
        w = min(window, n)
        c0 = int(g.integers(0, n - w + 1))
        p, i, x = csr_uniform(seed * 1000 + k, m, w, nnz_lo, min(nnz_hi, w))
        nodes.append((p, (i + c0).astype(np.int32), x))
    return nodes, n
