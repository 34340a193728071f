"""Shared test helpers: random CSR operands with the edge cases the reference
tests use (empty rows/columns, a dense column), and bit-exact comparisons."""
import numpy as np

from sparsediff_linalg_amd import workloads


def random_csr(rng, m, n, density, dense_col=None, empty_row=None, empty_col=None,
               values="normal"):
    mask = rng.random((m, n)) < density
    if dense_col is not None and n > 0:
        mask[:, dense_col] = True
    if empty_row is not None and m > 0:
        mask[empty_row, :] = False
    if empty_col is not None and n > 0:
        mask[:, empty_col] = False
    p = np.zeros(m + 1, np.int32)
    p[1:] = np.cumsum(mask.sum(1))
    rows, cols = np.nonzero(mask)
    i = cols.astype(np.int32)
    if values == "normal":
        x = rng.standard_normal(len(i))
    else:
        x = rng.uniform(-0.5, 0.5, len(i))
    return p, i, x


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def max_ulp(a, b):
    a = np.ascontiguousarray(a, np.float64).view(np.int64)
    b = np.ascontiguousarray(b, np.float64).view(np.int64)
    if a.size == 0:
        return 0
    return int(np.max(np.abs(a - b)))


def uniform_csr(seed, m, n, lo, hi):
    return workloads.csr_uniform(seed, m, n, lo, hi)


def row_term_bounds(Lm, Lp, Li, blocks, Am, Ap):
    """Terms of every output row of kron(I_blocks, L) A (the row's nnz bound)."""
    Lp = np.asarray(Lp, np.int64)
    lenA = np.diff(np.asarray(Ap, np.int64))
    nL = Am // blocks
    out = np.zeros(Lm * blocks, np.int64)
    cs = np.zeros(len(Li) + 1, np.int64)
    for b in range(blocks):
        np.cumsum(lenA[np.asarray(Li, np.int64) + b * nL], out=cs[1:])
        out[b * Lm:(b + 1) * Lm] = cs[Lp[1:]] - cs[Lp[:-1]]
    return out


def check_rows_parallel(fetch, Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax, row_ranges,
                        threads=16):
    """Checks device products' output rows against the oracle, block by
    block: fetch(r0, r1) returns a list with, for every device output under
    test, its (Cp from 0 as int64, Ci, Cx) of rows [r0, r1); the oracle
    computes the same rows once, in a thread pool (its C routine releases the
    GIL).  Structure and values must be bit-identical.  Returns the number of
    entries checked per output."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    Lp, Li, Ap, Ai = (np.ascontiguousarray(v, np.int32) for v in (Lp, Li, Ap, Ai))
    Lx, Ax = np.ascontiguousarray(Lx, np.float64), np.ascontiguousarray(Ax, np.float64)
    d = np.ascontiguousarray(d, np.float64) if d is not None else None
    bounds = np.concatenate([[0], np.cumsum(row_term_bounds(Lm, Lp, Li, blocks, Am, Ap))])
    threads = max(1, min(threads, os.cpu_count() or 1))
    checked = 0

    def job(r0, r1):
        cap = int(min(bounds[r1] - bounds[r0], (r1 - r0) * An))
        return O.spgemm_range(r0, r1, Lm, Lp, Li, Lx, blocks, d, Am, An, Ap, Ai, Ax, cap)

    with ThreadPoolExecutor(threads) as pool:
        pending = []
        ranges = list(row_ranges)
        for idx, (r0, r1) in enumerate(ranges):
            pending.append((r0, r1, pool.submit(job, r0, r1)))
            if len(pending) >= 2 * threads or idx == len(ranges) - 1:
                for a, b, fut in pending:
                    ref = fut.result()
                    assert ref is not None, (a, b)
                    for o, (gp, gi, gx) in enumerate(fetch(a, b)):
                        assert np.array_equal(gp, ref[0]), ("row pointers", o, a, b)
                        assert np.array_equal(gi, ref[1]), ("columns", o, a, b)
                        assert bits_equal(gx, ref[2]), ("values", o, a, b, max_ulp(gx, ref[2]))
                    checked += len(ref[1])
                pending = []
    return checked


def row_blocks(nrows, term_bounds, per_block=2_000_000):
    """Contiguous row ranges of about per_block terms each."""
    cs = np.concatenate([[0], np.cumsum(term_bounds)])
    out, r0 = [], 0
    while r0 < nrows:
        r1 = int(np.searchsorted(cs, cs[r0] + per_block, side="right")) - 1
        r1 = max(r0 + 1, min(r1, nrows))
        out.append((r0, r1))
        r0 = r1
    return out
