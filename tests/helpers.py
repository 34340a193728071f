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
