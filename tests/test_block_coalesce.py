"""The stacked-PD output coalesce by output row (sdl_block_coalesce_mode 0,
the default) against the per-cell pull lists (mode 1) on the GPU: both sum
each output cell over its source blocks in ascending order from +0.0
(stacked_pd_coalesce.c:326-381), so structures and values must agree bit for
bit.  Cases: A^T D A, B^T D A and B A of random stacks (a block with an
empty column list), the full 256-block workload, and
an output wider than the row coalesce's LDS row (which keeps the lists).
Values are also checked against dense FP64 products in test_dropin_pd.py."""
import zlib

import numpy as np
import pytest

from oracle import pd_oracle as PO
from tests import dropin as D
from tests.helpers import random_csr

pytestmark = pytest.mark.gpu


def make(M):
    if "blocks" not in M:
        return D.pd(M["m"], M["n"], M["rp"], M["cp"], M["X"])
    return D.spd(M["m"], M["n"], [D.pd(b["m"], b["n"], b["rp"], b["cp"], b["X"])
                                  for b in M["blocks"]])


def stack(rng, m, n, nb, lo, hi, empty_cols=()):
    """nb blocks on disjoint rows (a stack partitions rows), row and column
    counts log-uniform in [lo, hi], sorted lists (new_permuted_dense asserts
    them, permuted_dense.c:237-245); blocks may share columns."""
    bl = []
    rows = rng.permutation(m)
    r0 = 0
    for k in range(nb):
        m0, n0 = np.exp(rng.uniform(np.log(lo), np.log(hi), 2)).astype(int)
        m0 = min(m0, m - r0)
        if m0 <= 0:
            break
        rp = np.sort(rows[r0:r0 + m0])
        r0 += m0
        cp = np.sort(rng.choice(n, 0 if k in empty_cols else min(n, n0), replace=False))
        bl.append(PO.pd(m, n, rp.tolist(), cp.tolist(),
                        rng.standard_normal((len(rp), len(cp)))))
    return PO.spd(m, n, bl)


def run_both(alloc, fill):
    """Alloc + fill under pull lists, then by row; returns both CSR views."""
    L = D.lib()
    out = []
    prev = L.sdl_block_coalesce_mode(1)
    try:
        for mode in (1, 0):
            assert L.sdl_block_coalesce_mode(mode) in (0, 1)
            Cm = alloc()
            fill(Cm)
            out.append(D.mat_csr(Cm))
            L.free_matrix(Cm)
    finally:
        L.sdl_block_coalesce_mode(prev)
    return out


def same(a, b):
    (p1, i1, x1), (p0, i0, x0) = a, b
    assert np.array_equal(p1, p0), "row pointers differ"
    assert np.array_equal(i1, i0), "column indices differ"
    assert np.array_equal(np.asarray(x1).view(np.int64), np.asarray(x0).view(np.int64)), \
        "values differ"


@pytest.mark.parametrize("seed", [31, 32])
def test_atda_random_stacks(seed):
    L = D.lib()
    rng = np.random.default_rng(seed)
    m, n = 900, 500
    Ao = stack(rng, m, n, 20, 8, 120, empty_cols=(3,))
    Ah = make(Ao)
    d = rng.uniform(0.1, 10.0, m)
    lists, rows = run_both(lambda: L.ATA_matrices_alloc(Ah),
                           lambda Cm: L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm))
    same(lists, rows)
    p, i, _ = rows
    ep, ei = PO.csr_pattern(PO.ata(Ao))
    assert list(p) == ep and list(i) == ei
    L.free_matrix(Ah)


def operand(rng, k, m, n):
    """A handle of kind k (m x n): a stack, one PD block or a CSR matrix."""
    if k == "spd":
        return make(stack(rng, m, n, 14, 8, 100))
    if k == "pd":
        rp = np.sort(rng.choice(m, min(m, 90), replace=False)).tolist()
        cp = np.sort(rng.choice(n, min(n, 70), replace=False)).tolist()
        return make(PO.pd(m, n, rp, cp, rng.standard_normal((len(rp), len(cp)))))
    p, i, x = random_csr(rng, m, n, 0.05)
    return D.sparse(p, i, x, m, n)


@pytest.mark.parametrize("bk,ak", [("spd", "spd"), ("spd", "pd"), ("pd", "spd"),
                                   ("sparse", "spd"), ("spd", "sparse")])
def test_btda_stacks(bk, ak):
    L = D.lib()
    rng = np.random.default_rng(zlib.crc32(f"btda{bk}{ak}".encode()))
    m, nA, nB = 700, 400, 350
    Bh, Ah = operand(rng, bk, m, nB), operand(rng, ak, m, nA)
    d = rng.uniform(0.1, 10.0, m)
    lists, rows = run_both(lambda: L.BTA_matrices_alloc(Ah, Bh),
                           lambda Cm: L.BTDA_matrices_fill_values(Ah, D.dptr(d), Bh, Cm))
    same(lists, rows)
    L.free_matrix(Ah)
    L.free_matrix(Bh)


@pytest.mark.parametrize("bk", ["pd", "sparse"])
def test_ba_stacks(bk):
    L = D.lib()
    rng = np.random.default_rng(zlib.crc32(f"ba{bk}".encode()))
    mB, k, nA = 150, 700, 400
    Bh, Ah = operand(rng, bk, mB, k), operand(rng, "spd", k, nA)
    lists, rows = run_both(lambda: L.BA_matrices_alloc(Bh, Ah),
                           lambda Cm: L.BA_matrices_fill_values(Bh, Ah, Cm))
    same(lists, rows)
    L.free_matrix(Ah)
    L.free_matrix(Bh)


def test_full_size_stacked_blocks():
    """256 blocks, m0 / n0 log-uniform 64..512, col_perms in [0, 4096)."""
    from sparsediff_linalg_amd import workloads
    L = D.lib()
    w = workloads.stacked_pd_blocks(3)
    m, N, d = w["m"], w["n"], w["d"]
    Ah = make(PO.spd(m, N, [PO.pd(m, N, rp.tolist(), cp.tolist(), X) for rp, cp, X in w["A"]]))
    lists, rows = run_both(lambda: L.ATA_matrices_alloc(Ah),
                           lambda Cm: L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm))
    same(lists, rows)
    L.free_matrix(Ah)


def test_wide_output_keeps_lists():
    """n = 9000 > 8192 output columns: the row coalesce does not apply and
    both modes run the pull lists."""
    L = D.lib()
    rng = np.random.default_rng(33)
    m, n = 600, 9000
    Ah = make(stack(rng, m, n, 8, 20, 300))
    d = rng.uniform(0.1, 10.0, m)
    lists, rows = run_both(lambda: L.ATA_matrices_alloc(Ah),
                           lambda Cm: L.ATDA_matrices_fill_values(Ah, D.dptr(d), Cm))
    same(lists, rows)
    L.free_matrix(Ah)


def test_mode_argument_checked():
    L = D.lib()
    assert L.sdl_block_coalesce_mode(7) == -1
    prev = L.sdl_block_coalesce_mode(0)
    assert prev in (0, 1)
    assert L.sdl_block_coalesce_mode(prev) == 0
