"""ORACLE (test infrastructure only): restatement of the reference's structure
rules for permuted_dense / stacked_pd products, with values from dense FP64
products (textbook GEMM in numpy).

The reference computes these values with CBLAS dgemm (vendor unpinned,
CMakeLists.txt:127), whose summation order is not specified, so the checks are
to a tolerance, |C - C_ref| <= 1e-13 * (|B|^T |D| |A|) per entry; structure is
exact.  Each rule cites the reference function it restates.

Representation: PD = dict(m, n, rp, cp, X) with X an m0 x n0 array; SPD =
dict(m, n, blocks=[PD, ...]); sparse = scipy-free dense arrays are used for
values, CSR arrays for structure.
"""
from __future__ import annotations

import numpy as np


def pd(m, n, rp, cp, X=None):
    rp = list(rp)
    cp = list(cp)
    if X is None:
        X = np.zeros((len(rp), len(cp)))
    return dict(m=m, n=n, rp=rp, cp=cp, X=np.asarray(X, float).reshape(len(rp), len(cp)))


def spd(m, n, blocks):
    return dict(m=m, n=n, blocks=list(blocks))


def blocks(M):
    return M["blocks"] if "blocks" in M else [M]


def dense(M):
    out = np.zeros((M["m"], M["n"]))
    for b in blocks(M):
        if len(b["rp"]) and len(b["cp"]):
            out[np.ix_(b["rp"], b["cp"])] = b["X"]
    return out


def csr_pattern(M):
    """CSR p/i of the union of row-disjoint blocks (pd_blocks_csr_pattern,
    permuted_dense.c:344-367)."""
    rows = [None] * M["m"]
    for b in blocks(M):
        for r in b["rp"]:
            rows[r] = list(b["cp"])
    p = [0]
    i = []
    for r in range(M["m"]):
        if rows[r]:
            i.extend(rows[r])
        p.append(len(i))
    return p, i


def overlap(a, b):
    return len(set(a) & set(b)) > 0


def union(lists):
    return sorted(set().union(*[set(x) for x in lists])) if lists else []


# ------------------------------------------------------------- coalesce ----
def coalesce(m, n, raw):
    """stacked_pd_coalesce.c:244-308: rows grouped by signature (ascending
    raw-block list), signatures in row-scan order, col_perm = sorted union."""
    row_blocks = {}
    for k, b in enumerate(raw):
        for r in b["rp"]:
            row_blocks.setdefault(r, []).append(k)
    order = []
    sig_rows = {}
    for r in range(m):
        if r not in row_blocks:
            continue
        s = tuple(row_blocks[r])
        if s not in sig_rows:
            sig_rows[s] = []
            order.append(s)
        sig_rows[s].append(r)
    res = []
    for s in order:
        b = pd(m, n, sig_rows[s], union([raw[k]["cp"] for k in s]))
        b["src"] = list(s)  # contributing raw blocks (src_block_idx)
        res.append(b)
    return res


# ----------------------------------------------------------- structures ----
def bta_pd_pd(B, A):  # permuted_dense_linalg.c:118-146
    if not overlap(B["rp"], A["rp"]):
        return pd(B["n"], A["n"], [], [])
    return pd(B["n"], A["n"], B["cp"], A["cp"])


def bta_pd_spd(B, A):  # stacked_pd_linalg.c:250-316
    cps = [a["cp"] for a in blocks(A) if overlap(B["rp"], a["rp"])]
    return pd(B["n"], A["n"], B["cp"], union(cps))


def bta_pd_csc(B, Ap, Ai, An):  # permuted_dense_linalg.c:542-566 (A given as CSR)
    hit = set()
    rows = set(B["rp"])
    for r in range(len(Ap) - 1):
        if r in rows:
            hit.update(Ai[Ap[r]:Ap[r + 1]])
    return pd(B["n"], An, B["cp"], sorted(hit))


def bta_csc_pd(Bp, Bi, Bn, A):  # permuted_dense_linalg.c:646-667 (B given as CSR)
    hit = set()
    rows = set(A["rp"])
    for r in range(len(Bp) - 1):
        if r in rows:
            hit.update(Bi[Bp[r]:Bp[r + 1]])
    return pd(Bn, A["n"], sorted(hit), A["cp"])


def ba_csc_pd(Bp, Bi, Bm, A):  # B A with sparse B read by rows (matmul_dispatchers.c:147-194)
    rows = set(A["rp"])
    act = [r for r in range(Bm) if any(c in rows for c in Bi[Bp[r]:Bp[r + 1]])]
    return pd(Bm, A["n"], act, A["cp"])


def ba_pd_pd(B, A):  # permuted_dense_linalg.c:458-480
    if not overlap(B["cp"], A["rp"]):
        return pd(B["m"], A["n"], [], [])
    return pd(B["m"], A["n"], B["rp"], A["cp"])


def ba_pd_spd(B, A):  # stacked_pd_linalg.c:608-625 (BTA_pd_spd on B^T)
    cps = [a["cp"] for a in blocks(A) if overlap(B["cp"], a["rp"])]
    return pd(B["m"], A["n"], B["rp"], union(cps))


def ba_pd_csr(B, Sp, Si, Sn):  # permuted_dense_linalg.c:313-375
    hit = set()
    for r in B["cp"]:
        hit.update(Si[Sp[r]:Sp[r + 1]])
    return pd(B["m"], Sn, B["rp"], sorted(hit))


def ata(A):
    if "blocks" not in A:  # permuted_dense_linalg.c:87-99
        return pd(A["n"], A["n"], A["cp"], A["cp"])
    raw = [pd(A["n"], A["n"], a["cp"], a["cp"]) for a in blocks(A)]
    return spd(A["n"], A["n"], coalesce(A["n"], A["n"], raw))


def bta(A, B, A_csr=None, B_csr=None):
    """Structure of BTA_matrices_alloc(A, B) (matmul_dispatchers.c:48-60)."""
    if B_csr is None and "blocks" not in B:  # B PD
        if A_csr is not None:
            return bta_pd_csc(B, *A_csr)
        if "blocks" in A:
            return bta_pd_spd(B, A)
        return bta_pd_pd(B, A)
    An = A_csr[2] if A_csr is not None else A["n"]
    if B_csr is None:  # B SPD: blockwise over B's blocks, coalesce
        raw = []
        for b in blocks(B):
            if A_csr is not None:
                raw.append(bta_pd_csc(b, *A_csr))
            elif "blocks" in A:
                raw.append(bta_pd_spd(b, A))
            else:
                raw.append(bta_pd_pd(b, A))
        return spd(B["n"], An, coalesce(B["n"], An, raw))
    Bp, Bi, Bn = B_csr  # B sparse
    if "blocks" not in A:
        return bta_csc_pd(Bp, Bi, Bn, A)
    raw = [bta_csc_pd(Bp, Bi, Bn, a) for a in blocks(A)]
    return spd(Bn, A["n"], coalesce(Bn, A["n"], raw))


def ba(B, A, B_csr=None, A_csr=None):
    """Structure of BA_matrices_alloc(B, A) (matmul_dispatchers.c:197-205):
    B permuted_dense or sparse (given as CSR (p, i, m)); A any kind (sparse A
    as CSR (p, i, n))."""
    if B_csr is None:
        if "blocks" in B:
            raise ValueError("a stacked_pd B is not supported")
        if A_csr is not None:
            return ba_pd_csr(B, *A_csr)
        if "blocks" in A:
            return ba_pd_spd(B, A)
        return ba_pd_pd(B, A)
    Bp, Bi, Bm = B_csr
    if "blocks" not in A:
        return ba_csc_pd(Bp, Bi, Bm, A)
    raw = [ba_csc_pd(Bp, Bi, Bm, a) for a in blocks(A)]
    return spd(Bm, A["n"], coalesce(Bm, A["n"], raw))


def transpose(A):
    if "blocks" not in A:
        return pd(A["n"], A["m"], A["cp"], A["rp"])
    raw = [pd(A["n"], A["m"], a["cp"], a["rp"]) for a in blocks(A)]
    return spd(A["n"], A["m"], coalesce(A["n"], A["m"], raw))


def with_values(S, Dn):
    """Fill a structure's blocks from a dense result."""
    for b in blocks(S):
        if len(b["rp"]) and len(b["cp"]):
            b["X"] = Dn[np.ix_(b["rp"], b["cp"])]
        else:
            b["X"] = np.zeros((len(b["rp"]), len(b["cp"])))
    return S


def tol_bound(absB, absD, absA, scale=1e-13):
    """Per-entry bound 1e-13 * (|B|^T |D| |A|) of a B^T D A product."""
    return scale * (absB.T @ (absD[:, None] * absA)) + 1e-300
