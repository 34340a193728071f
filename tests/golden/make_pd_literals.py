"""Writes pd_literals.json: known-answer cases of the permuted_dense /
stacked_pd kinds, transcribed from the reference's own tests (inputs and the
expected outputs those tests assert; each case names its source lines).

Where a reference test compares against a computed product rather than a
literal (a dense triple product or the D-then-BTA decomposition), the
expected values here are that product worked by hand; `main` re-derives every
expected block from a dense numpy product and refuses to write on a mismatch.

Run: python tests/golden/make_pd_literals.py
"""
import json
import os

import numpy as np


def pd(m, n, rp, cp, X):
    return {"kind": "pd", "m": m, "n": n, "row_perm": rp, "col_perm": cp, "X": X}


def spd(m, n, blocks):
    return {"kind": "spd", "m": m, "n": n, "blocks": blocks}


def csr(m, n, p, i, x):
    return {"kind": "sparse", "m": m, "n": n, "p": p, "i": i, "x": x}


def out(kind, m, n, blocks, src_p=None, src_idx=None):
    o = {"kind": kind, "m": m, "n": n, "blocks": blocks}
    if src_p is not None:
        o["src_p"] = src_p
        o["src_idx"] = src_idx
    return o


def blk(rp, cp, X):
    return {"row_perm": rp, "col_perm": cp, "X": X}


P = "tests/test_permuted_dense.h"
S = "tests/test_stacked_pd.h"
X6 = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]

CASES = [
    # ---- permuted_dense ------------------------------------------------
    dict(name="DA_pd", source=P + ":136-162", op="DA",
         A=pd(5, 6, [1, 2, 4], [0, 3], X6), d=[7.0, -1.5, 0.0, 9.0, 2.5],
         C=out("pd", 5, 6, [blk([1, 2, 4], [0, 3], [-1.5, -3.0, 0.0, 0.0, 12.5, 15.0])])),
    dict(name="ATDA_pd", source=P + ":192-222", op="ATDA",
         A=pd(5, 6, [1, 2, 4], [0, 3], X6), d=[7.0, -1.5, 0.0, 9.0, 2.5],
         C=out("pd", 6, 6, [blk([0, 3], [0, 3], [61.0, 72.0, 72.0, 84.0])])),
    dict(name="BA_pd_csr", source=P + ":217-263", op="BA",
         B=pd(5, 6, [1, 2, 4], [0, 3], X6),
         A=csr(6, 4, [0, 1, 1, 2, 4, 4, 4], [1, 2, 1, 3], [10.0, 30.0, 20.0, 40.0]),
         C=out("pd", 5, 4, [blk([1, 2, 4], [1, 3], [50.0, 80.0, 110.0, 160.0, 170.0, 240.0])])),
    dict(name="BA_pd_csr_no_active", source=P + ":265-298", op="BA",
         B=pd(5, 6, [1, 2, 4], [0, 3], X6),
         A=csr(6, 2, [0, 0, 1, 1, 1, 1, 2], [0, 1], [100.0, 200.0]),
         C=out("pd", 5, 2, [blk([1, 2, 4], [], [])])),
    dict(name="BTA_pd_pd_matching_rows", source=P + ":469-507", op="BTA",
         A=pd(4, 4, [1, 3], [0, 2], [1.0, 2.0, 3.0, 4.0]),
         B=pd(4, 4, [1, 3], [1, 3], [5.0, 6.0, 7.0, 8.0]),
         C=out("pd", 4, 4, [blk([1, 3], [0, 2], [26.0, 38.0, 30.0, 44.0])])),
    dict(name="BTA_pd_pd_empty_overlap", source=P + ":509-542", op="BTDA",
         A=pd(4, 4, [0, 2], [0, 2], [1.0, 2.0, 3.0, 4.0]),
         B=pd(4, 4, [1, 3], [1, 3], [5.0, 6.0, 7.0, 8.0]), d=[1.0, 1.0, 1.0, 1.0],
         C=out("pd", 4, 4, [blk([], [], [])])),
    dict(name="BTA_pd_pd_partial_overlap", source=P + ":544-610", op="BTA",
         A=pd(8, 4, [1, 3, 5], [0, 2], X6),
         B=pd(8, 4, [3, 5, 7], [1, 3], [10.0, 20.0, 30.0, 40.0, 50.0, 60.0]),
         # common rows {3, 5}: [[10,30],[20,40]] @ [[3,4],[5,6]]
         C=out("pd", 4, 4, [blk([1, 3], [0, 2], [180.0, 220.0, 260.0, 320.0])])),
    dict(name="BTDA_pd_pd_decomposition", source=P + ":612-675", op="BTDA",
         A=pd(3, 4, [0, 1, 2], [0, 2], X6),
         B=pd(3, 4, [0, 1, 2], [1, 3], [7.0, 8.0, 9.0, 10.0, 11.0, 12.0]),
         d=[2.0, -1.0, 3.0],
         C=out("pd", 4, 4, [blk([1, 3], [0, 2], [152.0, 190.0, 166.0, 208.0])])),
    dict(name="BTDA_pd_pd_matching_rows", source=P + ":677-713", op="BTDA",
         A=pd(4, 4, [1, 3], [0, 2], [1.0, 2.0, 3.0, 4.0]),
         B=pd(4, 4, [1, 3], [1, 3], [5.0, 6.0, 7.0, 8.0]), d=[-2.0, 0.5, 100.0, 3.0],
         C=out("pd", 4, 4, [blk([1, 3], [0, 2], [65.5, 89.0, 75.0, 102.0])])),
    dict(name="BTDA_pd_pd_partial_overlap", source=P + ":715-754", op="BTDA",
         A=pd(8, 4, [1, 3, 5], [0, 2], X6),
         B=pd(8, 4, [3, 5, 7], [1, 3], [10.0, 20.0, 30.0, 40.0, 50.0, 60.0]),
         d=[9.0, 9.0, 9.0, 2.0, 9.0, -1.5, 9.0, 9.0],
         C=out("pd", 4, 4, [blk([1, 3], [0, 2], [-165.0, -190.0, -180.0, -200.0])])),
    dict(name="BA_pd_pd_full_block_B", source=P + ":756-794", op="BA",
         B=pd(2, 3, [0, 1], [0, 1, 2], X6),
         A=pd(3, 5, [0, 2], [1, 4], [7.0, 8.0, 9.0, 10.0]),
         C=out("pd", 2, 5, [blk([0, 1], [1, 4], [34.0, 38.0, 82.0, 92.0])])),
    dict(name="BA_pd_pd_general_B", source=P + ":796-830", op="BA",
         B=pd(2, 5, [0, 1], [1, 3], [1.0, 2.0, 3.0, 4.0]),
         A=pd(5, 4, [1, 2], [0, 3], [5.0, 6.0, 7.0, 8.0]),
         C=out("pd", 2, 4, [blk([0, 1], [0, 3], [5.0, 6.0, 15.0, 18.0])])),
    dict(name="BA_pd_csc", source=P + ":832-872", op="BA",
         B=pd(2, 5, [0, 1], [1, 3], [1.0, 2.0, 3.0, 4.0]),
         A=csr(5, 4, [0, 0, 2, 4, 4, 4], [0, 3, 0, 3], [5.0, 6.0, 7.0, 8.0]),
         C=out("pd", 2, 4, [blk([0, 1], [0, 3], [5.0, 6.0, 15.0, 18.0])])),
    dict(name="BA_pd_spd", source=P + ":874-920", op="BA",
         B=pd(3, 4, [0, 1, 2], [0, 1, 3], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0]),
         A=spd(4, 6, [pd(4, 6, [0, 1], [0, 4], [10.0, 11.0, 12.0, 13.0]),
                      pd(4, 6, [2, 3], [1, 4], [20.0, 21.0, 22.0, 23.0])]),
         C=out("pd", 3, 6, [blk([0, 1, 2], [0, 1, 4],
                                [34.0, 66.0, 106.0, 100.0, 132.0, 247.0, 166.0, 198.0, 388.0])])),
    dict(name="BA_pd_pd_fast_path", source=P + ":922-953", op="BA",
         B=pd(2, 4, [0, 1], [1, 3], [1.0, 2.0, 3.0, 4.0]),
         A=pd(4, 3, [1, 3], [0, 2], [5.0, 6.0, 7.0, 8.0]),
         C=out("pd", 2, 3, [blk([0, 1], [0, 2], [19.0, 22.0, 43.0, 50.0])])),
    # ---- stacked_pd ----------------------------------------------------
    dict(name="transpose_spd_overlap_coalesces", source=S + ":350-418", op="transpose",
         A=spd(4, 5, [pd(4, 5, [0, 1], [2, 4], [1.0, 2.0, 3.0, 4.0]),
                      pd(4, 5, [3], [0, 2], [10.0, 11.0])]),
         C=out("spd", 5, 4, [blk([0], [3], [10.0]), blk([2], [0, 1, 3], [1.0, 3.0, 11.0]),
                             blk([4], [0, 1], [2.0, 4.0])],
               src_p=[0, 1, 3, 4], src_idx=[1, 0, 1, 0])),
    dict(name="ATDA_spd_disjoint_cols", source=S + ":756-789", op="ATDA",
         A=spd(4, 4, [pd(4, 4, [0, 1], [0, 1], [1.0, 2.0, 3.0, 4.0]),
                      pd(4, 4, [2, 3], [2, 3], [5.0, 6.0, 7.0, 8.0])]),
         d=[1.0, 2.0, 3.0, 4.0],
         C=out("spd", 4, 4, [blk([0, 1], [0, 1], [19.0, 26.0, 26.0, 36.0]),
                             blk([2, 3], [2, 3], [271.0, 314.0, 314.0, 364.0])],
               src_p=[0, 1, 2], src_idx=[0, 1])),
    dict(name="ATDA_spd_overlapping_cols", source=S + ":699-754,791-833", op="ATDA",
         A=spd(4, 3, [pd(4, 3, [0, 1], [0, 1], [1.0, 2.0, 3.0, 4.0]),
                      pd(4, 3, [2, 3], [1, 2], [5.0, 6.0, 7.0, 8.0])]),
         d=[1.0, 2.0, 3.0, 4.0],
         C=out("spd", 3, 3, [blk([0], [0, 1], [19.0, 26.0]),
                             blk([1], [0, 1, 2], [26.0, 307.0, 314.0]),
                             blk([2], [1, 2], [314.0, 364.0])],
               src_p=[0, 1, 3, 4], src_idx=[0, 0, 1, 1])),
    dict(name="BA_pd_spd_overlapping_col_perms", source=S + ":1055-1103", op="BA",
         B=pd(2, 3, [0, 1], [0, 1, 2], X6),
         A=spd(3, 3, [pd(3, 3, [0, 1], [0, 1], [10.0, 11.0, 12.0, 13.0]),
                      pd(3, 3, [2], [1, 2], [20.0, 21.0])]),
         C=out("pd", 2, 3, [blk([0, 1], [0, 1, 2], [34.0, 97.0, 63.0, 100.0, 229.0, 126.0])])),
]


# ------------------------------------------------------------ re-derive ----
def dense(M):
    D = np.zeros((M["m"], M["n"]))
    if M["kind"] == "sparse":
        for r in range(M["m"]):
            for q in range(M["p"][r], M["p"][r + 1]):
                D[r, M["i"][q]] = M["x"][q]
        return D
    for b in (M["blocks"] if M["kind"] == "spd" else [M]):
        X = np.asarray(b["X"], float).reshape(len(b["row_perm"]), len(b["col_perm"]))
        if X.size:
            D[np.ix_(b["row_perm"], b["col_perm"])] = X
    return D


def product(c):
    op = c["op"]
    A = dense(c["A"])
    if op == "DA":
        return np.asarray(c["d"])[:, None] * A
    if op == "ATDA":
        return A.T @ (np.asarray(c["d"])[:, None] * A)
    if op == "transpose":
        return A.T
    B = dense(c["B"])
    if op == "BTA":
        return B.T @ A
    if op == "BTDA":
        return B.T @ (np.asarray(c["d"])[:, None] * A)
    if op == "BA":
        return B @ A
    raise ValueError(op)


def main():
    for c in CASES:
        ref = product(c)
        got = np.zeros_like(ref)
        cover = np.zeros(ref.shape, bool)
        for b in c["C"]["blocks"]:
            X = np.asarray(b["X"], float).reshape(len(b["row_perm"]), len(b["col_perm"]))
            if X.size:
                got[np.ix_(b["row_perm"], b["col_perm"])] = X
                cover[np.ix_(b["row_perm"], b["col_perm"])] = True
        assert np.array_equal(got, ref), c["name"]
        assert not np.any(ref[~cover]), c["name"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pd_literals.json")
    with open(path, "w") as f:
        json.dump({"_about": "permuted_dense / stacked_pd known answers from the reference's "
                             "tests; written by make_pd_literals.py", "cases": CASES}, f, indent=1)
    print(f"wrote {len(CASES)} cases to {path}")


if __name__ == "__main__":
    main()
