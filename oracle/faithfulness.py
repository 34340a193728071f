#!/usr/bin/env python3
"""ORACLE (test infrastructure only): faithfulness record of the C
restatement (liboracle.so) against the reference's own sparse sources built
in this container (oracle/_ref/libref_sparse.so), on the 20000 x 5000
B^T diag(d) A product (6 distinct uniform columns per row, seed 1).

Both run single-threaded, pinned to one core, on the same inputs: the alloc
(CSC mirror + symbolic Gustavson + plan) and the fill for 8 fresh d vectors,
each timed 7 times after a warm-up (median reported).  Structure and every
value of every fill must be bit-identical; the record states the time ratio
port / reference for each phase.  The reference build never leaves this
container, so the record is committed as a fixture
(tests/golden/oracle_faithfulness.json) and the CPU baseline on other hosts
times the restatement.

Usage: python oracle/faithfulness.py [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden",
                                                  "oracle_faithfulness.json"))
    ap.add_argument("--reps", type=int, default=7)
    args = ap.parse_args()
    from oracle import cpu_env
    from oracle import oracle as O
    from sparsediff_linalg_amd import workloads

    if not O.ref_available():
        sys.exit("faithfulness: oracle/_ref/libref_sparse.so is not built here")
    w = workloads.btda_20000x5000(1)
    ds = [workloads.uniform(50_000 + r, w["m"], 0.1, 10.0) for r in range(8)]
    args_w = (w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"], w["Bp"], w["Bi"], w["Bx"])

    # bit-identity: structure and values of every fill
    same = True
    for d in ds:
        Rp, Ri, Rx = O.ref_btda(*args_w, d)
        Pp, Pi, Px, _ = O.btda(*args_w, d)
        same &= (np.array_equal(np.asarray(Rp, np.int64), np.asarray(Pp, np.int64))
                 and np.array_equal(Ri, Pi)
                 and np.array_equal(Rx.view(np.int64), Px.view(np.int64)))
    nnz = int(len(Ri))

    res = {}
    with cpu_env.pinned() as facts:
        for kind, timer in (("reference", O.ref_fill_timer), ("port", O.port_fill_timer)):
            timer(*args_w, ds[:2])  # warm-up
            allocs, fills = [], []
            for _ in range(args.reps):
                out = timer(*args_w, ds)
                allocs.append(out[0])
                fills.extend(out[1])
            res[kind] = {"alloc_s_median": statistics.median(allocs),
                         "fill_s_median": statistics.median(fills),
                         "alloc_runs": len(allocs), "fill_runs": len(fills)}
    rec = {
        "what": ("C restatement (oracle/liboracle.so) vs the reference's own sparse sources "
                 "(oracle/_ref/libref_sparse.so, gcc -O3 -DNDEBUG), both single-thread"),
        "workload": ("C = B^T diag(d) A, A and B 20000x5000 CSR, 6 distinct uniform columns per "
                     "row, FP64 N(0,1), seed 1; 8 fresh d vectors U[0.1,10]"),
        "nnz_out": nnz,
        "bit_identical": bool(same),
        "reference": res["reference"],
        "port": res["port"],
        "ratio_alloc": res["port"]["alloc_s_median"] / res["reference"]["alloc_s_median"],
        "ratio_fill": res["port"]["fill_s_median"] / res["reference"]["fill_s_median"],
        "host": facts,
        "date": time.strftime("%Y-%m-%d"),
        "command": "python oracle/faithfulness.py",
    }
    with open(args.out, "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
