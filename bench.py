#!/usr/bin/env python3
"""Benchmark of the B^T diag(d) A refill hot path on one MI355X.

Workload: the B^T.D.A product of two 20000 x 5000 CSR matrices with exactly 6
distinct uniform columns per row (int32 indices, FP64 N(0,1) values, seed 1)
and d uniform [0.1, 10]; output C is 5000 x 5000 CSR, sorted and coalesced.
One step = one batched refill of C for K fresh d vectors (one symbolic pass,
shared by every step, is done before timing).  The step runs the library's
whole refill sequence: operand gather, d transposition and the refill
kernel, all device-resident.

Reported beside it:
  * the plain per-call GPU path: the same plan refilled one d at a time
    (one launch per d), device-resident;
  * the reference CPU path: the reference's own sparse sources
    (oracle/_ref, built from the reference tree) or the C restatement,
    single thread, on a bounded sample of refills.

Prints ONE JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("B^T·D·A nnz-out/s + % HBM peak; batched+predicted vs plain GPU at 1e-13 rel")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--K", type=int, default=64, help="d vectors per batched refill")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="approximate CPU time of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--lean", action="store_true",
                    help="counter-collection mode: pinned refill configuration, no autotune, "
                         "no plain path, no spot check")
    ap.add_argument("--rounds", type=int, default=50,
                    help="multi: fresh d vectors per instance (refill rounds in the graph)")
    ap.add_argument("--instances", type=int, default=4096, help="multi: number of instances")
    ap.add_argument("--multi-layout", default="tile", choices=["tile", "pairs"],
                    help="multi: bound values in tile layout or as operand pairs")
    ap.add_argument("--multi-k", type=int, default=8,
                    help="multi: refill rounds per launch (even; output ring of this many slots)")
    ap.add_argument("--workload", default="btda",
                    choices=["btda", "blocks-atda", "blocks-btda", "multi", "ata-large", "coo",
                             "learned", "sum", "dense-kron", "lower-coo", "write-ceiling",
                             "hessian", "tune", "mfma-ceiling"],
                    help="btda: the sparse refill (default line); blocks-*: the stacked "
                         "permuted-dense products (A^T D A, or B^T D A of two stacks)")
    ap.add_argument("--rmat-scale", type=int, default=20,
                    help="learned: R-MAT scale of the A^T A evaluated beside the banded one")
    ap.add_argument("--epochs", type=int, default=100, help="learned: training epochs")
    ap.add_argument("--coalesce-lists", action="store_true",
                    help="blocks: coalesce the partials through per-cell pull lists "
                         "instead of by output row")
    ap.add_argument("--gemm-slice", type=int, default=0,
                    help="blocks-*: reduction slice depth of the block GEMM (16 or 32; 0 = "
                         "library default)")
    ap.add_argument("--rank-refill", action="store_true",
                    help="ata-large: plan with SDL_PLAN_RANK_REFILL (rows of more than 1024 "
                         "entries re-ranked at every refill instead of recorded positions)")
    ap.add_argument("--sweep", action="store_true",
                    help="btda: also time every refill launch configuration (stderr table)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run the streaming copy kernel on this many MiB (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def pinned_cpu(fn, *a, pin=True, **kw):
    """Runs a CPU-baseline timing, pinned to one core when the baseline is
    single-threaded, and adds the host facts (CPU model, nproc, pinning) to the
    record it returns."""
    from oracle import cpu_env
    if not pin:
        rec = fn(*a, **kw)
        rec.update({"cpu_model": cpu_env.cpu_model(), "nproc": cpu_env.nproc(),
                    "pinning": "none (multi-threaded)"})
        return rec
    with cpu_env.pinned() as facts:
        rec = fn(*a, **kw)
    rec.update(facts)
    return rec


def measured_traffic(name):
    """HBM bytes per launch measured by rocprofv3 counters (FETCH_SIZE and
    WRITE_SIZE runs, scaled by a streaming-copy calibration), from the summary
    kept under profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the sparse product is sharded by rows/instances in the design but no multi-GPU "
              "path is built.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "multi":
        return main_multi(args)
    if args.workload == "ata-large":
        return main_ata_large(args)
    if args.workload == "coo":
        return main_coo(args)
    if args.workload == "learned":
        return main_learned(args)
    if args.workload == "sum":
        return main_sum(args)
    if args.workload == "dense-kron":
        return main_dense_kron(args)
    if args.workload == "lower-coo":
        return main_lower_coo(args)
    if args.workload == "write-ceiling":
        return main_write_ceiling(args)
    if args.workload == "tune":
        return main_tune(args)
    if args.workload == "mfma-ceiling":
        return main_mfma_ceiling(args)
    if args.workload == "hessian":
        return main_hessian(args)
    if args.workload != "btda":
        return main_blocks(args)
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, workloads
    from sparsediff_linalg_amd._lib import sdl_refill_cfg

    lib = sdl.load()
    K = args.K
    w = workloads.btda_20000x5000(1)
    m, nA, nB = w["m"], w["nA"], w["nB"]
    nnzA, nnzB = int(w["Ap"][-1]), int(w["Bp"][-1])

    dAp, dAi, dBp, dBi = (DeviceArray.from_numpy(w[k]) for k in ("Ap", "Ai", "Bp", "Bi"))
    dAx, dBx = DeviceArray.from_numpy(w["Ax"]), DeviceArray.from_numpy(w["Bx"])
    h = C.c_void_p()
    t0 = time.perf_counter()
    check(lib.sdl_btda_plan_create(m, nA, nB, dAp.ptr, dAi.ptr, dBp.ptr, dBi.ptr, 0, C.byref(h)),
          "plan")
    lib.sdl_synchronize()
    t_alloc = time.perf_counter() - t0
    nnz = lib.sdl_btda_plan_nnz(h)
    terms = lib.sdl_btda_plan_terms(h)
    check(lib.sdl_btda_plan_reserve(h, K), "reserve")

    # d pools (different d per step) and the output batch (rows padded to 64)
    NPOOL = 4
    pools = [DeviceArray.from_numpy(np.stack([workloads.uniform(10_000 + 1000 * p + r, m, 0.1, 10.0)
                                              for r in range(K)])) for p in range(NPOOL)]
    ldc = (nnz + 63) // 64 * 64
    dC = DeviceArray((K, ldc), np.float64)

    # autotune the refill launch shape on this plan (alloc-phase work)
    best = sdl_refill_cfg()
    best_us = C.c_float()
    if args.lean:
        best.variant, best.rchunk = 7, 8
        check(lib.sdl_btda_set_refill_cfg(h, C.byref(best)), "set cfg")
        args.plain_steps = 0
    else:
        check(lib.sdl_btda_autotune(h, dAx.ptr, dBx.ptr, pools[0].ptr, m, K, dC.ptr, ldc, 10,
                                    C.byref(best), C.byref(best_us)), "autotune")

    stream = lib.sdl_library_stream()
    sweep = refill_sweep(lib, h, dAx, dBx, pools[0], m, K, dC, ldc) if args.sweep else None
    if sweep is not None:
        check(lib.sdl_btda_set_refill_cfg(h, C.byref(best)), "set cfg")
    ev = [(lib.sdl_event_create(), lib.sdl_event_create()) for _ in range(min(args.steps, 50))]
    e_all0, e_all1 = lib.sdl_event_create(), lib.sdl_event_create()

    def step(i, evs=None):
        D = pools[i % NPOOL]
        a, b = evs if evs else (None, None)
        return lib.sdl_btda_refill_batched_ex(h, dAx.ptr, dBx.ptr, D.ptr, m, K, dC.ptr, ldc, 0,
                                              a, b, stream)

    for i in range(args.warmup):
        check(step(i))
    lib.sdl_synchronize()
    t_wall0 = time.perf_counter()
    lib.sdl_event_record(e_all0, stream)
    for i in range(args.steps):
        check(step(i))
    lib.sdl_event_record(e_all1, stream)
    lib.sdl_synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms_total = lib.sdl_event_elapsed_ms(e_all0, e_all1)
    ms_step = ms_total / args.steps
    # the refill kernel alone, bracketed by events on its stream (a separate
    # pass, so the event records do not sit in the timed steps)
    for i in range(len(ev)):
        check(step(i, ev[i]))
    lib.sdl_synchronize()
    main_ms = float(np.mean([lib.sdl_event_elapsed_ms(a, b) for a, b in ev]))

    # the same refills with the d vectors handed over k-major (SDL_BATCH_D_KMAJOR:
    # refill r of row k at D[k * ldk + r]), which skips the d transposition;
    # informative only -- value stays on the refill-major input above
    kmajor = None
    if not args.lean and best.variant >= 3:
        ldk = (K + 15) // 16 * 16
        pools_k = [DeviceArray.from_numpy(np.ascontiguousarray(
            np.pad(P.numpy().T, ((0, 0), (0, ldk - K))))) for P in pools]
        for i in range(args.warmup):
            check(lib.sdl_btda_refill_batched_ex(h, dAx.ptr, dBx.ptr, pools_k[i % NPOOL].ptr, ldk,
                                                 K, dC.ptr, ldc, 2, None, None, stream), "kmajor")
        lib.sdl_event_record(e_all0, stream)
        for i in range(args.steps):
            check(lib.sdl_btda_refill_batched_ex(h, dAx.ptr, dBx.ptr, pools_k[i % NPOOL].ptr, ldk,
                                                 K, dC.ptr, ldc, 2, None, None, stream), "kmajor")
        lib.sdl_event_record(e_all1, stream)
        lib.sdl_synchronize()
        ms_k = lib.sdl_event_elapsed_ms(e_all0, e_all1) / args.steps
        got = dC.numpy()[:, :nnz]
        for r in (0, K - 1):
            one = DeviceArray(max(nnz, 1), np.float64)
            D_r = pools[(args.steps - 1) % NPOOL]
            check(lib.sdl_btda_refill(h, dAx.ptr, dBx.ptr, D_r.offset(r * m), one.ptr, stream), "one")
            lib.sdl_synchronize()
            assert np.array_equal(one.numpy()[:nnz].view(np.int64), got[r].view(np.int64)), \
                "k-major batched refill differs from the single refill"
        kmajor = {"ms_per_step": ms_k, "nnz_out_per_s": K * nnz / (ms_k * 1e-3),
                  "note": ("d vectors given k-major in 4 rotating 10 MB pools: no "
                           "transposition, same values; every XCD's L2 re-reads the whole "
                           "d, from HBM for a cold pool, where the library's transposed "
                           "buffer is the same each step and stays in the MALL")}
        del pools_k
        # the refill-major steps ran last in the checks below: redo one
        check(step(len(ev) - 1), "refill")
        lib.sdl_synchronize()

    # correctness spot check of the last step against the single-refill path
    D_last = pools[(len(ev) - 1) % NPOOL]  # the event pass ran last
    out_batch = dC.numpy()[:, :nnz]
    one = DeviceArray(max(nnz, 1), np.float64)
    for r in (() if args.lean else (0, K - 1)):
        check(lib.sdl_btda_refill(h, dAx.ptr, dBx.ptr, D_last.offset(r * m), one.ptr, stream))
        lib.sdl_synchronize()
        assert np.array_equal(one.numpy()[:nnz].view(np.int64), out_batch[r].view(np.int64)), \
            "batched refill differs from the single refill"

    # plain per-call GPU path: one launch per d
    for i in range(2 if args.plain_steps > 0 else 0):
        for r in range(K):
            lib.sdl_btda_refill(h, dAx.ptr, dBx.ptr, pools[0].offset(r * m), dC.offset(r * ldc),
                                stream)
    lib.sdl_event_record(e_all0, stream)
    for i in range(args.plain_steps):
        D = pools[i % NPOOL]
        for r in range(K):
            lib.sdl_btda_refill(h, dAx.ptr, dBx.ptr, D.offset(r * m), dC.offset(r * ldc), stream)
    lib.sdl_event_record(e_all1, stream)
    lib.sdl_synchronize()
    plain_ms_batch = lib.sdl_event_elapsed_ms(e_all0, e_all1) / max(args.plain_steps, 1)
    # the same per-call path with the per-term-gather kernel (three indices and
    # three gathers per term) instead of the default tile refill
    plain_gather_ms = None
    if args.plain_steps > 0 and lib.sdl_btda_plan_single_refill(h, 0) == 1:
        lib.sdl_event_record(e_all0, stream)
        for i in range(args.plain_steps):
            D = pools[i % NPOOL]
            for r in range(K):
                lib.sdl_btda_refill(h, dAx.ptr, dBx.ptr, D.offset(r * m), dC.offset(r * ldc),
                                    stream)
        lib.sdl_event_record(e_all1, stream)
        lib.sdl_synchronize()
        plain_gather_ms = lib.sdl_event_elapsed_ms(e_all0, e_all1) / args.plain_steps
        lib.sdl_btda_plan_single_refill(h, 1)

    # write floor: the refill's store pattern over the same output rows with
    # constant values (no reads), the least time any refill of this batch can
    # take on this GPU; plain / floor bounds the batched-over-plain ratio
    floor_ms = None
    if not args.lean:
        for groups in (1, 8):
            for _ in range(3):
                check(lib.sdl_write_rows_kernel(dC.ptr, ldc, K, nnz, groups, stream), "floor")
            lib.sdl_event_record(e_all0, stream)
            for _ in range(20):
                check(lib.sdl_write_rows_kernel(dC.ptr, ldc, K, nnz, groups, stream), "floor")
            lib.sdl_event_record(e_all1, stream)
            lib.sdl_synchronize()
            t = lib.sdl_event_elapsed_ms(e_all0, e_all1) / 20
            floor_ms = t if floor_ms is None else min(floor_ms, t)

    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        src, dst = DeviceArray(nb // 8, np.float64), DeviceArray(nb // 8, np.float64)
        for _ in range(3):
            check(lib.sdl_copy_kernel(dst.ptr, src.ptr, nb, stream), "copy")
        lib.sdl_synchronize()

    # algorithmic bytes / flops (each array counted once)
    S = (4 * (nB + 1) + 12 * nnzB) + (4 * (m + 1) + 12 * nnzA) + (4 * (nB + 1) + 4 * nnz)
    bytes_K = S + K * (8 * m + 8 * nnz)
    flops_K = K * (nnzB + 2 * terms)
    value = K * nnz / (ms_step * 1e-3)
    achieved = bytes_K / (ms_step * 1e-3) / 1e9

    traffic = measured_traffic("btda_refill_20000x5000_traffic.json")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, nnz, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz-out/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("C = B^T diag(d) A, A and B 20000x5000 CSR, 6 distinct uniform "
                         "columns/row, FP64 N(0,1), d U[0.1,10], seed 1; one step = one batched "
                         f"refill of C for K={K} fresh d vectors"),
            "K": K, "nnz_out": nnz, "terms": terms, "m": m, "n_A": nA, "n_B": nB,
            "refill_cfg": {"variant": best.variant, "rchunk": best.rchunk},
            "symbolic_mode": "exact",
            "symbolic_note": ("the symbolic phase (exact or predicted) is the untimed alloc "
                              "phase; it yields the same structure and refill schedule, so the "
                              "timed refills are the same for either"),
            "wall_s": t_wall,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": refill_kernel_name(best),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_K,
            "main_kernel_us": main_ms * 1e3,
            "main_kernel_alg_GBs": bytes_K / (main_ms * 1e-3) / 1e9,
            "timing": ("achieved = algorithmic bytes / event time of the whole refill step "
                       "(preparation + refill kernel); main_kernel_us = events around the "
                       "refill kernel on its stream, in a second pass of the same steps"),
        },
        "gflops": flops_K / (ms_step * 1e-3) / 1e9,
        "plain_gpu": {
            "ms_per_K_refills": plain_ms_batch,
            "us_per_refill": plain_ms_batch * 1e3 / K,
            "nnz_out_per_s": K * nnz / (plain_ms_batch * 1e-3),
            "kernel": "k_num_pull_tile (one launch per d)",
            "per_term_gathers_us_per_refill": (None if plain_gather_ms is None
                                               else plain_gather_ms * 1e3 / K),
            "batched_vs_per_term_gathers": (None if plain_gather_ms is None
                                            else plain_gather_ms / ms_step),
        },
        "batched_vs_plain": plain_ms_batch / ms_step,
        "write_floor": None if floor_ms is None else {
            "us": floor_ms * 1e3, "GB_s": K * nnz * 8 / (floor_ms * 1e-3) / 1e9,
            "step_over_floor": ms_step / floor_ms,
            "max_batched_vs_plain": plain_ms_batch / floor_ms,
            "note": ("k_write_rows: the refill's K x nnz FP64 output stored in its own pattern "
                     "with constant values and no reads; a batched refill cannot beat it, so "
                     "plain / floor bounds batched_vs_plain on this GPU")},
        "d_kmajor": kmajor,
        "alloc_s": t_alloc,
        "cpu_baseline": cpu,
    }
    if sweep is not None:
        out["refill_sweep_us"] = sweep
    lib.sdl_btda_plan_destroy(h)
    print(json.dumps(out), flush=True)


def refill_kernel_name(cfg):
    if cfg.variant == 7:
        return "k_prep_batch (d transposition) + k_num_pull_tile_batched"
    if cfg.variant >= 3:
        return "k_prep_batch (operand pairs + d transposition) + k_num_pull_interleaved"
    return "k_num_pull_rchunk / k_num_pull_batched (direct gathers)"


def refill_sweep(lib, h, dAx, dBx, D, m, K, dC, ldc, reps=20):
    """Event time per batched refill of every candidate launch configuration
    (after two warm-ups); returns {"variant/rchunk": us}."""
    from sparsediff_linalg_amd import check
    from sparsediff_linalg_amd._lib import sdl_refill_cfg
    cands = (sdl_refill_cfg * 64)()
    nc = min(64, lib.sdl_refill_candidates(cands, 64))
    stream = lib.sdl_library_stream()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    res = {}
    for c in range(nc):
        cfg = cands[c]
        for _ in range(2):
            check(lib.sdl_btda_refill_batched_cfg(h, dAx.ptr, dBx.ptr, D.ptr, m, K, dC.ptr, ldc,
                                                  C.byref(cfg), stream), "sweep")
        lib.sdl_event_record(e0, stream)
        for _ in range(reps):
            check(lib.sdl_btda_refill_batched_cfg(h, dAx.ptr, dBx.ptr, D.ptr, m, K, dC.ptr, ldc,
                                                  C.byref(cfg), stream), "sweep")
        lib.sdl_event_record(e1, stream)
        lib.sdl_synchronize()
        us = lib.sdl_event_elapsed_ms(e0, e1) * 1e3 / reps
        key = f"{cfg.variant}/{cfg.rchunk}"
        res[key] = round(us, 2)
        print(f"sweep {key}: {us:.2f} us", file=sys.stderr, flush=True)
    return res


def main_hessian(args):
    """Hessian assembly across expression nodes, one Newton step per step: a
    synthetic Lagrangian over 200,000 variables with 12 chain-rule nodes
    (m_k log-uniform in [20,000, 120,000], 3-9 nnz per row in a node-specific
    window of 50,000 variables): every node's A_k^T diag(d_k) A_k refill, the
    ordered sum of the nodes and the lower-triangular export, issued call by
    call and replayed from one captured HIP graph.  The graph replay's output
    is checked whole against the oracle, bit for bit."""
    import sparsediff_linalg_amd as sdl
    from oracle import oracle as O
    from sparsediff_linalg_amd import DeviceArray, workloads
    from sparsediff_linalg_amd.hessian import Hessian

    lib = sdl.load()
    nodes, n = workloads.lagrangian_nodes(7)
    t0 = time.perf_counter()
    H = Hessian([(p, i) for p, i, _ in nodes], n)
    lib.sdl_synchronize()
    t_create = time.perf_counter() - t0
    rng = np.random.default_rng(5)
    ds = [workloads.uniform(600 + k, len(p) - 1, 0.1, 10.0) for k, (p, _, _) in enumerate(nodes)]
    Ax = [DeviceArray.from_numpy(x) for _, _, x in nodes]
    D = [DeviceArray.from_numpy(d) for d in ds]
    Hx = DeviceArray(max(H.nnz, 1), np.float64)
    Lx = DeviceArray(max(H.lower_nnz, 1), np.float64)
    stream = lib.sdl_library_stream()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(args.steps, 20)

    def timed(fn):
        for _ in range(max(args.warmup, 3)):
            fn()
        lib.sdl_synchronize()
        lib.sdl_event_record(e0, stream)
        for _ in range(steps):
            fn()
        lib.sdl_event_record(e1, stream)
        lib.sdl_synchronize()
        return lib.sdl_event_elapsed_ms(e0, e1) / steps

    # node refills in one launch over every node's tiles (0), or one launch
    # per node in order (1) or forked over 2 or 4 streams: eager and replayed;
    # the line reports the fastest replay (the choice is a setting)
    by_streams = {}
    for ns in (0, 1, 2, 4):
        H.set_streams(ns)
        me = timed(lambda: H.eval(Ax, D, Hx, Lx, stream))
        H.capture(Ax, D, Hx, Lx, stream)
        mg = timed(lambda: H.replay(stream))
        by_streams[ns] = {"eager_ms": me, "graph_ms": mg}
    best_ns = min(by_streams, key=lambda k: by_streams[k]["graph_ms"])
    H.set_streams(best_ns)
    ms_eager = timed(lambda: H.eval(Ax, D, Hx, Lx, stream))
    H.capture(Ax, D, Hx, Lx, stream)
    ms_graph = timed(lambda: H.replay(stream))
    # the last replay against the oracle, whole
    got_h, got_l = Hx.numpy()[:H.nnz], Lx.numpy()[:H.lower_nnz]
    from oracle import cpu_env
    ops = []
    with cpu_env.pinned() as facts:
        t0 = time.perf_counter()
        for (p, i, x), d in zip(nodes, ds):
            m = len(p) - 1
            Lp, Li, Lv = O.transpose(m, n, p, i, x)
            Cp, Ci, Cx, _ = O.spgemm(n, Lp, Li, Lv, 1, d, m, n, p, i, x)
            ops.append((Cp.astype(np.int32), Ci, Cx))
        Hp, Hi, Ox, _ = O.sum_n(n, n, ops)
        _, _, lx, _ = O.coo_lower(n, Hp, Hi, Ox)
        t_oracle = time.perf_counter() - t0
    exact = bool(np.array_equal(got_h.view(np.int64), Ox.view(np.int64)) and
                 np.array_equal(got_l.view(np.int64), lx.view(np.int64)))
    assert exact, "Hessian assembly differs from the oracle"
    terms = sum(int((np.diff(p).astype(np.int64) ** 2).sum()) for p, _, _ in nodes)
    node_nnz = sum(len(c[1]) for c in ops)
    # algorithmic bytes of one step: node refills (values, d, term lists read
    # once, outputs written), the sum (node outputs read, H written), the
    # export (H's lower values read and written)
    bytes_ = (sum(12 * len(i) + 8 * (len(p) - 1) for p, i, _ in nodes) + 12 * terms
              + 8 * node_nnz) + (8 * node_nnz + 8 * H.nnz) + 16 * H.lower_nnz
    out = {
        "metric": "Hessian assembly Newton steps/s (node refills + ordered sum + lower export)",
        "value": 1e3 / ms_graph, "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": max(args.warmup, 3), "ms_per_step": ms_graph, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("Lagrangian Hessian of 12 chain-rule nodes over 200,000 "
                                "variables (m_k in [20k, 120k], 3-9 nnz/row); one step = every "
                                "node's A^T D A refill, the ordered sum and the lower export"),
                   "nodes": len(nodes), "n": n, "terms": terms, "node_nnz": node_nnz,
                   "nnz_H": H.nnz, "lower_nnz": H.lower_nnz, "create_s": t_create,
                   "bit_exact_vs_oracle": exact},
        "eager_ms_per_step": ms_eager, "graph_vs_eager": ms_eager / ms_graph,
        "node_streams": best_ns, "by_node_streams": by_streams,
        "roofline": {"bound": "hbm", "kernel": "graph of the step's launches",
                     "achieved": bytes_ / (ms_graph * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": bytes_ / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_launch": bytes_},
        "cpu_baseline": None if args.no_cpu else {
            "value": 1.0 / t_oracle, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": (f"one whole step on the oracle, structure and values (per-node "
                       f"transpose + Gustavson, ordered sum, lower export) in {t_oracle:.2f} s"),
            **facts},
    }
    H.close()
    # a launch-bound DAG: 32 small nodes (m_k in [500, 4000], 3,000-variable
    # windows over 20,000 variables), bit-exact between graph and eager
    nodes_s, n_s = workloads.lagrangian_nodes(8, n=20_000, nnodes=32, m_lo=500, m_hi=4_000,
                                              window=3_000)
    Hs = Hessian([(p, i) for p, i, _ in nodes_s], n_s)
    Axs = [DeviceArray.from_numpy(x) for _, _, x in nodes_s]
    Ds = [DeviceArray.from_numpy(workloads.uniform(700 + k, len(p) - 1, 0.1, 10.0))
          for k, (p, _, _) in enumerate(nodes_s)]
    Hxs = DeviceArray(max(Hs.nnz, 1), np.float64)
    Lxs = DeviceArray(max(Hs.lower_nnz, 1), np.float64)
    ms_eager_s = timed(lambda: Hs.eval(Axs, Ds, Hxs, Lxs, stream))
    eager_vals = Hxs.numpy()
    Hs.capture(Axs, Ds, Hxs, Lxs, stream)
    ms_graph_s = timed(lambda: Hs.replay(stream))
    same = bool(np.array_equal(eager_vals.view(np.int64), Hxs.numpy().view(np.int64)))
    assert same, "graph replay differs from the eager evaluation"
    Hs.set_streams(4)
    Hs.capture(Axs, Ds, Hxs, Lxs, stream)
    ms_graph_s4 = timed(lambda: Hs.replay(stream))
    same = same and bool(np.array_equal(eager_vals.view(np.int64), Hxs.numpy().view(np.int64)))
    Hs.set_streams(1)
    Hs.capture(Axs, Ds, Hxs, Lxs, stream)
    ms_graph_s1 = timed(lambda: Hs.replay(stream))
    same = same and bool(np.array_equal(eager_vals.view(np.int64), Hxs.numpy().view(np.int64)))
    assert same, "one-stream graph replay differs from the eager evaluation"
    out["small_dag"] = {"workload": ("32 chain-rule nodes over 20,000 variables, m_k in "
                                     "[500, 4000]"), "nnz_H": Hs.nnz,
                        "eager_ms_per_step": ms_eager_s, "graph_ms_per_step": ms_graph_s,
                        "graph_vs_eager": ms_eager_s / ms_graph_s, "graph_equals_eager": same,
                        "graph_one_stream_ms_per_step": ms_graph_s1,
                        "graph_four_streams_ms_per_step": ms_graph_s4,
                        "node_refills": "one launch over every node's tiles (16 nodes per launch)"}
    Hs.close()
    print(json.dumps(out), flush=True)


def main_tune(args):
    """Tuned vs built-in plan thresholds (the symbolic phase's largest LDS hash
    table and bitmap density factor, the row-accumulator refill's row cuts) on
    the A^T A of the rank-Zipf 2,000,000 x 500,000 matrix, of the banded 2^20
    matrix (bandwidth 64) and of the R-MAT 2^20 graph.  Per product the tuner
    runs once (sdl_btda_tune: plan creation over 20 symbolic candidates, then
    refills over 9 row-cut candidates where values are given; a candidate must
    be 3 % cheaper than the incumbent), then plans are created afresh with the
    built-in and the recorded thresholds, interleaved three times (least), and
    timed: plan creation from device-resident structure (host clock around
    device syncs) plus one refill for a fresh d (device events).  The R-MAT product (9.7e9 output
    entries) is tuned and timed on plan creation alone."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, tuning, workloads
    from sparsediff_linalg_amd.predictor import Plan

    lib = sdl.load()
    stream = lib.sdl_library_stream()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()

    def measure(p, i, n, t, dAx, dd, dev):
        from sparsediff_linalg_amd.tuning import sym_tune
        h = C.c_void_p()
        lib.sdl_synchronize()
        t0 = time.perf_counter()
        check(lib.sdl_btda_plan_create_tuned(len(p) - 1, n, n, dev[0].ptr, dev[1].ptr, dev[0].ptr,
                                             dev[1].ptr, 0, C.byref(sym_tune(t)), C.byref(h)),
              "plan")
        lib.sdl_synchronize()
        plan_ms = (time.perf_counter() - t0) * 1e3
        pl = Plan.__new__(Plan)   # adopt the handle: rowptr / nnz / schedule readers
        pl.h, pl.lib, pl.nA, pl.nB = h, lib, n, n
        nnz = pl.nnz()
        rowptr = np.empty(n + 1, np.int64)
        lib.sdl_memcpy_dtoh(rowptr.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), rowptr.nbytes)
        ref_ms = 0.0
        if dAx is not None:
            Cx = DeviceArray(max(nnz, 1), np.float64)
            check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr, Cx.ptr, stream), "refill")
            lib.sdl_event_record(e0, stream)
            check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, dd.ptr, Cx.ptr, stream), "refill")
            lib.sdl_event_record(e1, stream)
            lib.sdl_synchronize()
            ref_ms = lib.sdl_event_elapsed_ms(e0, e1)
            del Cx
        schedule = "pull" if lib.sdl_btda_plan_is_pull(pl.h) else "row accumulators"
        pl.close()
        return plan_ms, ref_ms, nnz, hash(rowptr.tobytes()), schedule

    def zipf():
        p, i, x = workloads.zipf_csr(2, 2_000_000, 500_000)
        return p, i, x, 500_000

    def banded():
        p, i, x = workloads.banded_csr(1 << 20, 64)
        return p, i, x, 1 << 20

    def rmat():
        p, i, x = workloads.rmat_csr(args.rmat_scale)
        return p, i, None, 1 << args.rmat_scale

    products = [("A^T A, A 2,000,000 x 500,000 rank-Zipf(1.2) rows, mean 16", zipf),
                ("A^T A, banded 2^20 x 2^20, bandwidth 64", banded),
                (f"A^T A, R-MAT 2^{args.rmat_scale} nodes, 16 edges/node", rmat)]
    results = []
    tot_def = tot_tuned = 0.0
    for label, make in products:
        p, i, x, n = make()
        m = len(p) - 1
        dAx = DeviceArray.from_numpy(x) if x is not None else None
        dd = DeviceArray.from_numpy(workloads.uniform(77, m, 0.1, 10.0))
        tuning.clear()
        Plan((p, i), None, n, n).close()  # warm-up: kernels, allocator
        t0 = time.perf_counter()
        best, t_best, t_def = tuning.tune((p, i), None, n, n,
                                          values=(dAx, dAx) if dAx is not None else None,
                                          reps=3)
        tune_s = time.perf_counter() - t0
        print(f"[tune] {label}: tuner {tune_s:.1f} s -> {best}", file=sys.stderr, flush=True)
        default = tuning.default()
        same = best == default
        arms = (("built-in", default),) if same else (("built-in", default), ("tuned", best))
        runs = {"built-in": [], "tuned": []}
        dev = [DeviceArray.from_numpy(np.ascontiguousarray(a, np.int32)) for a in (p, i)]
        for _ in range(3):
            for key, t in arms:
                runs[key].append(measure(p, i, n, t, dAx, dd, dev))
        if same:
            # the tuner kept the built-in thresholds: one arm, no ratio of noise
            runs["tuned"] = runs["built-in"]
        # least of 3 interleaved runs (allocation stalls only add time)
        d_best = min(runs["built-in"], key=lambda r: r[0] + r[1])
        t_best_run = min(runs["tuned"], key=lambda r: r[0] + r[1])
        assert d_best[2] == t_best_run[2] and d_best[3] == t_best_run[3], label
        td, tt = d_best[0] + d_best[1], t_best_run[0] + t_best_run[1]
        rec = {"product": label, "nnz_out": d_best[2], "schedule": d_best[4],
               "tuned": best, "tuner_s": tune_s,
               "tuner_estimate_ms": {"tuned": t_best, "built_in": t_def},
               "built_in_ms": {"plan": d_best[0], "refill": d_best[1], "total": td},
               "tuned_ms": {"plan": t_best_run[0], "refill": t_best_run[1], "total": tt},
               "gain": td / tt, "structure_same": True,
               "tuner_kept_built_in": same}
        tot_def += td
        tot_tuned += tt
        results.append(rec)
        print(f"[tune] {json.dumps(rec)}", file=sys.stderr, flush=True)
        del dAx, dd
    tuning.clear()
    out = {
        "metric": "plan + refill time, built-in / tuned thresholds (tuning table)",
        "value": tot_def / tot_tuned, "unit": "x", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": tot_tuned, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32+f64", "data": "synthetic",
        "config": {"workload": ("plan creation + one refill of three A^T A products "
                                "(rank-Zipf 2M x 500k, banded 2^20 bw 64, R-MAT plan only), "
                                "tuned vs built-in thresholds, least of 3 interleaved"),
                   "products": results},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def mfma_ceiling(lib, iters=4000):
    """Sustained FP64 matrix rate of this GPU: register-only loops of
    v_mfma_f64_16x16x4_f64 (sdl_mfma_kernel), 16 workgroups of 4 wavefronts
    per CU, 1 to 16 independent accumulators per wavefront.  Returns the best
    TFLOP/s and the table."""
    from sparsediff_linalg_amd import DeviceArray, check
    blocks = 256 * 16
    out = DeviceArray(blocks * 256, np.float64)
    stream = lib.sdl_library_stream()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    table = {}
    for nacc in (1, 2, 4, 8, 16):
        it = max(1, iters * 4 // nacc)
        check(lib.sdl_mfma_kernel(nacc, blocks, it, out.ptr, stream), "mfma")
        lib.sdl_event_record(e0, stream)
        check(lib.sdl_mfma_kernel(nacc, blocks, it, out.ptr, stream), "mfma")
        lib.sdl_event_record(e1, stream)
        lib.sdl_synchronize()
        ms = lib.sdl_event_elapsed_ms(e0, e1)
        table[nacc] = blocks * 4 * it * nacc * 2048.0 / (ms * 1e-3) / 1e12
    return max(table.values()), table


def main_mfma_ceiling(args):
    """FP64 matrix-rate calibration (the ceiling a block-product GEMM can
    reach on this GPU) against the published dense FP64 MFMA peak."""
    import sparsediff_linalg_amd as sdl
    lib = sdl.load()
    best, table = mfma_ceiling(lib)
    out = {
        "metric": "sustained FP64 MFMA TFLOP/s (register-only v_mfma_f64_16x16x4_f64 loops)",
        "value": best, "unit": "TFLOP/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("4096 workgroups x 4 wavefronts, each wavefront 1-16 "
                                "independent 16x16 accumulators, no memory traffic"),
                   "tflops_by_accumulators": table},
        "roofline": {"bound": "mfma", "kernel": "k_mfma_f64<NACC>", "achieved": best,
                     "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                     "frac": best / FP64_MFMA_PEAK_TFS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main_write_ceiling(args):
    """Streaming-write calibration on the output size of the default line
    (K = 64 refills of the 20000 x 5000 product: 64 rows of 709,888 doubles,
    363 MB): constant non-temporal 16-byte and 8-byte stores, ordinary 16-byte
    stores, a refill's store pattern (all 64 rows per thread, or 8), and the
    streaming copy of the same bytes.  The best write rate is the ceiling a
    write-bound refill can reach on this GPU."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check
    lib = sdl.load()
    K, nnz = args.K, 709_850
    ldc = (nnz + 63) // 64 * 64
    nbytes = K * ldc * 8
    dst = DeviceArray((K, ldc), np.float64)
    src = DeviceArray((K, ldc), np.float64)
    stream = lib.sdl_library_stream()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(args.steps, 10)

    def timed(fn):
        for _ in range(3):
            check(fn(), "write")
        lib.sdl_event_record(e0, stream)
        for _ in range(steps):
            check(fn(), "write")
        lib.sdl_event_record(e1, stream)
        lib.sdl_synchronize()
        return lib.sdl_event_elapsed_ms(e0, e1) / steps

    rates = {}
    for name, fn in (
            ("store16_nt", lambda: lib.sdl_write_kernel(dst.ptr, nbytes, 0, stream)),
            ("store16", lambda: lib.sdl_write_kernel(dst.ptr, nbytes, 1, stream)),
            ("store8_nt", lambda: lib.sdl_write_kernel(dst.ptr, nbytes, 2, stream)),
            ("refill_rows_g1", lambda: lib.sdl_write_rows_kernel(dst.ptr, ldc, K, nnz, 1, stream)),
            ("refill_rows_g8", lambda: lib.sdl_write_rows_kernel(dst.ptr, ldc, K, nnz, 8, stream))):
        ms = timed(fn)
        b = nbytes if not name.startswith("refill") else K * nnz * 8
        rates[name] = {"ms": ms, "GB_s": b / (ms * 1e-3) / 1e9}
    ms = timed(lambda: lib.sdl_copy_kernel(dst.ptr, src.ptr, nbytes, stream))
    rates["copy"] = {"ms": ms, "GB_s": 2 * nbytes / (ms * 1e-3) / 1e9}
    best = max(v["GB_s"] for k, v in rates.items() if k != "copy")
    out = {
        "metric": "streaming write GB/s (calibration of the write-bound refill)",
        "value": best, "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 3,
        "ms_per_step": min(v["ms"] for k, v in rates.items() if k != "copy"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"constant stores over {K} x {ldc} doubles ({nbytes / 1e6:.1f} MB),"
                                " the output of the default batched refill"), "K": K},
        "rates": rates,
        "roofline": {"bound": "hbm", "kernel": "k_stream_write16 / k_write_rows",
                     "achieved": best, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": best / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)



def main_ata_large(args):
    """C = A^T diag(d) A for A 2,000,000 x 500,000 with rank-Zipf(1.2) row
    lengths in [1, 512] (mean 16), uniform columns, FP64 N(0,1): one symbolic
    pass (int64 plan; nnz(C) beyond INT_MAX) and refills for fresh d."""
    import scipy.sparse as sp

    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, workloads
    from sparsediff_linalg_amd.predictor import Plan

    lib = sdl.load()
    m, n = 2_000_000, 500_000
    Ap, Ai, Ax = workloads.zipf_csr(2, m, n)
    nnzA = int(Ap[-1])
    t0 = time.perf_counter()
    pl = Plan((Ap, Ai), None, n, n, mode=512 if args.rank_refill else 0)
    lib.sdl_synchronize()
    t_sym = time.perf_counter() - t0
    nnz, terms = pl.nnz(), lib.sdl_btda_plan_terms(pl.h)
    sched = pl.schedule()
    print(f"[ata-large] plan {t_sym:.3f} s, schedule {sched}", file=sys.stderr, flush=True)
    dAx = DeviceArray.from_numpy(Ax)
    ds = [DeviceArray.from_numpy(workloads.uniform(90 + i, m, 0.1, 10.0)) for i in range(2)]
    Cx = DeviceArray(nnz, np.float64)
    stream = lib.sdl_library_stream()
    check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, ds[0].ptr, Cx.ptr, stream), "refill")
    lib.sdl_synchronize()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(1, min(args.steps, 5))
    lib.sdl_event_record(e0, stream)
    for i in range(steps):
        check(lib.sdl_btda_refill(pl.h, dAx.ptr, dAx.ptr, ds[i % 2].ptr, Cx.ptr, stream), "refill")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms = lib.sdl_event_elapsed_ms(e0, e1) / steps
    # sanity: 32 rows against scipy (different summation order: 1e-13 relative)
    d_last = ds[(steps - 1) % 2].numpy()
    A = sp.csr_matrix((Ax, Ai, Ap), shape=(m, n))
    cp = np.empty(n + 1, np.int64)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), cp.nbytes)
    AT = A.T.tocsr()
    worst = 0.0
    for r in np.random.default_rng(0).choice(n, 32, replace=False):
        row = (AT[r].multiply(d_last[None, :])) @ A
        row = row.tocsr()
        a, b = int(cp[r]), int(cp[r + 1])
        gx = np.empty(b - a)
        lib.sdl_memcpy_dtoh(gx.ctypes.data, Cx.offset(a), gx.nbytes)
        dense = np.zeros(n)
        dense[row.indices] = row.data
        gi = np.empty(b - a, np.int32)
        lib.sdl_memcpy_dtoh(gi.ctypes.data, lib.sdl_btda_plan_colind(pl.h) + 4 * a, gi.nbytes)
        mag = np.abs(AT[r].multiply(d_last[None, :])) @ np.abs(A)
        magd = np.zeros(n)
        magd[mag.indices] = mag.data
        err = np.abs(gx - dense[gi]) / (magd[gi] + 1e-300)
        worst = max(worst, float(err.max()) if len(err) else 0.0)
    assert worst <= 1e-13, worst
    bytes_ = (4 * (m + 1) + 12 * nnzA) + 8 * m + (4 * (n + 1) + 12 * nnz)
    t = ms * 1e-3
    out = {
        "metric": METRIC + " (large A^T A)",
        "value": nnz / t, "unit": "nnz-out/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": ("C = A^T diag(d) A, A 2,000,000 x 500,000 CSR, rank-Zipf(1.2) row "
                         "lengths clipped [1,512] mean 16, uniform columns, FP64 N(0,1), seed 2; "
                         "one step = one refill for a fresh d"),
            "nnz_A": nnzA, "terms": terms, "nnz_out": nnz, "symbolic_s": t_sym,
            "schedule": "pull" if lib.sdl_btda_plan_is_pull(pl.h) else "row accumulators",
            "refill_schedule": sched,
            "sampled_max_rel_err_vs_scipy": worst,
        },
        "roofline": {"bound": "hbm",
                     "kernel": ("k_num_rank_pos<1024, 256, 8> (recorded positions, rows "
                                "staged in LDS) + k_num_rowacc_dense spills"
                                if sched["rank_positions"] else
                                "k_num_rowacc_rank<1024> (rows by rank) + k_num_rowacc_dense "
                                "spills"),
                     "achieved": bytes_ / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_ / t / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_num_rank_pos_ata_large_traffic.json"
                                                 if sched["rank_positions"] else
                                                 "k_num_rowacc_rank_ata_large_traffic.json"),
                     "alg_bytes_per_launch": bytes_},
        "gflops": (nnzA + 2 * terms) / t / 1e9,
        "symbolic_terms_per_s": terms / t_sym,
        "cpu_baseline": None if args.no_cpu else pinned_cpu(ata_cpu_baseline, m, n, Ap, Ai, Ax,
                                                             d_last, args.cpu_seconds),
    }
    pl.close()
    print(json.dumps(out), flush=True)


def ata_cpu_baseline(m, n, Ap, Ai, Ax, d, seconds):
    """The oracle's single-thread row-at-a-time Gustavson (structure and values
    of each row: mark-array count, sort, accumulator fill) over random output
    rows until about `seconds` of CPU time; nnz-out/s of those rows."""
    from oracle import oracle as O
    Lp, Li, Lx = O.transpose(m, n, Ap, Ai, Ax)
    rng = np.random.default_rng(11)
    done = nnz = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        rows = rng.choice(n, 200, replace=False)
        for cols, _ in O.spgemm_rows(rows, n, Lp, Li, Lx, 1, d, m, n, Ap, Ai, Ax):
            nnz += len(cols)
        done += len(rows)
    t = time.perf_counter() - t0
    return {"value": nnz / t, "unit": "nnz-out/s", "cores": 1, "kind": "port",
            "sample": f"{done} random output rows of the same product ({nnz} entries), "
                      f"{t:.1f} s on one core (oracle C, row at a time: symbolic + numeric)"}


def main_coo(args):
    """Coalesce 50M COO triplets (35M distinct keys uniform in 1M x 1M, 15M
    repeats of earlier keys, shuffled; FP64 N(0,1)) into CSR with duplicates
    summed in input order."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, workloads

    lib = sdl.load()
    m = n = 1_000_000
    rows, cols, x = workloads.coo_triplets(2)
    nin = len(rows)
    dr, dc, dx = (DeviceArray.from_numpy(a) for a in (rows, cols, x))
    lib.sdl_synchronize()
    h = C.c_void_p()
    t0 = time.perf_counter()
    check(lib.sdl_coo_plan_create(nin, dr.ptr, dc.ptr, m, n, C.byref(h)), "coo plan")
    lib.sdl_synchronize()
    t_plan = time.perf_counter() - t0
    nnz = lib.sdl_coo_plan_nnz(h)
    out = DeviceArray(nnz, np.float64)
    stream = lib.sdl_library_stream()
    check(lib.sdl_coo_fill(h, dx.ptr, out.ptr, stream), "fill")
    lib.sdl_synchronize()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(1, min(args.steps, 50))
    lib.sdl_event_record(e0, stream)
    for _ in range(steps):
        check(lib.sdl_coo_fill(h, dx.ptr, out.ptr, stream), "fill")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms = lib.sdl_event_elapsed_ms(e0, e1) / steps
    # check: every output value is the sum of its key's duplicates (numpy,
    # any order: 1e-13 of the summed magnitudes)
    cp = np.empty(m + 1, np.int64)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_coo_plan_rowptr(h), cp.nbytes)
    got = out.numpy()
    key = rows.astype(np.int64) * n + cols
    uk, inv = np.unique(key, return_inverse=True)
    ref = np.bincount(inv, weights=x)
    mag = np.bincount(inv, weights=np.abs(x))
    assert nnz == len(uk) and cp[-1] == nnz
    worst = float(np.max(np.abs(got - ref) / (mag + 1e-300)))
    assert worst <= 1e-13, worst
    bytes_ = 8 * nin + 8 * nnz                       # the value fill: values in, values out
    bytes_all = 16 * nin + 4 * (m + 1) + 12 * nnz    # the whole coalesce (structure + values)
    out_json = {
        "metric": "COO coalesce triplets/s (duplicates summed in input order)",
        "value": nin / (ms * 1e-3), "unit": "triplets/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("coalesce 50M COO triplets in 1M x 1M (35M distinct keys, "
                                "30% duplicates, shuffled), FP64 N(0,1), seed 2; one step = "
                                "one value fill of the coalesced CSR"),
                   "n_in": nin, "nnz_out": nnz, "plan_s": t_plan,
                   "max_rel_err_vs_numpy": worst},
        "roofline": {"bound": "hbm", "kernel": "k_coalesce_fill",
                     "achieved": bytes_ / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_launch": bytes_},
        "plan_triplets_per_s": nin / t_plan,
        "alg_bytes_whole_coalesce": bytes_all,
        "cpu_baseline": None if args.no_cpu else pinned_cpu(coo_cpu_baseline, rows, cols, x, m, n),
    }
    lib.sdl_coo_plan_destroy(h)
    print(json.dumps(out_json), flush=True)


def main_sum(args):
    """Hessian-assembly sum: four 1M x 1M operands of ~16 entries per row
    (each later operand shares half of operand 0's draws), summed into their
    union (~40M entries from 64M operand entries) from +0.0 in operand order.
    One step = one value fill through the device sum plan."""
    import sparsediff_linalg_amd as sdl
    from oracle import oracle as O
    from sparsediff_linalg_amd import DeviceArray, check, workloads

    lib = sdl.load()
    m = n = 1_000_000
    ops = workloads.sum_operands(6, m, n, 4, 16, 0.5)
    nops = len(ops)
    dev = [tuple(DeviceArray.from_numpy(a) for a in o) for o in ops]
    vp = C.c_void_p
    dP = (vp * nops)(*[d[0].ptr for d in dev])
    dI = (vp * nops)(*[d[1].ptr for d in dev])
    dX = (vp * nops)(*[d[2].ptr for d in dev])
    nz = (C.c_longlong * nops)(*[len(o[1]) for o in ops])
    items = int(sum(nz))
    lib.sdl_synchronize()
    h = vp()
    t0 = time.perf_counter()
    check(lib.sdl_sum_plan_create(nops, m, n, dP, dI, nz, 0, C.byref(h)), "sum plan")
    lib.sdl_synchronize()
    t_plan = time.perf_counter() - t0
    nnz = lib.sdl_sum_plan_nnz(h)
    masked = bool(lib.sdl_sum_plan_is_masked(h))
    out = DeviceArray(nnz, np.float64)
    stream = lib.sdl_library_stream()
    for _ in range(max(1, args.warmup)):
        check(lib.sdl_sum_fill(h, dX, None, out.ptr, stream), "fill")
    lib.sdl_synchronize()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(1, min(args.steps, 100))
    lib.sdl_event_record(e0, stream)
    for _ in range(steps):
        check(lib.sdl_sum_fill(h, dX, None, out.ptr, stream), "fill")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms = lib.sdl_event_elapsed_ms(e0, e1) / steps
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        src, dst = DeviceArray(nb // 8, np.float64), DeviceArray(nb // 8, np.float64)
        for _ in range(3):
            check(lib.sdl_copy_kernel(dst.ptr, src.ptr, nb, stream), "copy")
        lib.sdl_synchronize()
    # whole output against the oracle, bit for bit
    Cp, Ci, Cx, _ = O.sum_n(m, n, ops)
    cp = np.empty(m + 1, np.int64)
    ci = np.empty(nnz, np.int32)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_sum_plan_rowptr(h), cp.nbytes)
    lib.sdl_memcpy_dtoh(ci.ctypes.data, lib.sdl_sum_plan_colind(h), ci.nbytes)
    exact = (np.array_equal(cp, Cp) and np.array_equal(ci, Ci)
             and np.array_equal(out.numpy().view(np.int64), Cx.view(np.int64)))
    assert exact, "sum differs from the oracle"
    # the row-scaled fill (sum_scaled_matrices semantics for four operands): a
    # plan that keeps the rows
    h2 = vp()
    check(lib.sdl_sum_plan_create(nops, m, n, dP, dI, nz, 1, C.byref(h2)), "sum plan (rows)")
    dscale = np.random.default_rng(7).uniform(0.5, 2.0, (nops, m))
    dD = DeviceArray.from_numpy(dscale)
    out2 = DeviceArray(nnz, np.float64)
    check(lib.sdl_sum_fill(h2, dX, dD.ptr, out2.ptr, stream), "scaled fill")
    lib.sdl_synchronize()
    lib.sdl_event_record(e0, stream)
    for _ in range(steps):
        check(lib.sdl_sum_fill(h2, dX, dD.ptr, out2.ptr, stream), "scaled fill")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms_scaled = lib.sdl_event_elapsed_ms(e0, e1) / steps
    _, _, Sx, _ = O.sum_n(m, n, ops, d=dscale)
    exact_scaled = np.array_equal(out2.numpy().view(np.int64), Sx.view(np.int64))
    assert exact_scaled, "scaled sum differs from the oracle"
    lib.sdl_sum_plan_destroy(h2)
    bytes_ = 8 * items + 8 * nnz               # operand values in, sums out
    # plan bytes read per fill: operand masks + tile bases, or segments + item order
    meta = nnz + 4 * nops * ((nnz + 255) // 256) if masked else 4 * (nnz + 1) + 4 * items
    out_json = {
        "metric": "Hessian-assembly sum entries/s (output entries filled per second)",
        "value": nnz / (ms * 1e-3), "unit": "nnz-out/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("sum of 4 operands 1M x 1M, ~16 entries per row each, operands "
                                "1-3 sharing half of operand 0's draws, FP64 N(0,1), seed 6; one "
                                "step = one value fill of the union"),
                   "operand_entries": items, "nnz_out": nnz, "plan_s": t_plan,
                   "bit_exact_vs_oracle": exact},
        "row_scaled_fill": {"kernel": "k_sum_fill_mask<4,true>", "ms": ms_scaled,
                            "hbm_frac_alg": (8 * items + 8 * nnz + 8 * nops * m + 4 * nnz)
                            / (ms_scaled * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "bit_exact_vs_oracle": exact_scaled},
        "roofline": {"bound": "hbm",
                     "kernel": "k_sum_fill_mask" if masked else "k_sum_fill_src<false>",
                     "achieved": bytes_ / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_sum_fill_mask_hessian_sum_traffic.json"),
                     "alg_bytes_per_launch": bytes_, "plan_bytes_per_launch": meta},
        "cpu_baseline": None if args.no_cpu else pinned_cpu(sum_cpu_baseline, ops, m, n,
                                                             args.cpu_seconds),
    }
    lib.sdl_sum_plan_destroy(h)
    print(json.dumps(out_json), flush=True)


def main_lower_coo(args):
    """Solver hand-off: the lower-triangular COO of a Hessian-shaped product
    output kept in HBM.  A = 2^20 x 2^20 banded (bandwidth 32, FP64 N(0,1)),
    C = A^T diag(d) A (2^20 rows, ~129 entries per row); one step = one refresh
    of the exported values from C.x (sdl_lower_coo_refresh)."""
    import sparsediff_linalg_amd as sdl
    from oracle import oracle as O
    from sparsediff_linalg_amd import DeviceArray, check, workloads
    from sparsediff_linalg_amd.predictor import Plan

    lib = sdl.load()
    m, bw = 1 << 20, 32
    Ap, Ai, Ax = workloads.banded_csr(m, bw)
    pl = Plan((Ap, Ai), None, m, m)
    nnz = pl.nnz()
    dA = DeviceArray.from_numpy(Ax)
    dd = DeviceArray.from_numpy(workloads.uniform(31, m, 0.1, 10.0))
    Cx = DeviceArray(nnz, np.float64)
    check(lib.sdl_btda_refill(pl.h, dA.ptr, dA.ptr, dd.ptr, Cx.ptr, None), "refill")
    lib.sdl_synchronize()
    h = C.c_void_p()
    t0 = time.perf_counter()
    check(lib.sdl_lower_coo_create(m, lib.sdl_btda_plan_rowptr(pl.h), lib.sdl_btda_plan_colind(pl.h),
                                   C.byref(h)), "lower coo")
    t_create = time.perf_counter() - t0
    cnt = lib.sdl_lower_coo_nnz(h)
    x = DeviceArray(cnt, np.float64)
    stream = lib.sdl_library_stream()
    for _ in range(max(1, args.warmup)):
        check(lib.sdl_lower_coo_refresh(h, Cx.ptr, x.ptr, stream), "refresh")
    lib.sdl_synchronize()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(1, min(args.steps, 100))
    lib.sdl_event_record(e0, stream)
    for _ in range(steps):
        check(lib.sdl_lower_coo_refresh(h, Cx.ptr, x.ptr, stream), "refresh")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms = lib.sdl_event_elapsed_ms(e0, e1) / steps
    # whole export against the oracle's (rows, cols, value_map, values)
    cp = np.empty(m + 1, np.int64)
    lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(pl.h), cp.nbytes)
    ci = np.empty(nnz, np.int32)
    lib.sdl_memcpy_dtoh(ci.ctypes.data, lib.sdl_btda_plan_colind(pl.h), ci.nbytes)
    cx = Cx.numpy()
    r, c, v, vm = O.coo_lower(m, cp.astype(np.int32), ci, cx)

    def dev(ptr, n, dt):
        a = np.empty(n, dt)
        lib.sdl_memcpy_dtoh(a.ctypes.data, ptr, a.nbytes)
        return a

    exact = (cnt == len(r) and np.array_equal(dev(lib.sdl_lower_coo_rows(h), cnt, np.int32), r)
             and np.array_equal(dev(lib.sdl_lower_coo_cols(h), cnt, np.int32), c)
             and np.array_equal(dev(lib.sdl_lower_coo_value_map(h), cnt, np.int64), vm)
             and np.array_equal(x.numpy().view(np.int64), v.view(np.int64)))
    assert exact, "lower-triangular export differs from the oracle"
    bytes_ = 16 * cnt  # values read from C.x, written to the export
    cpu = None
    if not args.no_cpu:
        if O.ref_available():
            k, fills, t = O.ref_lower_coo_timer(m, m, cp.astype(np.int32), ci, cx, args.cpu_seconds)
            cpu = {"value": k * fills / t, "unit": "entries/s", "cores": 1, "kind": "reference",
                   "sample": (f"whole export, {fills} refreshes in {t:.2f} s on one core "
                              "(oracle/_ref refresh_lower_triangular_coo)")}
        else:
            t0 = time.perf_counter()
            O.coo_lower(m, cp.astype(np.int32), ci, cx)
            t = time.perf_counter() - t0
            cpu = {"value": cnt / t, "unit": "entries/s", "cores": 1, "kind": "port",
                   "sample": f"whole export (structure and values) in {t:.2f} s, the oracle"}
    out_json = {
        "metric": "lower-triangular COO refresh entries/s (solver hand-off)",
        "value": cnt / (ms * 1e-3), "unit": "entries/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("lower-triangular COO of C = A^T diag(d) A, A 2^20 x 2^20 banded "
                                "(bandwidth 32), C with ~129 entries per row; one step = one "
                                "refresh of the exported values"),
                   "nnz_C": nnz, "exported": cnt, "create_s": t_create,
                   "bit_exact_vs_oracle": exact},
        "roofline": {"bound": "hbm", "kernel": "k_lower_refresh",
                     "achieved": bytes_ / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": bytes_},
        "cpu_baseline": cpu,
    }
    lib.sdl_lower_coo_destroy(h)
    pl.close()
    print(json.dumps(out_json), flush=True)


def main_dense_kron(args):
    """The dense Kronecker-vector products of dense.h on device arrays:
    v = (Y kron I_m) w and v = (I_n kron X^T) w with m = 4096, k = 256,
    n = 4096 (FP64 N(0,1)); one step = both products.  Each output is an
    ordered sum of rounded products (the reference's arithmetic), so the
    kernels use neither MFMA nor fused multiply-add."""
    import sparsediff_linalg_amd as sdl
    from oracle import oracle as O
    from sparsediff_linalg_amd import DeviceArray, check

    lib = sdl.load()
    m, k, n = 4096, 256, 4096
    rng = np.random.default_rng(12)
    Y, X, w = rng.standard_normal(k * n), rng.standard_normal(m * k), rng.standard_normal(m * n)
    dY, dX, dw = (DeviceArray.from_numpy(a) for a in (Y, X, w))
    dv1, dv2 = DeviceArray(m * k, np.float64), DeviceArray(k * n, np.float64)
    stream = lib.sdl_library_stream()

    def step():
        check(lib.sdl_y_kron_i_vec(m, k, n, dY.ptr, dw.ptr, dv1.ptr, stream), "y_kron_i")
        check(lib.sdl_i_kron_xt_vec(m, k, n, dX.ptr, dw.ptr, dv2.ptr, stream), "i_kron_xt")

    for _ in range(max(1, args.warmup)):
        step()
    lib.sdl_synchronize()
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    steps = max(1, min(args.steps, 50))
    lib.sdl_event_record(e0, stream)
    for _ in range(steps):
        step()
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms = lib.sdl_event_elapsed_ms(e0, e1) / steps
    # sampled outputs against the oracle, bit for bit (64 rows of W / columns)
    v1, v2 = dv1.numpy(), dv2.numpy()
    rows = np.arange(0, m, m // 64)
    Ys = Y.reshape(n, k).T  # Y[j + col*k]
    W = w.reshape(n, m).T   # W[row, col] = w[row + col*m]
    exact = True
    for r in rows:
        ref = np.zeros(k)
        for c in range(n):
            ref = ref + Ys[:, c] * W[r, c]
        exact = exact and np.array_equal(v1[r + np.arange(k) * m].view(np.int64), ref.view(np.int64))
    cols = np.arange(0, n, n // 16)
    Xm = X.reshape(k, m).T  # X[row + j*m]
    for c in cols:
        ref = np.zeros(k)
        for r in range(m):
            ref = ref + Xm[r, :] * W[r, c]
        exact = exact and np.array_equal(v2[np.arange(k) + c * k].view(np.int64), ref.view(np.int64))
    assert exact, "kron products differ from the ordered sums"
    flops = 2 * 2.0 * m * k * n
    out_json = {
        "metric": "dense Kronecker-vector products FP64 GFLOP/s (ordered sums)",
        "value": flops / (ms * 1e-3) / 1e9, "unit": "GFLOP/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("v = (Y kron I_m) w and v = (I_n kron X^T) w, m = 4096, k = 256, "
                                "n = 4096, FP64 N(0,1), seed 12; one step = both products"),
                   "sampled_outputs_bit_exact": exact},
        "roofline": {"bound": "fp64", "kernel": "k_ordered_dots<2,16,32>",
                     "achieved": flops / (ms * 1e-3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": flops / (ms * 1e-3) / 1e12 / 78.6, "traffic": None,
                     "note": ("separate FP64 multiply and add per term (the reference's rounding): "
                              "the attainable ceiling is half the FMA peak")},
        "cpu_baseline": None if args.no_cpu else pinned_cpu(dense_kron_cpu_baseline, m, k, n, Y, X,
                                                             w),
    }
    print(json.dumps(out_json), flush=True)


def dense_kron_cpu_baseline(m, k, n, Y, X, w):
    """The reference's own Y_kron_I_vec / I_kron_XT_vec (oracle/_ref, one core)
    on the first 512 columns of W (the same per-output work; n scaled 1/8)."""
    from oracle import oracle as O
    ns = 512
    Ys, ws = Y[:k * ns], w[:m * ns]
    fn1 = O.ref_lib().Y_kron_I_vec if O.ref_available() else None
    fn2 = O.ref_lib().I_kron_XT_vec if O.ref_available() else None
    t0 = time.perf_counter()
    O.y_kron_i_vec(m, k, ns, Ys, ws, fn1)
    O.i_kron_xt_vec(m, k, ns, X, ws, fn2)
    t = time.perf_counter() - t0
    return {"value": 2 * 2.0 * m * k * ns / t / 1e9, "unit": "GFLOP/s", "cores": 1,
            "kind": "reference" if fn1 else "port",
            "sample": f"both products with n = {ns} (1/8 of the columns) in {t:.2f} s on one core"}


def sum_cpu_baseline(ops, m, n, seconds):
    """The reference's own fill (memset, then accumulator over each operand
    through the maps of sum_n_csr_alloc; matrix_sum.c:142-155) from
    oracle/_ref on one core, whole workload, repeated for about `seconds`;
    the oracle's fill when the reference build is absent."""
    from oracle import oracle as O
    nops = len(ops)
    if O.ref_available():
        r = O.ref_lib()
        keep = []
        views = [O._csr_view(m, n, p, i, x, keep) for p, i, x in ops]
        flat = np.zeros(sum(v.nnz for v in views), np.int32)
        parr = (C.POINTER(O._CSR) * nops)(*[C.pointer(v) for v in views])
        Cc = r.sum_n_csr_alloc(parr, nops, O.I(flat))
        nnz = Cc.contents.nnz
        Cx = np.zeros(nnz, np.float64)
        xs = [np.ascontiguousarray(x, np.float64) for _, _, x in ops]
        offs = np.concatenate([[0], np.cumsum([v.nnz for v in views])])
        maps = [np.ascontiguousarray(flat[offs[k]:offs[k + 1]]) for k in range(nops)]
        fills, t = 0, 0.0
        while t < seconds or fills == 0:
            t0 = time.perf_counter()
            Cx.fill(0.0)
            for k in range(nops):
                r.accumulator(O.D(xs[k]), views[k].nnz, O.I(maps[k]), O.D(Cx))
            t += time.perf_counter() - t0
            fills += 1
        r.free_CSR_matrix(Cc)
        kind, what = "reference", "oracle/_ref (the reference's accumulator fill)"
    else:
        t0 = time.perf_counter()
        _, _, Cx, _ = O.sum_n(m, n, ops)
        t, fills, nnz = time.perf_counter() - t0, 1, len(Cx)
        kind, what = "port", "the oracle (structure and values)"
    return {"value": nnz * fills / t, "unit": "nnz-out/s", "cores": 1, "kind": kind,
            "sample": f"whole workload, {fills} fills in {t:.2f} s on one core, {what}"}


def main_learned(args):
    """Learned row binning, evaluated on the R-MAT and banded products: the
    symbolic phase of A^T A for a banded matrix (2^20 rows, bandwidth 64) and
    an R-MAT graph (a, b, c = 0.57, 0.19, 0.19, 16 edges per node, duplicates
    coalesced), plus a held-out product of the stacked-block law (1024 blocks,
    column permutations in 2^20 columns).  The exact path sizes each row's
    count accumulator from min(terms, column span); the predicted path lets a
    network trained on the GPU from the exact counts of the 20000 x 5000
    B^T D A product, rank-Zipf A^T A products and the stacked-block A^T A
    only (none of the evaluated families) lower the bound of the rows where a
    prediction could pay, by expected cost.  Structures compared with the
    exact plan's."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import workloads
    from sparsediff_linalg_amd.predictor import Plan, Predictor, training_products, training_set

    lib = sdl.load()
    t0 = time.perf_counter()
    X, Y, n = training_set(training_products(small=False))
    t_label = time.perf_counter() - t0
    mlp = Predictor(seed=1)
    t0 = time.perf_counter()
    loss = mlp.train(X, Y, n, epochs=args.epochs)
    lib.sdl_synchronize()
    t_train = time.perf_counter() - t0
    acc = float(np.mean(mlp.predict(X, n) == Y.numpy()))
    net = {"training": "20000x5000 B^T D A, rank-Zipf A^T A (3 sizes), stacked-block A^T A",
           "rows": n, "final_loss": loss, "exact_bin_train": acc, "label_s": t_label,
           "train_s": t_train}
    print(f"[learned] trained: {net}", file=sys.stderr, flush=True)

    def dtoh(ptr, count, dt):
        out = np.empty(count, dt)
        if count:
            lib.sdl_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes)
        return out

    def same_structure(a, b, nrows):
        pa = dtoh(lib.sdl_btda_plan_rowptr(a.h), nrows + 1, np.int64)
        pb = dtoh(lib.sdl_btda_plan_rowptr(b.h), nrows + 1, np.int64)
        if not np.array_equal(pa, pb):
            return False
        nnz = int(pa[-1])
        chunk = nnz if nnz <= 10 ** 8 else 10 ** 6
        starts = [0] if nnz <= 10 ** 8 else np.random.default_rng(0).integers(0, nnz - chunk, 50)
        for s0 in starts:
            ia = dtoh(lib.sdl_btda_plan_colind(a.h) + 4 * int(s0), chunk, np.int32)
            ib = dtoh(lib.sdl_btda_plan_colind(b.h) + 4 * int(s0), chunk, np.int32)
            if not np.array_equal(ia, ib):
                return False
        return True

    def timed(p, i, nn, **kw):
        Plan((p, i), None, nn, nn, **kw).close()  # warm-up (kernels, allocator)
        best = None
        for _ in range(3):
            lib.sdl_synchronize()
            t0 = time.perf_counter()
            pl = Plan((p, i), None, nn, nn, **kw)
            lib.sdl_synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[1]:
                if best:
                    best[0].close()
                best = (pl, dt)
            else:
                pl.close()
        return best

    def wide_blocks():
        wb = workloads.stacked_pd_blocks(14, nblocks=1024, ncols=1 << 20)
        return workloads.blocks_as_csr(wb)

    sc = args.rmat_scale
    products = [("banded 2^20 x 2^20, bandwidth 64", lambda: workloads.banded_csr(1 << 20, 64),
                 1 << 20, True),
                (f"R-MAT 2^{sc} nodes, 16 edges/node", lambda: workloads.rmat_csr(sc), 1 << sc,
                 True),
                ("stacked blocks as CSR, 1024 blocks, columns in 2^20 (held out)", wide_blocks,
                 1 << 20, False)]
    results = []
    tot_terms = tot_exact = tot_pred = 0.0
    for label, make, nn, eval_set in products:
        p, i, _ = make()
        ex, t_ex = timed(p, i, nn)
        terms = lib.sdl_btda_plan_terms(ex.h)
        pr, t_pr = timed(p, i, nn, mode=1, mlp=mlp)
        same = same_structure(ex, pr, nn)
        assert same, label
        st = pr.pred_stats()
        rec = {"product": f"A^T A, {label}", "nnz_A": int(p[-1]), "terms": terms,
               "nnz_out": ex.nnz(), "exact_symbolic_s": t_ex, "predicted_symbolic_s": t_pr,
               "gain": t_ex / t_pr, "structure_same": same,
               "misbin_rate_evaluated": st["misbin"] / max(st["candidates"], 1),
               "recount_rate_lowered": st["overflow"] / max(st["predicted"] + st["overflow"], 1),
               **st}
        ex.close()
        pr.close()
        if eval_set:
            tot_terms += terms
            tot_exact += t_ex
            tot_pred += t_pr
        results.append(rec)
        print(f"[learned] {json.dumps(rec)}", file=sys.stderr, flush=True)
    out = {
        "metric": "symbolic A^T A terms/s with predicted row binning (learned binning)",
        "value": tot_terms / tot_pred, "unit": "terms/s", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": tot_pred * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": ("symbolic A^T A of a banded matrix (2^20 rows, bandwidth 64) "
                                f"and an R-MAT graph (2^{sc} nodes); one step = both symbolic "
                                "phases with the predicted row binning (best of 3)"),
                   "products": results, "network": net},
        "exact_terms_per_s": tot_terms / tot_exact,
        "gain_vs_exact": tot_exact / tot_pred,
        "roofline": None,
        "cpu_baseline": None if args.no_cpu else pinned_cpu(symbolic_cpu_baseline,
                                                             args.cpu_seconds),
    }
    mlp.close()
    print(json.dumps(out), flush=True)


def symbolic_cpu_baseline(seconds):
    """The oracle's single-thread two-pass symbolic Gustavson (mark-array count,
    fill, per-row sort) over chunks of 4096 output rows of the banded A^T A
    (2^20 rows, bandwidth 64) at random offsets, until about `seconds`."""
    import ctypes as Cc

    from oracle import oracle as O
    from sparsediff_linalg_amd import workloads
    n = 1 << 20
    p, i, _ = workloads.banded_csr(n, 64)
    Lp, Li, _ = O.transpose(n, n, p, i)
    Lp = np.asarray(Lp, np.int32)
    Li = np.asarray(Li, np.int32)
    Ap, Ai = np.ascontiguousarray(p, np.int32), np.ascontiguousarray(i, np.int32)
    o = O.lib()
    P = Cc.POINTER(Cc.c_int)
    lens = np.diff(Ap).astype(np.int64)
    rng = np.random.default_rng(5)
    rows = terms = 0
    t = 0.0
    while t < seconds:
        a = int(rng.integers(0, n - 4096))
        lp = np.ascontiguousarray(Lp[a:a + 4097] - Lp[a], np.int32)
        li = np.ascontiguousarray(Li[Lp[a]:Lp[a + 4096]], np.int32)
        terms += int(lens[li].sum())
        cp = np.zeros(4097, np.int64)
        t0 = time.perf_counter()
        nnz = o.orc_spgemm_symbolic(4096, lp.ctypes.data_as(P), li.ctypes.data_as(P), 1, n, n,
                                    Ap.ctypes.data_as(P), Ai.ctypes.data_as(P),
                                    cp.ctypes.data_as(Cc.POINTER(Cc.c_longlong)), None)
        ci = np.zeros(max(nnz, 1), np.int32)
        o.orc_spgemm_symbolic(4096, lp.ctypes.data_as(P), li.ctypes.data_as(P), 1, n, n,
                              Ap.ctypes.data_as(P), Ai.ctypes.data_as(P),
                              cp.ctypes.data_as(Cc.POINTER(Cc.c_longlong)), ci.ctypes.data_as(P))
        t += time.perf_counter() - t0
        rows += 4096
    return {"value": terms / t, "unit": "terms/s", "cores": 1, "kind": "port",
            "sample": f"{rows} output rows of the banded A^T A (2^20 rows, bandwidth 64) in chunks "
                      f"of 4096, two-pass symbolic (count, fill + sort), {t:.1f} s on one core "
                      f"(oracle C)"}


def multi_cpu_baseline(w, seconds):
    """The reference's single-thread path (oracle/_ref when built, else the C
    restatement) per instance: alloc once, two refills with fresh d, over
    instances in random order until about `seconds` of refill time."""
    from oracle import oracle as O
    from sparsediff_linalg_amd import workloads
    rng = np.random.default_rng(9)
    kind = "reference" if O.ref_available() else "port"
    nnz_total = fills_n = 0
    fill_s = 0.0
    done = 0
    for k in rng.permutation(len(w["m"])):
        if fill_s >= seconds:
            break
        r0, r1 = int(w["roff"][k]), int(w["roff"][k + 1])
        c0, c1 = int(w["coff"][k]), int(w["coff"][k + 1])
        ops = []
        for tag in ("A", "B"):
            p, i, x = w[tag]
            a, b = int(p[r0]), int(p[r1])
            ops.append(((p[r0:r1 + 1] - a).astype(np.int32), (i[a:b] - c0).astype(np.int32),
                        np.ascontiguousarray(x[a:b])))
        (Ap, Ai, Ax), (Bp, Bi, Bx) = ops
        m, nn = r1 - r0, c1 - c0
        ds = [workloads.uniform(60_000 + 2 * int(k) + j, m, 0.1, 10.0) for j in range(2)]
        if kind == "reference":
            _, fills, cx = O.ref_fill_timer(m, nn, nn, Ap, Ai, Ax, Bp, Bi, Bx, ds)
            nnz = len(cx)
        else:
            _, fills = O.port_fill_timer(m, nn, nn, Ap, Ai, Ax, Bp, Bi, Bx, ds)
            nnz = len(O.btda(m, nn, nn, Ap, Ai, Ax, Bp, Bi, Bx, ds[0])[1])
        fill_s += sum(fills)
        nnz_total += nnz * len(fills)
        fills_n += len(fills)
        done += 1
    return {"value": nnz_total / fill_s, "unit": "nnz-out/s", "cores": 1, "kind": kind,
            "sample": f"{done} instances in random order, alloc once and 2 refills each with fresh "
                      f"d ({fills_n} refills, {fill_s:.1f} s of refill time on one core)"}


def coo_cpu_baseline(rows, cols, x, m, n, count=5_000_000):
    """The oracle's single-thread coalesce (counting sort by row in input
    order, per-row column sort, ordered sums: structure and values) of the
    first `count` triplets."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    O.coo_coalesce(rows[:count], cols[:count], x[:count], m, n)
    t = time.perf_counter() - t0
    return {"value": count / t, "unit": "triplets/s", "cores": 1, "kind": "port",
            "sample": f"first {count} triplets of the same stream, coalesced whole (structure "
                      f"and values) in {t:.2f} s on one core (oracle C)"}


def main_multi(args):
    """Many uneven B^T D A instances (m log-uniform on [100, 50000], n = m/4,
    4..12 nnz/row): one symbolic pass over the block-diagonal product, then
    `rounds` fresh d vectors per instance refilled by one replay of a captured
    HIP graph (every round overwrites C).  Beside it, the per-call path: one
    launch per (instance, d)."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, workloads
    from sparsediff_linalg_amd.multi import MultiProduct, kmajor_ld

    lib = sdl.load()
    R = args.rounds
    K = args.multi_k
    t0 = time.perf_counter()
    w = workloads.multi_instances(4, ninst=args.instances)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    mp = MultiProduct(w, K=K, tile=args.multi_layout == "tile")
    lib.sdl_synchronize()
    t_sym = time.perf_counter() - t0
    M = mp.M
    # d vectors: round r is uniform[0.1, 10] with seed 40000 + r.  A
    # round-major pool of the first rounds (autotuning, per-call path) and the
    # k-major pool of all rounds (the graph), assembled in row chunks.
    host = [workloads.uniform(40_000 + r, M, 0.1, 10.0) for r in range(R)]
    nrm = max(K, min(R, 5))
    D = DeviceArray(nrm * M, np.float64)
    for r in range(nrm):
        lib.sdl_memcpy_htod(D.offset(r * M), host[r % R].ctypes.data, host[r % R].nbytes)
    ldt = kmajor_ld(R)
    DT = DeviceArray(M * ldt, np.float64)
    chunk = 1 << 20
    for a in range(0, M, chunk):
        b = min(M, a + chunk)
        blk = np.zeros((b - a, ldt))
        for r in range(R):
            blk[:, r] = host[r][a:b]
        lib.sdl_memcpy_htod(DT.offset(a * ldt), blk.ctypes.data, blk.nbytes)
    del host
    tuned = mp.autotune(D, reps=3)
    stream = lib.sdl_library_stream()
    graph = mp.capture(DT, ldt, R)
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    lib.sdl_graph_launch(graph, stream)  # warm-up
    lib.sdl_synchronize()
    reps = max(1, args.steps if args.steps <= 5 else 3)
    lib.sdl_event_record(e0, stream)
    for _ in range(reps):
        lib.sdl_graph_launch(graph, stream)
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms_graph = lib.sdl_event_elapsed_ms(e0, e1) / reps
    last_slot = (R - 1) % K
    final = mp.values(last_slot)
    # per-call path: one launch per (instance, d), round-major d
    plain_rounds = min(R, 5)
    mp.plain_round(D, 0)
    lib.sdl_synchronize()
    lib.sdl_event_record(e0, stream)
    for r in range(plain_rounds):
        mp.plain_round(D, r)
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms_plain_round = lib.sdl_event_elapsed_ms(e0, e1) / plain_rounds
    last = workloads.uniform(40_000 + R - 1, M, 0.1, 10.0)
    lib.sdl_memcpy_htod(D.ptr, last.ctypes.data, last.nbytes)
    mp.plain_round(D, 0)  # round R-1's d, now at the pool's first slot
    lib.sdl_synchronize()
    same = all(np.array_equal(a.view(np.int64), b.view(np.int64))
               for a, b in zip(mp.values(0), final))
    assert same, "per-call refills differ from the graph's"
    nnzA, nnzB = int(w["A"][0][-1]), int(w["B"][0][-1])
    nB = int(w["coff"][-1])
    S = (4 * (nB + 1) + 12 * nnzB) + (4 * (M + 1) + 12 * nnzA) + (4 * (nB + 1) + 4 * mp.nnz)
    bytes_R = S + R * (8 * M + 8 * mp.nnz)
    flops_R = R * (nnzB + 2 * mp.terms)
    t = ms_graph * 1e-3
    achieved = bytes_R / t / 1e9
    out = {
        "metric": METRIC + " (many instances)",
        "value": R * mp.nnz / t,
        "unit": "nnz-out/s",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": ms_graph,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.instances} independent B^T diag(d) A instances, m log-uniform "
                         "[100,50000], n=m/4, 4..12 nnz/row uniform columns, FP64 N(0,1), seed 4; "
                         f"one step = one replay of a HIP graph of {R} refill rounds of every "
                         "instance (each round a fresh d, overwriting C)"),
            "instances": args.instances, "rows": M, "nnz_out": mp.nnz, "terms": mp.terms,
            "groups": len(mp.g), "rounds": R, "rounds_per_launch": K, "gen_s": t_gen,
            "symbolic_s": t_sym, "tuned": tuned, "bound_layout": args.multi_layout,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_num_pull_tile_batched (bound tile layout)" if args.multi_layout == "tile"
                       else "k_num_pull_interleaved (bound operand pairs)")
                      + ", K rounds per launch per group",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "alg_bytes_per_launch": bytes_R,
        },
        "gflops": flops_R / t / 1e9,
        "plain_gpu": {"ms_per_round": ms_plain_round,
                      "launches_per_round": sum(len(g.ranges) for g in mp.g),
                      "nnz_out_per_s": mp.nnz / (ms_plain_round * 1e-3)},
        "batched_vs_plain": ms_plain_round / (ms_graph / R),
        "cpu_baseline": None if args.no_cpu else pinned_cpu(multi_cpu_baseline, w, args.cpu_seconds),
    }
    lib.sdl_graph_destroy(graph)
    mp.close()
    print(json.dumps(out), flush=True)


FP64_MFMA_PEAK_TFS = 78.6


def main_blocks(args):
    """Stacked permuted-dense products: 256 dense FP64 blocks, m0 and n0
    log-uniform on [64, 512], row partition, col_perms n0-subsets of [0, 4096).
    One step = one device-resident refill of the product for a fresh d (the
    grouped MFMA GEMM over all block pairs plus the ordered coalesce)."""
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, workloads
    from sparsediff_linalg_amd._lib import MAT_p

    lib = sdl.load()
    if args.gemm_slice:
        check(0 if lib.sdl_block_gemm_slice(args.gemm_slice) > 0 else 1, "gemm slice")
    if args.coalesce_lists:
        check(0 if lib.sdl_block_coalesce_mode(1) >= 0 else 1, "coalesce mode")
    w = workloads.stacked_pd_blocks(3)
    m, N = w["m"], w["n"]
    btda = args.workload == "blocks-btda"

    def stack(blocks):
        hs = []
        for rp, cp, X in blocks:
            X = np.ascontiguousarray(X)
            hs.append(lib.new_permuted_dense(m, N, len(rp), len(cp), sdl.ptr(rp), sdl.ptr(cp),
                                             sdl.ptr(X)))
        arr = (MAT_p * len(hs))(*hs)
        return lib.new_stacked_pd(m, N, len(hs), C.cast(arr, C.POINTER(MAT_p)), None, None)

    A = stack(w["A"])
    B = stack(w["B"]) if btda else None
    d = w["d"]
    t0 = time.perf_counter()
    Cm = lib.BTA_matrices_alloc(A, B) if btda else lib.ATA_matrices_alloc(A)
    t_alloc = time.perf_counter() - t0
    left = B if btda else A

    def dropin_fill(dv):
        if btda:
            lib.BTDA_matrices_fill_values(A, sdl.ptr(dv), B, Cm)
        else:
            lib.ATDA_matrices_fill_values(A, sdl.ptr(dv), Cm)

    dropin_fill(d)  # uploads the operand mirrors once
    t0 = time.perf_counter()
    dropin_fill(d)
    t_dropin = time.perf_counter() - t0

    stats = (C.c_double * 8)()
    check(lib.sdl_block_stats(Cm, stats), "stats")
    flops = stats[0] + stats[1]
    out_cells = int(stats[5])
    NPOOL = 4
    pools = [DeviceArray.from_numpy(workloads.uniform(20_000 + p, m, 0.1, 10.0)) for p in range(NPOOL)]
    stream = lib.sdl_library_stream()
    # per-step events around the block products (GEMM) and the whole refill
    ev = [(lib.sdl_event_create(), lib.sdl_event_create(), lib.sdl_event_create())
          for _ in range(min(args.steps, 50))]
    e0, e1 = lib.sdl_event_create(), lib.sdl_event_create()
    for i in range(args.warmup):
        check(lib.sdl_block_refill_device(Cm, left, A, pools[i % NPOOL].ptr, None, None), "refill")
    lib.sdl_synchronize()
    lib.sdl_event_record(e0, stream)
    for i in range(args.steps):
        a, g, b = ev[i] if i < len(ev) else (None, None, None)
        check(lib.sdl_block_refill_device_split(Cm, left, A, pools[i % NPOOL].ptr, a, g, b),
              "refill")
    lib.sdl_event_record(e1, stream)
    lib.sdl_synchronize()
    ms_step = lib.sdl_event_elapsed_ms(e0, e1) / args.steps
    ms_gemm = float(np.mean([lib.sdl_event_elapsed_ms(a, g) for a, g, _ in ev]))
    ms_coalesce = float(np.mean([lib.sdl_event_elapsed_ms(g, b) for _, g, b in ev]))
    sustained, _ = mfma_ceiling(lib, iters=2000)

    # sampled check of the last refill against the blocks (numpy, FP64)
    d_last = pools[(args.steps - 1) % NPOOL].numpy()
    check(lib.sdl_block_publish(Cm), "publish")
    csr = Cm.contents.to_csr(Cm).contents
    nnz = csr.nnz
    p = np.ctypeslib.as_array(csr.p, (csr.m + 1,)).copy()
    ci = np.ctypeslib.as_array(csr.i, (max(nnz, 1),))[:nnz].copy()
    cx = np.ctypeslib.as_array(csr.x, (max(nnz, 1),))[:nnz].copy()
    rng = np.random.default_rng(0)
    Lb = w["B"] if btda else w["A"]
    inv = []
    for (rpL, cpL, XL), (rpA, cpA, XA) in zip(Lb, w["A"]):
        iL = np.full(N, -1, np.int64)
        iL[cpL] = np.arange(len(cpL))
        iA = np.full(N, -1, np.int64)
        iA[cpA] = np.arange(len(cpA))
        inv.append((iL, iA))
    worst = 0.0
    rows_of = np.repeat(np.arange(csr.m), np.diff(p))
    for q in rng.choice(nnz, min(nnz, 300), replace=False):
        r, c = rows_of[q], ci[q]
        ref, mag = 0.0, 0.0
        for k, ((rpL, cpL, XL), (rpA, cpA, XA)) in enumerate(zip(Lb, w["A"])):
            iL, iA = inv[k]
            if iL[r] >= 0 and iA[c] >= 0:
                t = XL[:, iL[r]] * d_last[rpA] * XA[:, iA[c]]
                ref += t.sum()
                mag += np.abs(t).sum()
        err = abs(cx[q] - ref) / (mag + 1e-300)
        worst = max(worst, err)
    assert worst <= 1e-13, f"block product differs from the blocks: rel {worst}"

    XA_cells = sum(X.size for _, _, X in w["A"])
    XB_cells = sum(X.size for _, _, X in w["B"]) if btda else 0
    bytes_alg = 8 * (XA_cells + XB_cells + m + out_cells)
    tfs = flops / (ms_step * 1e-3) / 1e12
    gemm_tfs = stats[0] / (ms_gemm * 1e-3) / 1e12
    gbs = bytes_alg / (ms_step * 1e-3) / 1e9
    cpu = None if args.no_cpu else pinned_cpu(blocks_cpu_baseline, w, btda, args.cpu_seconds,
                                              pin=False)
    out = {
        "metric": "stacked permuted-dense B^T·D·A GFLOP/s + % FP64 MFMA peak",
        "value": flops / (ms_step * 1e-3) / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("stacked permuted-dense: 256 FP64 blocks, m0,n0 log-uniform [64,512], "
                         "row partition, col_perms n0-subsets of [0,4096), seed 3; "
                         + ("C = B^T diag(d) A, B on A's rows with its own col_perms"
                            if btda else "C = A^T diag(d) A") + "; one step = one refill"),
            "m": m, "n": N, "out_cells": out_cells, "raw_cells": int(stats[4]),
            "gemm_tiles": int(stats[2]), "gemm_tasks": int(stats[3]),
            "coalesce_terms": int(stats[6]), "flops": flops, "alloc_s": t_alloc,
            "gemm_slice": args.gemm_slice or 16,
            "coalesce": "pull lists" if args.coalesce_lists else "by output row",
            "dropin_fill_ms_incl_pcie": t_dropin * 1e3, "sampled_max_rel_err": worst,
        },
        "roofline": {
            "bound": "mfma", "kernel": "k_ggemm<16> (block products; timed apart by events)",
            "achieved": gemm_tfs, "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
            "frac": gemm_tfs / FP64_MFMA_PEAK_TFS, "traffic": None,
            "sustained_mfma_tflops": sustained, "frac_of_sustained": gemm_tfs / sustained,
            "gemm_ms": ms_gemm, "coalesce_ms": ms_coalesce,
            "step_tflops": tfs, "alg_bytes_per_launch": bytes_alg, "alg_GBs": gbs,
            "hbm_frac": gbs / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
    }
    for M in (Cm, A, B):
        if M is not None:
            lib.free_matrix(M)
    print(json.dumps(out), flush=True)


def blocks_cpu_baseline(w, btda, seconds):
    """Host BLAS (numpy) on a bounded sample of the block pairs: per block
    X_L^T (d X_A), the dgemm the reference issues per pair; the coalesce is
    not included."""
    import threadpoolctl
    d = w["d"]
    Lb = w["B"] if btda else w["A"]
    flops, t = 0.0, 0.0
    k = 0
    while t < seconds and k < 10 * len(Lb):
        (rpL, _, XL), (rpA, _, XA) = Lb[k % len(Lb)], w["A"][k % len(Lb)]
        t0 = time.perf_counter()
        Y = d[rpA][:, None] * XA
        _ = XL.T @ Y
        t += time.perf_counter() - t0
        flops += 2.0 * XL.shape[0] * XL.shape[1] * XA.shape[1]
        k += 1
    info = threadpoolctl.threadpool_info()
    threads = max([i.get("num_threads", 1) for i in info if i.get("user_api") == "blas"] or [1])
    return {"value": flops / t / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{k} block products X_L^T diag(d) X_A with numpy/OpenBLAS dgemm "
                      f"({threads} threads), {t:.2f} s; coalesce not timed"}


def cpu_baseline(w, nnz, seconds):
    """Single-thread CPU path on a bounded sample: the C restatement of the
    reference's sparse path (oracle/liboracle.so; its bit-identity and timing
    against the reference's own sources are recorded in
    tests/golden/oracle_faithfulness.json), pinned to one core.  One alloc per
    pass of 64 fresh d vectors, then one fill per d, until about `seconds` of
    fill time; the value is from the median fill time (after one warm-up
    pass)."""
    try:
        from oracle import cpu_env
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"error": str(e)}
    import statistics

    from sparsediff_linalg_amd import workloads
    ds = [workloads.uniform(50_000 + r, w["m"], 0.1, 10.0) for r in range(64)]
    args_w = (w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"], w["Bp"], w["Bi"], w["Bx"])
    fills_all, allocs = [], []
    with cpu_env.pinned() as facts:
        O.port_fill_timer(*args_w, ds[:4])  # warm-up
        t_start = time.perf_counter()
        while sum(fills_all) < seconds and time.perf_counter() - t_start < 3 * seconds:
            a, fills = O.port_fill_timer(*args_w, ds)
            allocs.append(a)
            fills_all.extend(fills)
    med = statistics.median(fills_all)
    return {
        "value": nnz / med,
        "unit": "nnz-out/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"20000x5000 B^T D A product on the C restatement of the reference's sparse "
                   f"path (bit-identical to the reference build, time within 20 %): alloc (CSC "
                   f"mirror + Gustavson + plan) once per pass of 64 d vectors, "
                   f"{len(fills_all)} single-thread refills with fresh d, median fill "
                   f"{med * 1e3:.3f} ms, median alloc {statistics.median(allocs) * 1e3:.1f} ms"),
        "us_per_refill": med * 1e6,
        **facts,
    }


if __name__ == "__main__":
    main()
