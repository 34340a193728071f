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
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run the streaming copy kernel on this many MiB (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the sparse product is sharded by rows/instances in the design but no multi-GPU "
              "path is built.", file=sys.stderr)
        sys.exit(2)
    import sparsediff_linalg_amd as sdl
    from sparsediff_linalg_amd import DeviceArray, check, workloads
    from sparsediff_linalg_amd._lib import sdl_refill_cfg

    lib = sdl.load()
    K = args.K
    w = workloads.btda_config1(1)
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
        best.variant, best.rchunk = 3, 8
        check(lib.sdl_btda_set_refill_cfg(h, C.byref(best)), "set cfg")
        args.plain_steps = 0
    else:
        check(lib.sdl_btda_autotune(h, dAx.ptr, dBx.ptr, pools[0].ptr, m, K, dC.ptr, ldc, 10,
                                    C.byref(best), C.byref(best_us)), "autotune")

    stream = lib.sdl_library_stream()
    ev = [(lib.sdl_event_create(), lib.sdl_event_create()) for _ in range(args.steps)]
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
        check(step(i, ev[i]))
    lib.sdl_event_record(e_all1, stream)
    lib.sdl_synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms_total = lib.sdl_event_elapsed_ms(e_all0, e_all1)
    ms_step = ms_total / args.steps
    main_ms = float(np.mean([lib.sdl_event_elapsed_ms(a, b) for a, b in ev]))

    # correctness spot check of the last step against the single-refill path
    D_last = pools[(args.steps - 1) % NPOOL]
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

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "btda_refill_config1_traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

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
            "workload": ("config 1: C = B^T diag(d) A, A and B 20000x5000 CSR, 6 distinct uniform "
                         "columns/row, FP64 N(0,1), d U[0.1,10], seed 1; one step = one batched "
                         f"refill of C for K={K} fresh d vectors"),
            "K": K, "nnz_out": nnz, "terms": terms, "m": m, "n_A": nA, "n_B": nB,
            "refill_cfg": {"variant": best.variant, "rchunk": best.rchunk},
            "wall_s": t_wall,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "batched refill sequence (k_gather_terms + k_transpose_d + "
                      "k_num_pull_interleaved)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_K,
            "main_kernel_us": main_ms * 1e3,
            "main_kernel_alg_GBs": bytes_K / (main_ms * 1e-3) / 1e9,
        },
        "gflops": flops_K / (ms_step * 1e-3) / 1e9,
        "plain_gpu": {
            "ms_per_K_refills": plain_ms_batch,
            "us_per_refill": plain_ms_batch * 1e3 / K,
            "nnz_out_per_s": K * nnz / (plain_ms_batch * 1e-3),
        },
        "batched_vs_plain": plain_ms_batch / ms_step,
        "alloc_s": t_alloc,
        "cpu_baseline": cpu,
    }
    lib.sdl_btda_plan_destroy(h)
    print(json.dumps(out), flush=True)


def cpu_baseline(w, nnz, seconds):
    """Single-thread CPU path on a bounded sample: one alloc per pass, then
    fills with fresh d (same product as the GPU workload), until about
    `seconds` of fill time."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"error": str(e)}
    from sparsediff_linalg_amd import workloads
    ds = [workloads.uniform(50_000 + r, w["m"], 0.1, 10.0) for r in range(64)]
    kind = "reference" if O.ref_available() else "port"
    fill_s = 0.0
    count = 0
    alloc_s = None
    t_start = time.perf_counter()
    while fill_s < seconds and time.perf_counter() - t_start < 3 * seconds:
        if kind == "reference":
            a, fills, _ = O.ref_fill_timer(w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"],
                                            w["Bp"], w["Bi"], w["Bx"], ds)
        else:
            a, fills = O.port_fill_timer(w["m"], w["nA"], w["nB"], w["Ap"], w["Ai"], w["Ax"],
                                         w["Bp"], w["Bi"], w["Bx"], ds)
        alloc_s = a if alloc_s is None else min(alloc_s, a)
        fill_s += sum(fills)
        count += len(fills)
    return {
        "value": count * nnz / fill_s,
        "unit": "nnz-out/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"config 1 product, alloc (CSC mirror + Gustavson + plan) once per pass of 64 "
                   f"d vectors, {count} single-thread refills with fresh d; fill time "
                   f"{fill_s:.2f} s, alloc {alloc_s * 1e3:.1f} ms"),
        "us_per_refill": fill_s / count * 1e6,
    }


if __name__ == "__main__":
    main()
