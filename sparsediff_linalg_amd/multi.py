"""Many independent B^T diag(d) A instances refilled together.

The instances form one block-diagonal product: output rows (columns of B)
are grouped by instance, so one plan's symbolic pass serves every instance
and each refill is one launch over all of them.  Instances are split into
contiguous groups whose term and entry counts stay within int32 (one plan
each).  K rounds (fresh d vectors) of a group are one launch of the
interleaved refill, reading the d vectors from a k-major pool (row k holds
every round's d[k]) and writing a ring of K output slots, so each term's
operands are read once per K rounds; all launches of all groups are captured
into one HIP graph and replayed as one launch.  The per-call path for
comparison refills one instance for one d per launch (the instance's entry
range of the same plan).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceArray, check, load, sdl_refill_cfg

BOUND = 1  # SDL_BATCH_VALUES_BOUND
KMAJOR = 2  # SDL_BATCH_D_KMAJOR
PLAN_PULL = 256  # SDL_PLAN_PULL: term lists whatever the terms / nnz ratio


def kmajor_ld(rounds):
    """Row stride of a k-major pool of `rounds` d vectors: the refill reads
    runs of up to 16 rounds from any even start, so 16 past the rounded end."""
    return (rounds + 15) // 16 * 16 + 16


class Group:
    def __init__(self, lib, w, i0, i1, K):
        self.i0, self.i1 = i0, i1
        r0, r1 = int(w["roff"][i0]), int(w["roff"][i1])
        c0, c1 = int(w["coff"][i0]), int(w["coff"][i1])
        self.r0, self.m, self.c0, self.n = r0, r1 - r0, c0, c1 - c0
        dev = []
        for tag in ("A", "B"):
            p, i, x = w[tag]
            a, b = int(p[r0]), int(p[r1])
            dev.append((DeviceArray.from_numpy((p[r0:r1 + 1] - a).astype(np.int32)),
                        DeviceArray.from_numpy((i[a:b] - c0).astype(np.int32)),
                        DeviceArray.from_numpy(np.ascontiguousarray(x[a:b]))))
        self.A, self.B = dev
        h = C.c_void_p()
        check(lib.sdl_btda_plan_create(self.m, self.n, self.n, self.A[0].ptr, self.A[1].ptr,
                                       self.B[0].ptr, self.B[1].ptr, PLAN_PULL, C.byref(h)),
              "plan")
        self.h = h
        self.nnz = lib.sdl_btda_plan_nnz(h)
        self.terms = lib.sdl_btda_plan_terms(h)
        self.pull = bool(lib.sdl_btda_plan_is_pull(h))
        cp = np.empty(self.n + 1, np.int64)
        check(lib.sdl_memcpy_dtoh(cp.ctypes.data, lib.sdl_btda_plan_rowptr(h), cp.nbytes), "Cp")
        # entry range of every instance (its output rows are its B columns)
        self.ranges = [(int(cp[w["coff"][k] - c0]), int(cp[w["coff"][k + 1] - c0]))
                       for k in range(i0, i1)]
        self.ldc = (max(self.nnz, 1) + 63) // 64 * 64
        self.C = None

    def finish(self, lib, K, tile=False):
        """Refill workspace, bound operand values (operand pairs, or with tile
        the tile layout: each term's A value and each L entry's B value) and
        the ring of K output slots (after every group's plan, so no symbolic
        pass runs beside them)."""
        if self.pull:
            reserve = lib.sdl_btda_plan_reserve_tile if tile else lib.sdl_btda_plan_reserve
            check(reserve(self.h, K), "reserve")
            check(lib.sdl_btda_bind_values(self.h, self.A[2].ptr, self.B[2].ptr, None), "bind")
        self.C = DeviceArray(K * self.ldc, np.float64)  # ring of K output slots


class MultiProduct:
    def __init__(self, w, max_terms=1_200_000_000, K=8, tile=False):
        """tile: bind the operand values in tile layout (the refills then run
        the tile form of the interleaved kernel, reading d once per L entry
        of a tile instead of once per term)."""
        assert K == 1 or K % 2 == 0, "K must be 1 or even"
        self.tile = tile
        self.K = K
        self.lib = lib = load()
        self.w = w
        self.M = int(w["roff"][-1])
        # terms of every instance: sum over its rows of len_A(k) * len_B(k)
        la = np.diff(w["A"][0]).astype(np.int64)
        lb = np.diff(w["B"][0]).astype(np.int64)
        per_row = la * lb
        starts = w["roff"][:-1]
        inst_terms = np.add.reduceat(per_row, starts) if len(per_row) else np.zeros(0, np.int64)
        self.groups = []
        i0, acc = 0, 0
        for k, t in enumerate(inst_terms):
            if acc + t > max_terms and k > i0:
                self.groups.append((i0, k))
                i0, acc = k, 0
            acc += int(t)
        self.groups.append((i0, len(inst_terms)))
        self.g = [Group(lib, w, a, b, K) for a, b in self.groups]
        lib.sdl_synchronize()
        for g in self.g:
            g.finish(lib, K, tile)
        self.nnz = sum(g.nnz for g in self.g)
        self.terms = sum(g.terms for g in self.g)
        lib.sdl_synchronize()

    def autotune(self, D, reps=5):
        """Time the candidate refill launch shapes on every group (K refills
        from the round-major pool D, which must hold at least K rounds)."""
        best = []
        for g in self.g:
            cfg, us = sdl_refill_cfg(), C.c_float()
            check(self.lib.sdl_btda_autotune(g.h, g.A[2].ptr, g.B[2].ptr, D.offset(g.r0),
                                             self.M, self.K, g.C.ptr, g.ldc, reps,
                                             C.byref(cfg), C.byref(us)), "autotune")
            best.append((cfg.variant, cfg.rchunk, us.value))
        return best

    def refill_round(self, D, r, stream=None):
        """One refill of every group for round r of a round-major pool D
        (round r at D + r*M) into output slot 0 (one launch per group)."""
        for g in self.g:
            check(self.lib.sdl_btda_refill_batched_ex(
                g.h, g.A[2].ptr, g.B[2].ptr, D.offset(r * self.M + g.r0), self.M, 1, g.C.ptr,
                g.ldc, BOUND, None, None, stream), "refill")

    def refill_rounds(self, DT, ldt, r0, k, stream=None):
        """Rounds r0 .. r0+k-1 (k <= K) of every group from the k-major pool DT
        (row stride ldt) into output slots 0 .. k-1: one launch per group."""
        assert k <= self.K and r0 % 2 == 0
        for g in self.g:
            check(self.lib.sdl_btda_refill_batched_ex(
                g.h, g.A[2].ptr, g.B[2].ptr, DT.offset(g.r0 * ldt + r0), ldt, k, g.C.ptr, g.ldc,
                BOUND | KMAJOR, None, None, stream), "refill")

    def capture(self, DT, ldt, rounds, stream=None):
        """A graph of `rounds` refill rounds of every group, K per launch
        (slot j of the ring holds the latest round r with r % K == j)."""
        check(self.lib.sdl_graph_begin(stream), "graph begin")
        for r0 in range(0, rounds, self.K):
            self.refill_rounds(DT, ldt, r0, min(self.K, rounds - r0), stream)
        g = C.c_void_p()
        check(self.lib.sdl_graph_end(stream, C.byref(g)), "graph end")
        return g

    def plain_round(self, D, r, stream=None):
        """One launch per instance: its entry range refilled for d round r."""
        for g in self.g:
            d = D.offset(r * self.M + g.r0)
            for e0, e1 in g.ranges:
                if e1 > e0:
                    check(self.lib.sdl_btda_refill_range(g.h, g.A[2].ptr, g.B[2].ptr, d, g.C.ptr,
                                                         e0, e1, stream), "range refill")

    def values(self, slot=0):
        out = []
        for g in self.g:
            v = np.empty(g.nnz, np.float64)
            if g.nnz:
                self.lib.sdl_memcpy_dtoh(v.ctypes.data, g.C.offset(slot * g.ldc), v.nbytes)
            out.append(v)
        return out

    def close(self):
        for g in self.g:
            if g.h:
                self.lib.sdl_btda_plan_destroy(g.h)
                g.h = None
