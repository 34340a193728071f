"""Host side of the learned row binning: build a training set from exact
symbolic plans (features and exact-count bins computed on the GPU), train the
16 -> 32 -> 6 MLP with the library's kernels, save / load its parameters.

    python -m sparsediff_linalg_amd.predictor --out row_predictor.bin

trains on the products of the sparse B^T D A workload (20000 x 5000, 6
nnz/row), the A^T A of rank-Zipf(1.2) matrices (mean row length 16, several
sizes) and the A^T A of the stacked permuted-dense operand read as CSR, and
writes the parameters.  With --structured the set also holds the A^T A of
banded matrices (bandwidths 4, 16, 48, 96 on 2^15 rows), of strided bands
(2^16 rows of 65 or 129 consecutive columns starting 2 or 8 columns apart, in
2^17 or 2^21 columns) and of an R-MAT graph (2^13 nodes): structured column patterns whose nnz follows the
column span rather than the uniform collision estimate (the file shipped beside this
module is made this way).
"""
from __future__ import annotations

import argparse
import ctypes as C
import os

import numpy as np

from ._lib import DeviceArray, check, load

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "row_predictor.bin")
NIN = 16


class Plan:
    """A device B^T diag(d) A plan (A^T A when B is A)."""

    def __init__(self, A, B=None, nA=None, nB=None, mode=0, mlp=None, tune=None):
        lib = load()
        Ap, Ai = A[0], A[1]
        m = len(Ap) - 1
        Bp, Bi = (B[0], B[1]) if B is not None else (Ap, Ai)
        self.nA = int(nA)
        self.nB = int(nB if nB is not None else nA)
        self.arrays = [DeviceArray.from_numpy(np.ascontiguousarray(a, np.int32))
                       for a in (Ap, Ai, Bp, Bi)]
        dAp, dAi, dBp, dBi = (a.ptr for a in self.arrays)
        h = C.c_void_p()
        if tune is not None:
            from .tuning import sym_tune
            check(lib.sdl_btda_plan_create_tuned(m, self.nA, self.nB, dAp, dAi, dBp, dBi, mode,
                                                 C.byref(sym_tune(tune)), C.byref(h)),
                  "tuned plan")
        elif mlp is not None or mode == 1:
            check(lib.sdl_btda_plan_create_predicted(m, self.nA, self.nB, dAp, dAi, dBp, dBi,
                                                     mlp.handle if mlp else None, C.byref(h)),
                  "predicted plan")
        else:
            check(lib.sdl_btda_plan_create(m, self.nA, self.nB, dAp, dAi, dBp, dBi, mode,
                                           C.byref(h)), "plan")
        self.h = h
        self.lib = lib

    @property
    def rows(self):
        return self.nB

    def nnz(self):
        return self.lib.sdl_btda_plan_nnz(self.h)

    def structure(self):
        nnz = self.nnz()
        p = np.empty(self.nB + 1, np.int32)
        i = np.empty(max(nnz, 1), np.int32)
        check(self.lib.sdl_btda_plan_structure(self.h, p.ctypes.data_as(C.POINTER(C.c_int)),
                                               i.ctypes.data_as(C.POINTER(C.c_int))), "structure")
        return p, i[:nnz]

    def features(self):
        F = DeviceArray((self.nB, NIN), np.float32)
        check(self.lib.sdl_btda_plan_features(self.h, F.ptr), "features")
        return F

    def bins(self):
        Y = DeviceArray(self.nB, np.int32)
        check(self.lib.sdl_btda_plan_row_bins(self.h, Y.ptr), "bins")
        return Y

    def schedule(self):
        """The refill schedule: pull, by-rank rows, recorded positions, listed
        multi-term items, rows left to the dense accumulator."""
        o = (C.c_longlong * 5)()
        check(self.lib.sdl_btda_plan_schedule(self.h, o), "schedule")
        return dict(pull=bool(o[0]), rank_rows=o[1], rank_positions=bool(o[2]), items=o[3],
                    dense_rows=o[4])

    def tune(self):
        """The symbolic / refill thresholds this plan was created with."""
        from ._lib import sdl_sym_tune
        t = sdl_sym_tune()
        check(self.lib.sdl_btda_plan_tune(self.h, C.byref(t)), "plan tune")
        return t.as_dict()

    def pred_stats(self):
        o = (C.c_longlong * 5)()
        check(self.lib.sdl_btda_plan_pred_stats(self.h, o), "pred stats")
        return dict(predicted=o[0], overflow=o[1], misbin=o[2], exact=o[3], candidates=o[4])

    def close(self):
        if self.h:
            self.lib.sdl_btda_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Predictor:
    """The library's row-size MLP."""

    def __init__(self, seed=1):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.sdl_mlp_create(seed, C.byref(h)), "mlp create")
        self.handle = h

    @classmethod
    def from_file(cls, path=DEFAULT_PATH):
        p = cls()
        check(p.lib.sdl_mlp_load(p.handle, path.encode()), "mlp load")
        return p

    def params(self):
        out = np.empty(self.lib.sdl_mlp_num_params(), np.float32)
        check(self.lib.sdl_mlp_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))),
              "get params")
        return out

    def set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        check(self.lib.sdl_mlp_set_params(self.handle, P.ctypes.data_as(C.POINTER(C.c_float))),
              "set params")

    def set_tau(self, tau):
        check(self.lib.sdl_mlp_set_tau(self.handle, tau), "set tau")

    def train(self, X, Y, n, epochs=40, batch=4096, lr=3e-3, seed=7):
        loss = C.c_float()
        check(self.lib.sdl_mlp_train(self.handle, X.ptr, Y.ptr, n, epochs, batch, lr, seed,
                                     C.byref(loss)), "train")
        return loss.value

    def predict(self, X, n):
        out = DeviceArray(n, np.uint8)
        check(self.lib.sdl_mlp_predict(self.handle, X.ptr, n, out.ptr), "predict")
        return out.numpy()

    def save(self, path):
        check(self.lib.sdl_mlp_save(self.handle, path.encode()), "mlp save")

    def close(self):
        if self.handle:
            self.lib.sdl_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_products(small=False, structured=False):
    """(name, A, B, nA, nB) products whose exact plans label the training set:
    the sparse B^T D A workload, A^T A of rank-Zipf matrices and of the
    stacked permuted-dense operand as CSR; with `structured`, also A^T A of
    banded matrices and of an R-MAT graph."""
    from . import workloads as W
    out = []
    w1 = W.btda_20000x5000(1, m=4000 if small else 20000, n=1000 if small else 5000)
    out.append(("btda", (w1["Ap"], w1["Ai"]), (w1["Bp"], w1["Bi"]), w1["nA"], w1["nB"]))
    for m, n in ((20000, 5000), (50000, 20000)) if small else ((200000, 50000), (100000, 10000),
                                                               (400000, 200000)):
        p, i, _ = W.zipf_csr(2 + m // 1000, m, n)
        out.append((f"zipf{m}x{n}", (p, i), None, n, n))
    wb = W.stacked_pd_blocks(3, nblocks=32 if small else 256)
    p, i, _ = W.blocks_as_csr(wb)
    out.append(("blocks", (p, i), None, wb["n"], wb["n"]))
    if structured:
        nb = 1 << (13 if small else 15)
        for bw in (4, 16, 48, 96):
            p, i, _ = W.banded_csr(nb, bw)
            out.append((f"banded{bw}", (p, i), None, nb, nb))
        for width, stride in ((65, 2), (129, 8)):
            for ncols in (4 * nb, 64 * nb):
                p, i, _ = W.strided_band_csr(2 * nb, ncols, width, stride)
                out.append((f"strided{width}x{stride}n{ncols}", (p, i), None, ncols, ncols))
        sc = 11 if small else 13
        p, i, _ = W.rmat_csr(sc, seed=3)
        out.append((f"rmat{sc}", (p, i), None, 1 << sc, 1 << sc))
    return out


def training_set(products):
    """Features and labels of the products' non-empty output rows (device
    arrays; rows without terms take no accumulator)."""
    Fs, Ys = [], []
    for _, A, B, nA, nB in products:
        pl = Plan(A, B, nA, nB)
        F = pl.features().numpy()
        keep = F[:, 0] > 0  # log2(1 + term bound)
        Fs.append(F[keep])
        Ys.append(pl.bins().numpy()[keep])
        pl.close()
    X = np.ascontiguousarray(np.concatenate(Fs), np.float32)
    Y = np.ascontiguousarray(np.concatenate(Ys), np.int32)
    return DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y), len(X)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--structured", action="store_true")
    args = ap.parse_args()
    X, Y, n = training_set(training_products(args.small, args.structured))
    mlp = Predictor(seed=1)
    loss = mlp.train(X, Y, n, epochs=args.epochs)
    pred = mlp.predict(X, n)
    y = Y.numpy()
    print(f"rows {n}, final loss {loss:.4f}, exact bin {np.mean(pred == y):.3f}, "
          f"under-predicted {np.mean(pred < y):.4f}, label counts {np.bincount(y, minlength=6)}")
    mlp.save(args.out)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
