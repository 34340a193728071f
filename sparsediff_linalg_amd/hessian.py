"""Host side of the Hessian assembly across expression nodes
(sdl_hessian_* in include/sdl_device.h): the Newton-step loop of a caller such
as SparseDiffEngine, which refills every chain-rule node, sums the nodes and
refreshes the lower-triangular export for the solver, here as device launches
issued call by call or replayed from one captured HIP graph."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceArray, check, load


class Hessian:
    """H = sum_k A_k^T diag(d_k) A_k over nodes A_k (host CSR structures
    (Ap, Ai), m_k x n), and H's lower triangle."""

    def __init__(self, nodes, n):
        self.lib = lib = load()
        self.n = n
        self.k = len(nodes)
        self.dev = [(DeviceArray.from_numpy(np.ascontiguousarray(p, np.int32)),
                     DeviceArray.from_numpy(np.ascontiguousarray(i, np.int32))) for p, i in nodes]
        m = (C.c_int * self.k)(*[len(p) - 1 for p, _ in nodes])
        P = (C.c_void_p * self.k)(*[a.ptr for a, _ in self.dev])
        I_ = (C.c_void_p * self.k)(*[b.ptr for _, b in self.dev])
        h = C.c_void_p()
        check(lib.sdl_hessian_create(self.k, n, m, P, I_, C.byref(h)), "hessian create")
        self.h = h
        self.nnz = lib.sdl_hessian_nnz(h)
        self.lower_nnz = lib.sdl_hessian_lower_nnz(h)

    def set_streams(self, streams):
        """Streams the node refills fork over (1 = in order on one stream)."""
        check(self.lib.sdl_hessian_set_streams(self.h, int(streams)), "hessian streams")

    def _ptrs(self, arrs):
        return (C.c_void_p * self.k)(*[a.ptr for a in arrs]) if arrs is not None else None

    def eval(self, Ax, D, Hx, Lx=None, stream=None):
        """Ax, D: per-node device arrays of values and d; Hx / Lx outputs."""
        check(self.lib.sdl_hessian_eval(self.h, self._ptrs(Ax), self._ptrs(D), Hx.ptr,
                                        Lx.ptr if Lx is not None else None, stream), "eval")

    def capture(self, Ax, D, Hx, Lx=None, stream=None):
        check(self.lib.sdl_hessian_capture(self.h, self._ptrs(Ax), self._ptrs(D), Hx.ptr,
                                           Lx.ptr if Lx is not None else None, stream),
              "capture")

    def replay(self, stream=None):
        check(self.lib.sdl_hessian_replay(self.h, stream), "replay")

    def structure(self):
        """(Hp int64, Hi, lower rows, lower cols) on the host."""
        def dtoh(ptr, count, dt):
            out = np.empty(count, dt)
            if count:
                check(self.lib.sdl_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes), "dtoh")
            return out
        lib = self.lib
        return (dtoh(lib.sdl_hessian_rowptr(self.h), self.n + 1, np.int64),
                dtoh(lib.sdl_hessian_colind(self.h), self.nnz, np.int32),
                dtoh(lib.sdl_hessian_lower_rows(self.h), self.lower_nnz, np.int32),
                dtoh(lib.sdl_hessian_lower_cols(self.h), self.lower_nnz, np.int32))

    def close(self):
        if self.h:
            self.lib.sdl_hessian_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
