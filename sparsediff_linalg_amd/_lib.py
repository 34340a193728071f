"""ctypes binding of libsdl_amd.so, the MI355X drop-in for sparsediff_linalg's
sparse-product hot path.

The library is the product: every compute entry point runs HIP kernels on the
GPU.  There is no CPU fallback; loading fails loudly when the built library is
missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdl_amd.so")

_lib = None

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
c_ll_p = C.POINTER(C.c_longlong)


class CSR_matrix(C.Structure):
    """Layout of the reference's CSR_matrix (include/sparsediff_linalg/CSR_matrix.h:31-39)."""

    _fields_ = [("p", c_int_p), ("i", c_int_p), ("x", c_double_p),
                ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int)]


class CSC_matrix(C.Structure):
    """Layout of the reference's CSC_matrix (include/sparsediff_linalg/CSC_matrix.h:33-41)."""

    _fields_ = [("p", c_int_p), ("i", c_int_p), ("x", c_double_p),
                ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int)]


class COO_matrix(C.Structure):
    """Layout of the reference's COO_matrix (include/sparsediff_linalg/COO_matrix.h:35-44)."""

    _fields_ = [("rows", c_int_p), ("cols", c_int_p), ("x", c_double_p),
                ("value_map", c_int_p), ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int)]


class matrix(C.Structure):
    pass


class sdl_refill_cfg(C.Structure):
    _fields_ = [("variant", C.c_int), ("rchunk", C.c_int)]


class sdl_sym_tune(C.Structure):
    _fields_ = [("hash_max", C.c_int), ("dens", C.c_int), ("rowacc_cut", C.c_int),
                ("rank_cut", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_to_csr_fn = C.CFUNCTYPE(C.POINTER(CSR_matrix), C.POINTER(matrix))
_gather_alloc_fn = C.CFUNCTYPE(C.POINTER(matrix), C.POINTER(matrix), c_int_p, C.c_int)
_gather_fill_fn = C.CFUNCTYPE(None, C.POINTER(matrix), C.POINTER(matrix))
matrix._fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int), ("x", c_double_p),
                   ("kind", C.c_int), ("values_version", C.c_uint64),
                   ("to_csr", _to_csr_fn), ("row_gather_alloc", _gather_alloc_fn),
                   ("row_gather_fill_values", _gather_fill_fn),
                   ("row_reduce_alloc", _gather_alloc_fn),
                   ("row_reduce_fill_values", _gather_fill_fn)]

MATRIX_SPARSE, MATRIX_PERMUTED_DENSE, MATRIX_STACKED_PD = 0, 1, 2

CSR_p = C.POINTER(CSR_matrix)
CSC_p = C.POINTER(CSC_matrix)
COO_p = C.POINTER(COO_matrix)
MAT_p = C.POINTER(matrix)
vp = C.c_void_p

# name -> (restype, argtypes); every symbol here is declared in include/*.h
SIGNATURES = {
    # runtime (sdl_device.h)
    "sdl_init": (None, []),
    "sdl_last_error": (C.c_char_p, []),
    "sdl_library_stream": (vp, []),
    "sdl_synchronize": (None, []),
    "sdl_device_alloc_count": (C.c_longlong, []),
    "sdl_device_free_count": (C.c_longlong, []),
    "sdl_dev_alloc": (vp, [C.c_size_t]),
    "sdl_dev_free": (None, [vp]),
    "sdl_memcpy_htod": (C.c_int, [vp, vp, C.c_size_t]),
    "sdl_memcpy_dtoh": (C.c_int, [vp, vp, C.c_size_t]),
    "sdl_memset_d": (C.c_int, [vp, C.c_int, C.c_size_t]),
    "sdl_event_create": (vp, []),
    "sdl_event_record": (None, [vp, vp]),
    "sdl_event_elapsed_ms": (C.c_float, [vp, vp]),
    "sdl_event_destroy": (None, [vp]),
    "sdl_device_name": (C.c_int, [C.c_char_p, C.c_int]),
    "sdl_copy_kernel": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "sdl_write_kernel": (C.c_int, [vp, C.c_size_t, C.c_int, vp]),
    "sdl_mfma_kernel": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp]),
    "sdl_write_rows_kernel": (C.c_int, [vp, C.c_longlong, C.c_int, C.c_longlong, C.c_int, vp]),
    # device-resident B^T D A plans
    "sdl_btda_plan_create": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int,
                                       C.POINTER(vp)]),
    "sdl_btda_plan_destroy": (None, [vp]),
    "sdl_btda_plan_nnz": (C.c_longlong, [vp]),
    "sdl_btda_plan_terms": (C.c_longlong, [vp]),
    "sdl_btda_plan_is_pull": (C.c_int, [vp]),
    "sdl_btda_plan_rowptr": (vp, [vp]),
    "sdl_btda_plan_colind": (vp, [vp]),
    "sdl_btda_plan_structure": (C.c_int, [vp, c_int_p, c_int_p]),
    "sdl_btda_plan_stats": (C.c_int, [vp, c_ll_p]),
    "sdl_btda_plan_schedule": (C.c_int, [vp, c_ll_p]),
    "sdl_btda_refill": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "sdl_btda_refill_batched": (C.c_int, [vp, vp, vp, vp, C.c_longlong, C.c_int, vp,
                                          C.c_longlong, vp]),
    "sdl_btda_refill_batched_cfg": (C.c_int, [vp, vp, vp, vp, C.c_longlong, C.c_int, vp,
                                              C.c_longlong, C.POINTER(sdl_refill_cfg), vp]),
    "sdl_refill_candidates": (C.c_int, [C.POINTER(sdl_refill_cfg), C.c_int]),
    "sdl_btda_autotune": (C.c_int, [vp, vp, vp, vp, C.c_longlong, C.c_int, vp, C.c_longlong,
                                    C.c_int, C.POINTER(sdl_refill_cfg), C.POINTER(C.c_float)]),
    "sdl_btda_set_refill_cfg": (C.c_int, [vp, C.POINTER(sdl_refill_cfg)]),
    "sdl_sym_tune_default": (C.c_int, [C.POINTER(sdl_sym_tune)]),
    "sdl_btda_plan_create_tuned": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int,
                                             C.POINTER(sdl_sym_tune), C.POINTER(vp)]),
    "sdl_btda_plan_tune": (C.c_int, [vp, C.POINTER(sdl_sym_tune)]),
    "sdl_btda_tune": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp,
                                C.c_int, C.POINTER(sdl_sym_tune), C.POINTER(C.c_float),
                                C.POINTER(C.c_float)]),
    "sdl_btda_tune_record_refill": (C.c_int, [vp, C.POINTER(sdl_refill_cfg)]),
    "sdl_tune_table_size": (C.c_int, []),
    "sdl_tune_table_clear": (None, []),
    "sdl_tune_table_save": (C.c_int, [C.c_char_p]),
    "sdl_tune_table_load": (C.c_int, [C.c_char_p]),
    "sdl_btda_plan_reserve": (C.c_int, [vp, C.c_int]),
    "sdl_btda_bind_values": (C.c_int, [vp, vp, vp, vp]),
    "sdl_btda_refill_batched_ex": (C.c_int, [vp, vp, vp, vp, C.c_longlong, C.c_int, vp,
                                             C.c_longlong, C.c_uint, vp, vp, vp]),
    "sdl_btda_refill_range": (C.c_int, [vp, vp, vp, vp, vp, C.c_longlong, C.c_longlong, vp]),
    "sdl_btda_refill_batched_static": (C.c_int, [vp, vp, vp, vp, C.c_longlong, C.c_int, vp,
                                                 C.c_longlong, vp]),
    "sdl_static_scratch_bytes": (C.c_longlong, []),
    "sdl_hessian_create": (C.c_int, [C.c_int, C.c_int, c_int_p, C.POINTER(vp), C.POINTER(vp),
                                     C.POINTER(vp)]),
    "sdl_hessian_destroy": (None, [vp]),
    "sdl_hessian_set_streams": (C.c_int, [vp, C.c_int]),
    "sdl_hessian_nnz": (C.c_longlong, [vp]),
    "sdl_hessian_rowptr": (vp, [vp]),
    "sdl_hessian_colind": (vp, [vp]),
    "sdl_hessian_lower_nnz": (C.c_longlong, [vp]),
    "sdl_hessian_lower_rows": (vp, [vp]),
    "sdl_hessian_lower_cols": (vp, [vp]),
    "sdl_hessian_eval": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), vp, vp, vp]),
    "sdl_hessian_capture": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), vp, vp, vp]),
    "sdl_hessian_replay": (C.c_int, [vp, vp]),
    "sdl_coo_plan_create": (C.c_int, [C.c_longlong, vp, vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "sdl_coo_plan_nnz": (C.c_longlong, [vp]),
    "sdl_coo_plan_rowptr": (vp, [vp]),
    "sdl_coo_plan_colind": (vp, [vp]),
    "sdl_coo_fill": (C.c_int, [vp, vp, vp, vp]),
    "sdl_coo_plan_destroy": (None, [vp]),
    "sdl_sum_plan_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(vp),
                                      c_ll_p, C.c_int, C.POINTER(vp)]),
    "sdl_sum_plan_nnz": (C.c_longlong, [vp]),
    "sdl_sum_plan_rowptr": (vp, [vp]),
    "sdl_sum_plan_colind": (vp, [vp]),
    "sdl_sum_plan_map": (vp, [vp]),
    "sdl_sum_plan_is_masked": (C.c_int, [vp]),
    "sdl_sum_fill": (C.c_int, [vp, C.POINTER(vp), vp, vp, vp]),
    "sdl_sum_add": (C.c_int, [vp, C.c_int, vp, vp, vp]),
    "sdl_sum_plan_destroy": (None, [vp]),
    "sdl_lower_coo_create": (C.c_int, [C.c_int, vp, vp, C.POINTER(vp)]),
    "sdl_lower_coo_nnz": (C.c_longlong, [vp]),
    "sdl_lower_coo_rows": (vp, [vp]),
    "sdl_lower_coo_cols": (vp, [vp]),
    "sdl_lower_coo_value_map": (vp, [vp]),
    "sdl_lower_coo_refresh": (C.c_int, [vp, vp, vp, vp]),
    "sdl_lower_coo_destroy": (None, [vp]),
    "sdl_y_kron_i_vec": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
    "sdl_i_kron_xt_vec": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]),
    "sdl_graph_begin": (C.c_int, [vp]),
    "sdl_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "sdl_graph_launch": (C.c_int, [vp, vp]),
    "sdl_graph_destroy": (None, [vp]),
    # learned row binning
    "sdl_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(vp)]),
    "sdl_mlp_destroy": (None, [vp]),
    "sdl_mlp_num_params": (C.c_int, []),
    "sdl_mlp_get_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "sdl_mlp_set_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "sdl_mlp_set_tau": (C.c_int, [vp, C.c_float]),
    "sdl_mlp_forward": (C.c_int, [vp, vp, C.c_longlong, vp]),
    "sdl_mlp_predict": (C.c_int, [vp, vp, C.c_longlong, vp]),
    "sdl_mlp_gradients": (C.c_int, [vp, vp, vp, C.c_longlong, vp, C.POINTER(C.c_float)]),
    "sdl_mlp_train": (C.c_int, [vp, vp, vp, C.c_longlong, C.c_int, C.c_int, C.c_float,
                                C.c_uint64, C.POINTER(C.c_float)]),
    "sdl_mlp_save": (C.c_int, [vp, C.c_char_p]),
    "sdl_mlp_load": (C.c_int, [vp, C.c_char_p]),
    "sdl_set_row_predictor": (C.c_int, [vp]),
    "sdl_btda_plan_create_predicted": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                                 C.POINTER(vp)]),
    "sdl_btda_plan_features": (C.c_int, [vp, vp]),
    "sdl_btda_plan_row_bins": (C.c_int, [vp, vp]),
    "sdl_btda_plan_pred_stats": (C.c_int, [vp, c_ll_p]),
    # device refills of block-kind product outputs (matrix * passed as pointers)
    "sdl_block_refill_device": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "sdl_block_refill_device_split": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "sdl_block_stats": (C.c_int, [vp, c_double_p]),
    "sdl_block_device_values": (vp, [vp]),
    "sdl_block_publish": (C.c_int, [vp]),
    "sdl_block_gemm_slice": (C.c_int, [C.c_int]),
    "sdl_btda_plan_single_refill": (C.c_int, [vp, C.c_int]),
    "sdl_btda_plan_reserve_tile": (C.c_int, [vp, C.c_int]),
    "sdl_sum_plan_fill_kind": (C.c_int, [vp]),
    "sdl_sum_scatter_ratio": (C.c_double, [C.c_double]),
    "sdl_block_coalesce_mode": (C.c_int, [C.c_int]),
    "sdl_btda_plan_window_rows": (C.c_int, [vp, c_ll_p]),
    "sdl_symbolic_window": (C.c_int, [C.c_int]),
    "sdl_symbolic_window_lanes": (C.c_int, [C.c_int]),
    "sdl_rowacc_window": (C.c_int, [C.c_int]),
    "sdl_symbolic_lane_groups": (C.c_int, [C.c_int]),
    # generators
    "sdl_gen_row_lengths_uniform": (C.c_longlong, [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                                   c_int_p]),
    "sdl_gen_fill_rows": (None, [C.c_uint64, C.c_int, C.c_int, c_int_p, c_int_p, c_double_p]),
    "sdl_gen_uniform": (None, [C.c_uint64, C.c_longlong, C.c_double, C.c_double, c_double_p]),
    "sdl_gen_normal": (None, [C.c_uint64, C.c_longlong, c_double_p]),
}


def _declare(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load libsdl_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(this package has no CPU path)")
        lib = C.CDLL(LIB_PATH, mode=os.RTLD_LOCAL)
        _declare(lib, SIGNATURES)
        _declare_extra(lib)
        _lib = lib
    return _lib


def _declare_extra(lib):
    from . import _api
    _declare(lib, _api.SIGNATURES)


def ptr(a: np.ndarray, ctype=None):
    """ctypes pointer to a contiguous numpy array's data."""
    assert a.flags["C_CONTIGUOUS"]
    if ctype is None:
        ctype = {np.dtype(np.int32): C.c_int, np.dtype(np.float64): C.c_double,
                 np.dtype(np.int64): C.c_longlong, np.dtype(np.float32): C.c_float,
                 np.dtype(np.uint32): C.c_uint32}[a.dtype]
    return a.ctypes.data_as(C.POINTER(ctype))


class DeviceArray:
    """A device allocation owned by Python (plumbing for tests and bench)."""

    def __init__(self, shape, dtype):
        self.lib = load()
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = self.lib.sdl_dev_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError("sdl_dev_alloc failed")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            d.lib.sdl_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes)
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.lib.sdl_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes)
        return out

    def offset(self, elems):
        return self.ptr + int(elems) * self.dtype.itemsize

    def copy_from(self, a):
        """Overwrite the whole allocation with host array a (same size), in
        place: device pointers held elsewhere (a captured graph) stay valid."""
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"copy_from: {a.nbytes} bytes into {self.nbytes}")
        if a.nbytes:
            check(self.lib.sdl_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "htod")

    def free(self):
        if self.ptr:
            self.lib.sdl_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def last_error():
    return load().sdl_last_error().decode()


def check(rc, what=""):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")


def device_name():
    buf = C.create_string_buffer(256)
    load().sdl_device_name(buf, 256)
    return buf.value.decode()
