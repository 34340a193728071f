"""ctypes signatures of the reference-compatible drop-in entry points
(include/sparsediff_linalg/*.h)."""
import ctypes as C

from ._lib import CSR_p, CSC_p, COO_p, MAT_p, c_int_p, c_double_p  # noqa: F401

SIGNATURES = {}
