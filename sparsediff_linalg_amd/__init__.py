"""MI355X-native drop-in for sparsediff_linalg's sparse-product hot path.

The product is the C-ABI library libsdl_amd.so (include/sparsediff_linalg/*.h
for the reference-compatible API, include/sdl_device.h for the device-resident
extension).  This Python package is the host-side binding used by the tests
and the benchmark.
"""
from ._lib import (CSR_matrix, CSC_matrix, COO_matrix, matrix, DeviceArray, load, ptr,  # noqa: F401
                   check, last_error, device_name, LIB_PATH)
