"""The calibration kernels' argument checks (CPU: rejected before any device
call) and, on the GPU, that the FP64 MFMA loop does what its flop count says
(every accumulator of x * y summed over the loop) and that the streaming
kernels write what they claim."""
import ctypes as C

import numpy as np
import pytest

from sparsediff_linalg_amd import DeviceArray, load


def test_mfma_kernel_rejects_bad_arguments():
    lib = load()
    fake = C.c_void_p(16)  # never dereferenced: every call below is rejected first
    for nacc, blocks, iters in ((3, 1, 1), (0, 1, 1), (32, 1, 1), (4, 0, 1), (4, 1, 0)):
        assert lib.sdl_mfma_kernel(nacc, blocks, iters, fake, None) == 1
    assert lib.sdl_mfma_kernel(4, 1, 1, None, None) == 1


def test_write_kernels_reject_bad_arguments():
    lib = load()
    fake = C.c_void_p(16)
    assert lib.sdl_write_kernel(fake, 24, 0, None) == 1      # not a multiple of 16
    assert lib.sdl_write_kernel(fake, 32, 3, None) == 1      # no mode 3
    assert lib.sdl_copy_kernel(fake, fake, 24, None) == 1
    assert lib.sdl_write_rows_kernel(None, 8, 1, 8, 1, None) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("nacc", [1, 4, 16])
def test_mfma_kernel_value(nacc):
    """Every accumulator gains D = A B per round, A[i][k] the x of lane i + 16k
    and B[k][j] the y of lane j + 16k; lane l stores the sum over its
    accumulators of its four D entries (rows l / 16 + 4v, column l % 16):
    checked against a float64 model of that layout (also pinning the lane
    layout the block GEMM's epilogue assumes)."""
    lib = load()
    blocks, iters = 2, 3
    out = DeviceArray(blocks * 256, np.float64)
    assert lib.sdl_mfma_kernel(nacc, blocks, iters, out.ptr, None) == 0
    lib.sdl_synchronize()
    got = out.numpy()
    assert np.all(np.isfinite(got))
    # lane t of a wave holds x_t = 1 + t e-9, y_t = 1 - t e-9 (t = threadIdx.x);
    # D[i][j] += sum_k A[i][k] B[k][j] with A[i][k] = x of lane i + 16 k and
    # B[k][j] = y of lane j + 16 k; lane l holds D rows l/16 + 4v, column l%16
    tid = np.arange(256)
    x = 1.0 + tid * 1e-9
    y = 1.0 - tid * 1e-9
    want = np.empty(256)
    for t in range(256):
        wbase = (t // 64) * 64
        l = t % 64
        j = l % 16
        s = 0.0
        for v in range(4):
            i = l // 16 + 4 * v
            s += sum(x[wbase + i + 16 * k] * y[wbase + j + 16 * k] for k in range(4))
        want[t] = s * iters * nacc
    assert np.allclose(got[:256], want, rtol=1e-12)
    assert np.allclose(got[256:], want, rtol=1e-12)
