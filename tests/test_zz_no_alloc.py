"""lym_dot_noalloc (the allocation-free variant of lym_dot, static device
scratch) against lym_dot through the C-ABI: two calls in a row, the second
on a smaller input, complex128 and complex64; the same kernels and grid give
bitwise the same sums."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,prec", [(torch.complex128, 0), (torch.complex64, 1)])
def test_dot_noalloc_matches_dot(dtype, prec):
    from tfleonardym_amd import _lib
    lib = _lib.lib()
    dev = _lib.require_device()
    g = torch.Generator(device="cpu").manual_seed(11 + prec)
    stream = _lib.stream_handle(dev)
    for n in (3_000_000, 40_001):  # the second call on a smaller input
        x = torch.randn(n, dtype=dtype, generator=g).to(dev)
        y = torch.randn(n, dtype=dtype, generator=g).to(dev)
        res_dev = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
        _lib.check(lib.lym_dot_noalloc(prec, n, _lib.ptr(x), _lib.ptr(y), _lib.ptr(res_dev), stream),
                   "dot_noalloc")
        torch.cuda.synchronize()
        got = res_dev.cpu().numpy()
        want = (ctypes.c_double * 2)()
        scratch = torch.empty(int(lib.lym_reduce_scratch_bytes(n)), dtype=torch.uint8, device=dev)
        _lib.check(lib.lym_dot(prec, n, _lib.ptr(x), _lib.ptr(y), want, _lib.ptr(scratch), stream), "dot")
        assert got[0] == want[0] and got[1] == want[1], (n, got, list(want))
        ref = np.vdot(x.cpu().numpy().astype(np.complex128), y.cpu().numpy().astype(np.complex128))
        assert abs(complex(got[0], got[1]) - ref) <= 1e-10 * max(1.0, abs(ref))


def test_dot_noalloc_rejects_bad_arguments():
    from tfleonardym_amd import _lib
    lib = _lib.lib()
    dev = _lib.require_device()
    x = torch.zeros(4, dtype=torch.complex128, device=dev)
    rc = lib.lym_dot_noalloc(0, -1, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), _lib.stream_handle(dev))
    assert rc == -1  # LYM_ERR_ARG
    rc = lib.lym_dot_noalloc(2, 4, _lib.ptr(x), _lib.ptr(x), _lib.ptr(x), _lib.stream_handle(dev))
    assert rc == -1
