"""Global complex inner product (algebra_utils.py:5-11).

``dot(v1, v2) = sum conj(v1) v2`` over all indices, accumulated in FP64 on
the device by the deterministic reduction kernel of liblym.  Raises
RuntimeError on a shape mismatch, as the reference does; returns a Python
complex (the reference returns a scalar tensor that callers use the same way).
"""
import ctypes

import torch

from . import _lib
from .tensors import to_device


def dot(v1, v2):
    if tuple(v1.shape) != tuple(v2.shape):
        raise RuntimeError("Incompatible vector shapes!")
    a = to_device(v1)
    b = to_device(v2, device=a.device)
    out = (ctypes.c_double * 2)()
    scratch = _lib.reduce_scratch(a.device)
    _lib.check(_lib.lib().lym_dot(_lib.PREC_F64, a.numel(), _lib.ptr(a), _lib.ptr(b), out,
                                  _lib.ptr(scratch), _lib.stream_handle(a.device)), "dot")
    return complex(out[0], out[1])


def axpby(a, x, b, y):
    """y <- a x + b y in place (complex scalars), on the device.  x and y
    must be distinct tensors (the kernel takes them as non-aliasing)."""
    if x.data_ptr() == y.data_ptr():
        raise RuntimeError("axpby: x and y must not alias")
    a = complex(a)
    b = complex(b)
    _lib.check(_lib.lib().lym_axpby(_lib.PREC_F64, y.numel(), a.real, a.imag, _lib.ptr(x), b.real, b.imag,
                                    _lib.ptr(y), _lib.stream_handle(y.device)), "axpby")
    return y


def lincomb(a, x, b, y):
    """a x + b y as a new tensor."""
    out = y.clone()
    return axpby(a, x, b, out)
