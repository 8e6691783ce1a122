"""SU(N) projection of a link field, drop-in for unitary_sqrt_projection
(smearing.py:18-31): U = X (X^dag X)^-1/2 / det(...)^(1/N) per link, in the
lym_unitary_projection kernel.  The APE smearing of that module is a gauge
observable outside the fermion path and is not part of this package."""
import torch

from . import _lib
from .tensors import to_device


def unitary_sqrt_projection(link_config):
    X = to_device(link_config)
    if X.dim() != 4 or X.shape[1] != X.shape[2]:
        raise RuntimeError("link field must have shape (D, Nc, Nc, V)")
    out = torch.empty_like(X)
    _lib.check(_lib.lib().lym_unitary_projection(int(X.shape[0]), int(X.shape[1]), int(X.shape[3]), _lib.ptr(X),
                                                 _lib.ptr(out), _lib.stream_handle(X.device)), "unitary_projection")
    return out
