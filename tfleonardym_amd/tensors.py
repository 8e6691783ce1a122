"""Field handling at the Python boundary: any array-like (torch tensor on any
device, NumPy array, or an object exporting DLPack) becomes a contiguous
device tensor of the reference's layout and dtype."""
import numpy as np
import torch

from . import _lib


def to_device(v, dtype=torch.complex128, device=None):
    """Contiguous tensor on the HIP device (zero-copy when already there)."""
    dev = device or _lib.require_device()
    if isinstance(v, torch.Tensor):
        t = v
    elif isinstance(v, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(v))
    elif hasattr(v, "__dlpack__"):
        t = torch.from_dlpack(v)
    else:
        t = torch.as_tensor(np.asarray(v))
    if not t.is_complex():
        t = t.to(torch.float64)
    return t.to(device=dev, dtype=dtype).contiguous()


def numel_complex(t):
    return int(t.numel())


def to_numpy(t):
    return t.detach().cpu().numpy()
