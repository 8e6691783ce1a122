"""ctypes binding of liblym.so (the C ABI declared in include/lym.h).

There is no fallback: if the library is missing or fails to load, importing
the operator modules raises, so nothing silently runs off the GPU path.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblym.so")

PREC_F64, PREC_F32, PREC_F16 = 0, 1, 2
REP_FUNDAMENTAL, REP_ADJOINT = 0, 1
MODE_PROJECTED, MODE_SIMPLE = 0, 1
LINKS_FULL, LINKS_SU = 0, 1
ABI_VERSION = 6  # LYM_ABI_VERSION of include/lym.h

OK = 0
NOT_CONVERGED = 1
BREAKDOWN = 2
STATUS_NAMES = {0: "running", 1: "converged", 2: "max_iterations", 3: "breakdown"}


class LymWilson(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("nc", ctypes.c_int32),
        ("representation", ctypes.c_int32),
        ("hermitian", ctypes.c_int32),
        ("dims", ctypes.c_int32 * 4),
        ("kappa", ctypes.c_double),
        ("bc_re", ctypes.c_double),
        ("bc_im", ctypes.c_double),
        ("links", ctypes.c_void_p),
        ("links_right", ctypes.c_void_p),
        ("links_f32", ctypes.c_void_p),
        ("links_right_f32", ctypes.c_void_p),
        ("link_format", ctypes.c_int32),
        ("precision_phase", ctypes.c_int32),
    ]


class LymTileChoice(ctypes.Structure):
    """lym_tile_choice: LDS tile (a0, a1, k; a0 = 0: untiled), hops in flight
    and the order in which an XCD's blocks walk the tiles."""
    _fields_ = [
        ("a0", ctypes.c_int32),
        ("a1", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("hif", ctypes.c_int32),
        ("nsite", ctypes.c_int32),
        ("tuned", ctypes.c_int32),
        ("walk", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("us", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LymKernelTime(ctypes.Structure):
    """lym_kernel_time: one kernel class of a timed solve."""
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("launches", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("total_us", ctypes.c_double),
        ("bytes", ctypes.c_double),
    ]


class LymSolveStats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("nshift", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
        ("final_rr", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("conv_iter", ctypes.c_int32 * 64),
        ("true_rel_resid", ctypes.c_double * 64),
        ("initial_rel_resid", ctypes.c_double * 64),
        ("refine_cycles", ctypes.c_int32 * 64),
        ("refine_iterations", ctypes.c_int32),
        ("precision_phase", ctypes.c_int32),
    ]

    def as_dict(self, nshift=None):
        n = self.nshift if nshift is None else nshift
        return {
            "status": STATUS_NAMES.get(self.status, str(self.status)),
            "iterations": int(self.iterations),
            "restarts": int(self.restarts),
            "final_rr": float(self.final_rr),
            "seconds": float(self.seconds),
            "conv_iter": [int(self.conv_iter[i]) for i in range(n)],
            "true_rel_resid": [float(self.true_rel_resid[i]) for i in range(n)],
            "initial_rel_resid": [float(self.initial_rel_resid[i]) for i in range(n)],
            "refine_cycles": [int(self.refine_cycles[i]) for i in range(n)],
            "refine_iterations": int(self.refine_iterations),
            "precision_phase": int(self.precision_phase),
        }


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes); every symbol include/lym.h declares
SIGNATURES = {
    "lym_abi_version": (_i32, []),
    "lym_init": (_i32, [_i32]),
    "lym_last_error": (ctypes.c_char_p, []),
    "lym_wilson_apply": (_i32, [ctypes.POINTER(LymWilson), _i32, _i32, _vp, _vp, _vp]),
    "lym_wilson_apply_square": (_i32, [ctypes.POINTER(LymWilson), _i32, _f64, _vp, _vp, _vp, _vp]),
    "lym_reduce_scratch_bytes": (_sz, [_i64]),
    "lym_dot": (_i32, [_i32, _i64, _vp, _vp, ctypes.POINTER(_f64), _vp, _vp]),
    "lym_dot_noalloc": (_i32, [_i32, _i64, _vp, _vp, _vp, _vp]),
    "lym_axpby": (_i32, [_i32, _i64, _f64, _f64, _vp, _f64, _f64, _vp, _vp]),
    "lym_gamma5": (_i32, [_i32, _i32, _i64, _vp, _vp, _vp]),
    "lym_spin_axpby": (_i32, [_i32, _i64, ctypes.POINTER(_f64), _vp, _vp, _vp, _vp]),
    "lym_convert": (_i32, [_i32, _vp, _i32, _vp, _i64, _vp]),
    "lym_links_f32_bytes": (_sz, [_i32, _i32, _i64]),
    "lym_pack_links_f32": (_i32, [_i32, _i32, _i64, _vp, _vp, _vp]),
    "lym_link_deviation": (_i32, [_vp, _i32, _vp, _vp]),
    "lym_lie_derivative": (_i32, [ctypes.POINTER(LymWilson), _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                  ctypes.POINTER(_f64), _i32, _i32, _vp, _vp]),
    "lym_slice_sums": (_i32, [ctypes.POINTER(LymWilson), _i32, _i32, ctypes.POINTER(_vp), _vp, _sz,
                              ctypes.POINTER(_f64), _vp]),
    "lym_unpack_ustar": (_i32, [_vp, _i64, _i32, _vp, _vp]),
    "lym_unitary_projection": (_i32, [_i32, _i32, _i64, _vp, _vp, _vp]),
    "lym_tune_wilson": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    "lym_tile_info": (_i32, [_vp, _i32, _vp]),
    "lym_solver_workspace_bytes": (_sz, [ctypes.POINTER(LymWilson), _i32, _i32]),
    "lym_mr_multishift": (_i32, [ctypes.POINTER(LymWilson), _i32, ctypes.POINTER(_f64), _i32, _vp,
                                 ctypes.POINTER(_vp), _i32, _f64, _f64, _vp, _sz,
                                 ctypes.POINTER(LymSolveStats), _vp]),
    "lym_cgm": (_i32, [ctypes.POINTER(LymWilson), _i32, _i32, ctypes.POINTER(_f64), _vp, ctypes.POINTER(_vp),
                       _f64, _i32, _vp, _sz, ctypes.POINTER(LymSolveStats), _vp]),
    "lym_bicgstab_workspace_bytes": (_sz, [ctypes.POINTER(LymWilson)]),
    "lym_bicgstab": (_i32, [ctypes.POINTER(LymWilson), _i32, _f64, _vp, _vp, _vp, _i32, _f64, _vp, _sz,
                            ctypes.POINTER(LymSolveStats), _vp]),
    "lym_kernel_timing": (_i32, [_i32]),
    "lym_kernel_timing_read": (_i32, [ctypes.POINTER(LymKernelTime), _i32, ctypes.POINTER(_i32)]),
    "lym_cgm_batched_workspace_bytes": (_sz, [ctypes.POINTER(LymWilson), _i32, _i32, _i32]),
    "lym_cgm_batched": (_i32, [ctypes.POINTER(LymWilson), _i32, _i32, _i32, ctypes.POINTER(_f64),
                               ctypes.POINTER(_vp), ctypes.POINTER(_vp), _f64, _i32, _vp, _sz,
                               ctypes.POINTER(LymSolveStats), _vp]),
}

_lib = None


def load(path=LIB_PATH):
    """Load liblym.so and bind every entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built: run `python -m tfleonardym_amd.build` (hipcc, gfx950). "
            "There is no CPU fallback for the Wilson operator.")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.lym_abi_version() != ABI_VERSION:
        raise ImportError("liblym.so ABI version mismatch; rebuild")
    _lib = lib
    return lib


def lib():
    return load()


def check(rc, what="liblym"):
    """Map a C status to Python: <0 raises RuntimeError, >=0 is returned."""
    if rc < 0:
        msg = load().lym_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg} (code {rc})")
    return rc


def require_device():
    if not torch.cuda.is_available():
        raise RuntimeError("the Wilson-fermion kernels need a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


# Largest |stored - rebuilt| link entry for which the kernels may rebuild the
# last row of every link (LINKS_SU).  Unitarised FP64 configurations sit near
# 1e-15; at 1e-14 the rebuilt rows change D psi by well under 1e-14 relative.
SU_TOLERANCE = 1e-14


def su_link_format(desc, device=None):
    """LINKS_SU if every complex128 link of `desc` is special unitary to
    SU_TOLERANCE (measured on the device by lym_link_deviation), else
    LINKS_FULL.  LYM_FULL_LINKS=1 in the environment forces LINKS_FULL."""
    if os.environ.get("LYM_FULL_LINKS") == "1":
        return LINKS_FULL
    dev = device if device is not None else require_device()
    out = torch.empty(1, dtype=torch.float64, device=dev)
    saved = desc.link_format
    desc.link_format = LINKS_FULL
    try:
        check(load().lym_link_deviation(ctypes.byref(desc), PREC_F64, ptr(out), stream_handle(dev)),
              "link deviation")
    finally:
        desc.link_format = saved
    return LINKS_SU if float(out.item()) <= SU_TOLERANCE else LINKS_FULL


def pack_links_f32(links):
    """complex64 copy of complex128 links (ndim, nc, nc, *geometry) in the
    kernels' packed layout (lym_pack_links_f32), as a flat float32 tensor on
    the links' device."""
    ndim, nc = int(links.shape[0]), int(links.shape[1])
    volume = int(links.numel() // (ndim * nc * nc))
    nbytes = int(load().lym_links_f32_bytes(ndim, nc, volume))
    if nbytes == 0:
        raise ValueError(f"pack_links_f32: unsupported links of shape {tuple(links.shape)}")
    out = torch.empty(nbytes // 4, dtype=torch.float32, device=links.device)
    src = links.contiguous()
    check(load().lym_pack_links_f32(ndim, nc, volume, ptr(src), ptr(out), stream_handle(links.device)),
          "pack links")
    return out


_scratch = {}


def reduce_scratch(device):
    """Small per-device scratch tensor for lym_dot (reused across calls on the
    current stream)."""
    key = (device.index, torch.cuda.current_stream(device).cuda_stream)
    t = _scratch.get(key)
    if t is None:
        n = int(load().lym_reduce_scratch_bytes(0))
        t = torch.zeros(n + 512, dtype=torch.uint8, device=device)
        _scratch[key] = t
    return t


def kernel_timing(enable):
    """Start (clearing the counters) or stop the kernel timing of the fast
    multishift CG on this host thread (lym_kernel_timing)."""
    check(lib().lym_kernel_timing(1 if enable else 0), "kernel_timing")


def kernel_timing_read():
    """Per kernel class of the timed solves: {name: dict(launches, us,
    bytes)} (lym_kernel_timing_read)."""
    n = ctypes.c_int(0)
    check(lib().lym_kernel_timing_read(None, 0, ctypes.byref(n)), "kernel_timing_read")
    arr = (LymKernelTime * max(1, n.value))()
    check(lib().lym_kernel_timing_read(arr, n.value, ctypes.byref(n)), "kernel_timing_read")
    return {arr[i].name.decode(): dict(launches=int(arr[i].launches), us=float(arr[i].total_us),
                                       bytes=float(arr[i].bytes)) for i in range(n.value)}
