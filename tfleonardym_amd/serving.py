"""Continuous-batching solve server for multishift Wilson solves.

Requests name a resident gauge field, a kappa, a shift list and a source;
the server returns x_s = (Q^2 + s)^-1 b for every shift at FP64 true relative
residual `tolerance`.  Behaviour:

* Gauge fields stay resident on the device (complex128 plus a complex64 copy).
  A request only brings its source.
* Admission is by slots and by HBM footprint.  Each request's footprint is
  its solver workspace + solutions + source, asked of liblym before it runs.
  Requests are admitted in arrival order, first fit: a request that does not
  fit yet waits while later, smaller ones may start.  A request waiting longer
  than `max_wait_s` blocks later admissions until it fits, so it cannot
  starve.
* Each slot is a worker thread with its own HIP stream.  ctypes releases the
  GIL during a solve, so solves of different requests overlap on the device,
  and a slot takes new work as soon as its solve ends.
* Every request has its own workspace and the kernels' reductions have a
  fixed order, so a result is bitwise the same whether the request runs alone
  or in a mix (tests/test_gpu_serving.py).

Latency is measured from `arrival` (set at submit unless given) to the moment
the result is ready on the host side (device synchronised).
"""
import collections
import ctypes
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .tensors import to_device


@dataclass
class SolveRequest:
    field: str
    kappa: float
    shifts: list
    source: object
    tolerance: float = 1e-10
    rid: int = -1
    arrival: float = None


@dataclass
class SolveResult:
    rid: int
    solutions: list
    stats: dict
    status: str
    arrival: float
    start: float
    done: float
    footprint: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.done - self.arrival


class GaugeStore:
    """Resident gauge fields keyed by name (4D/2D, fundamental representation)."""

    def __init__(self, device=None):
        self.device = device or _lib.require_device()
        self._entries = {}

    def add(self, key, gauge_field, geometry, bc_factor=None):
        U = to_device(gauge_field, device=self.device)
        if U.dim() != 4 or U.shape[1] != U.shape[2]:
            raise RuntimeError("gauge field must have shape (D, Nc, Nc, V)")
        geometry = tuple(int(L) for L in geometry)
        if int(np.prod(geometry)) != U.shape[-1] or len(geometry) != U.shape[0]:
            raise RuntimeError("geometry does not match the gauge field")
        U32 = torch.empty(U.shape, dtype=torch.complex64, device=self.device)
        _lib.check(_lib.lib().lym_convert(_lib.PREC_F64, _lib.ptr(U), _lib.PREC_F32, _lib.ptr(U32), U.numel(),
                                          _lib.stream_handle(self.device)), "convert")
        torch.cuda.synchronize(self.device)
        d = _lib.LymWilson()
        d.ndim = U.shape[0]
        d.nc = U.shape[1]
        d.representation = _lib.REP_FUNDAMENTAL
        d.hermitian = 1
        for mu, L in enumerate(geometry):
            d.dims[mu] = L
        bc = 1.0 if bc_factor is None else complex(bc_factor)
        d.bc_re = float(np.real(bc))
        d.bc_im = float(np.imag(bc))
        d.links = U.data_ptr()
        d.links_f32 = U32.data_ptr()
        d.link_format = _lib.su_link_format(d, self.device)
        self._entries[key] = dict(U=U, U32=U32, desc=d, geometry=geometry)
        return key

    def desc(self, key, kappa):
        e = self._entries[key]
        d = _lib.LymWilson()
        ctypes.memmove(ctypes.byref(d), ctypes.byref(e["desc"]), ctypes.sizeof(d))
        d.kappa = float(kappa)
        return d

    def spinor_shape(self, key):
        e = self._entries[key]
        ns = 4 if len(e["geometry"]) == 4 else 2
        return (ns, e["U"].shape[1], e["U"].shape[-1])

    def resident_bytes(self):
        return sum(e["U"].numel() * 16 + e["U32"].numel() * 8 for e in self._entries.values())

    def keys(self):
        return list(self._entries)


class SolveServer:
    """Continuous-admission multishift solve server (see module docstring)."""

    # "half": one-shift requests run the 16-bit-storage CG with FP64 reliable
    # updates, multishift requests the mixed CG-M (16-bit storage is built
    # for one shift only)
    _PREC = {"mixed": _lib.PREC_F32, "double": _lib.PREC_F64, "half": _lib.PREC_F16}

    def __init__(self, store, slots=4, hbm_budget_bytes=None, precision_mode="mixed", max_iter=50000,
                 max_wait_s=2.0):
        self.store = store
        self.slots = int(slots)
        self.prec = self._PREC[precision_mode]
        self.max_iter = int(max_iter)
        self.max_wait_s = float(max_wait_s)
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(store.device)
            hbm_budget_bytes = int(0.8 * free)
        self.budget = int(hbm_budget_bytes)
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._inflight_bytes = 0
        self._inflight = 0
        self._stop = False
        self._next_id = 0
        self._threads = []
        self.max_inflight_seen = 0

    # ---- lifecycle ---------------------------------------------------------
    def start(self):
        for i in range(self.slots):
            t = threading.Thread(target=self._worker, name=f"lym-slot-{i}", daemon=True)
            t.start()
            self._threads.append(t)
        return self

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._threads = []

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # ---- admission -----------------------------------------------------------
    def _prec_for(self, nshift):
        """Storage precision of a request with `nshift` shifts."""
        return self.prec if (self.prec != _lib.PREC_F16 or nshift == 1) else _lib.PREC_F32

    def footprint(self, req):
        """Device bytes a request holds while it runs."""
        d = self.store.desc(req.field, req.kappa)
        n = len(req.shifts)
        ws = int(_lib.lib().lym_solver_workspace_bytes(ctypes.byref(d), self._prec_for(n), n))
        if ws == 0:
            raise RuntimeError(_lib.lib().lym_last_error().decode())
        elems = int(np.prod(self.store.spinor_shape(req.field)))
        return ws + (n + 1) * elems * 16

    def submit(self, req):
        fut = Future()
        fp = self.footprint(req)
        if fp > self.budget:
            fut.set_exception(RuntimeError(f"request needs {fp} bytes, above the budget {self.budget}"))
            return fut
        with self._cv:
            if req.rid < 0:
                req.rid = self._next_id
            self._next_id += 1
            if req.arrival is None:
                req.arrival = time.perf_counter()
            self._pending.append((req, fp, fut))
            self._cv.notify_all()
        return fut

    def _take(self):
        """Under the lock: first request (arrival order) that fits, or None."""
        now = time.perf_counter()
        for i, (req, fp, fut) in enumerate(self._pending):
            fits = self._inflight_bytes + fp <= self.budget
            if fits:
                del self._pending[i]
                return req, fp, fut
            if now - req.arrival > self.max_wait_s:
                return None  # the oldest waiter blocks later admissions until it fits
        return None

    # ---- workers ---------------------------------------------------------------
    def _worker(self):
        stream = torch.cuda.Stream(device=self.store.device)
        while True:
            with self._cv:
                item = None
                while not self._stop:
                    item = self._take()
                    if item is not None:
                        break
                    self._cv.wait(timeout=0.05)
                if item is None:
                    return
                req, fp, fut = item
                self._inflight_bytes += fp
                self._inflight += 1
                self.max_inflight_seen = max(self.max_inflight_seen, self._inflight)
            try:
                with torch.cuda.stream(stream):
                    res = self._solve(req, fp, stream)
                fut.set_result(res)
            except Exception as e:  # surfaces on the future
                fut.set_exception(e)
            finally:
                with self._cv:
                    self._inflight_bytes -= fp
                    self._inflight -= 1
                    self._cv.notify_all()

    def _solve(self, req, fp, stream):
        t0 = time.perf_counter()
        lib = _lib.lib()
        d = self.store.desc(req.field, req.kappa)
        shifts = [float(s) for s in req.shifts]
        n = len(shifts)
        dev = self.store.device
        src = to_device(req.source, device=dev)
        if tuple(src.shape) != self.store.spinor_shape(req.field):
            raise RuntimeError(f"source shape {tuple(src.shape)} does not match field {req.field}")
        prec = self._prec_for(n)
        nbytes = int(lib.lym_solver_workspace_bytes(ctypes.byref(d), prec, n))
        ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
        xs = [torch.empty_like(src) for _ in shifts]
        sh = (ctypes.c_double * n)(*shifts)
        xp = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        st = _lib.LymSolveStats()
        rc = _lib.check(lib.lym_cgm(ctypes.byref(d), prec, n, sh, _lib.ptr(src), xp, float(req.tolerance),
                                    self.max_iter, _lib.ptr(ws), ws.numel(), ctypes.byref(st),
                                    ctypes.c_void_p(stream.cuda_stream)), "serve")
        stream.synchronize()
        done = time.perf_counter()
        return SolveResult(rid=req.rid, solutions=xs, stats=st.as_dict(n),
                           status={0: "converged", 1: "max_iterations", 2: "breakdown"}.get(rc, str(rc)),
                           arrival=req.arrival, start=t0, done=done, footprint=fp)


def latency_summary(results):
    lat = np.array([r.latency for r in results])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "n": int(len(lat))}
