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
* Requests with the same field, kappa, shift list and tolerance that wait
  together are merged (up to `max_batch`) into one lym_cgm_batched call: one
  launch sequence per iteration covers all of them, and the blocks of the
  group's sources on one site tile run back to back on one XCD, so their link
  reads share that XCD's L2.
* Each slot is a worker thread with its own HIP stream.  ctypes releases the
  GIL during a solve, so solves of different requests overlap on the device,
  and a slot takes new work as soon as its solve ends.  A source that is
  already a device tensor is read after the work queued before submit on the
  caller's current stream (an event recorded at submit).
* A request solved on its own has its own workspace, and the kernels'
  reductions have a fixed order, so its result is bitwise the same whether it
  runs alone or beside others (tests/test_gpu_serving.py).  A merged request
  meets the same tolerance; its iterates follow the group's precision
  switches, so its bits can differ from a solve on its own.

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
        U32 = _lib.pack_links_f32(U)
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
        return sum(e["U"].numel() * 16 + e["U32"].numel() * e["U32"].element_size() for e in self._entries.values())

    def keys(self):
        return list(self._entries)


class SolveServer:
    """Continuous-admission multishift solve server (see module docstring)."""

    # "half": r, p and Q p in 16-bit storage with per-site norms (one shift:
    # CG with FP64 reliable updates; several: the three-phase CG-M)
    _PREC = {"mixed": _lib.PREC_F32, "double": _lib.PREC_F64, "half": _lib.PREC_F16}

    def __init__(self, store, slots=4, hbm_budget_bytes=None, precision_mode="mixed", max_iter=50000,
                 max_wait_s=2.0, max_batch=8):
        self.store = store
        self.slots = int(slots)
        self.prec = self._PREC[precision_mode]
        self.max_iter = int(max_iter)
        self.max_wait_s = float(max_wait_s)
        self.max_batch = max(1, int(max_batch))
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
        self.batches = collections.Counter()  # group size -> launches

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
        """Storage precision of a request (the store's fields are fundamental,
        for which every precision is built at every shift count)."""
        return self.prec

    @staticmethod
    def _group_key(req):
        """Requests with equal keys share one operator and shift list and are
        solved together in one lym_cgm_batched call."""
        return (req.field, float(req.kappa), tuple(float(s) for s in req.shifts), float(req.tolerance))

    def footprint(self, req, nrhs=1):
        """Device bytes `nrhs` requests like `req` hold while they run together:
        the solver workspace (lym_cgm_batched's for a group), solutions and
        sources."""
        d = self.store.desc(req.field, req.kappa)
        n = len(req.shifts)
        prec = self._prec_for(n)
        lib = _lib.lib()
        if nrhs == 1:
            ws = int(lib.lym_solver_workspace_bytes(ctypes.byref(d), prec, n))
        else:
            ws = int(lib.lym_cgm_batched_workspace_bytes(ctypes.byref(d), prec, n, nrhs))
        if ws == 0:
            raise RuntimeError(lib.lym_last_error().decode())
        elems = int(np.prod(self.store.spinor_shape(req.field)))
        return ws + nrhs * (n + 1) * elems * 16

    def submit(self, req):
        fut = Future()
        fp = self.footprint(req)
        if fp > self.budget:
            fut.set_exception(RuntimeError(f"request needs {fp} bytes, above the budget {self.budget}"))
            return fut
        # a source already on the device may still be written by work queued on
        # the caller's stream: the solve's stream waits for this event
        ready = None
        src = req.source
        if isinstance(src, torch.Tensor) and src.is_cuda:
            ready = torch.cuda.Event()
            ready.record(torch.cuda.current_stream(src.device))
        with self._cv:
            if req.rid < 0:
                req.rid = self._next_id
            self._next_id += 1
            if req.arrival is None:
                req.arrival = time.perf_counter()
            self._pending.append((req, fp, fut, ready))
            self._cv.notify_all()
        return fut

    def _take(self):
        """Under the lock: the first request (arrival order) that fits, with up
        to max_batch - 1 later requests of the same group key as far as the
        group's footprint fits; or None."""
        now = time.perf_counter()
        for i, (req, fp, fut, ready) in enumerate(self._pending):
            if self._inflight_bytes + fp <= self.budget:
                group = [i]
                gfp = fp
                if self.max_batch > 1:
                    key = self._group_key(req)
                    for j in range(i + 1, len(self._pending)):
                        if len(group) == self.max_batch:
                            break
                        if self._group_key(self._pending[j][0]) != key:
                            continue
                        f = self.footprint(req, len(group) + 1)
                        if self._inflight_bytes + f > self.budget:
                            break
                        group.append(j)
                        gfp = f
                items = [self._pending[j] for j in group]
                for j in reversed(group):
                    del self._pending[j]
                return items, gfp
            if now - req.arrival > self.max_wait_s:
                return None  # the oldest waiter blocks later admissions until it fits
        return None

    # ---- workers ---------------------------------------------------------------
    def _worker(self):
        stream = torch.cuda.Stream(device=self.store.device)
        while True:
            with self._cv:
                taken = None
                while not self._stop:
                    taken = self._take()
                    if taken is not None:
                        break
                    self._cv.wait(timeout=0.05)
                if taken is None:
                    return
                items, fp = taken
                self._inflight_bytes += fp
                self._inflight += len(items)
                self.max_inflight_seen = max(self.max_inflight_seen, self._inflight)
                self.batches[len(items)] += 1
            try:
                with torch.cuda.stream(stream):
                    results = self._solve(items, fp, stream)
                for (_, _, fut, _), res in zip(items, results):
                    fut.set_result(res)
            except Exception as e:  # surfaces on the futures
                for _, _, fut, _ in items:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self._cv:
                    self._inflight_bytes -= fp
                    self._inflight -= len(items)
                    self._cv.notify_all()

    def _solve(self, items, fp, stream):
        """One lym_cgm call (one request) or one lym_cgm_batched call (a group
        of requests with one operator and shift list) on the slot's stream."""
        t0 = time.perf_counter()
        lib = _lib.lib()
        req0 = items[0][0]
        d = self.store.desc(req0.field, req0.kappa)
        shifts = [float(s) for s in req0.shifts]
        n = len(shifts)
        nrhs = len(items)
        dev = self.store.device
        srcs = []
        for req, _, _, ready in items:
            if ready is not None:
                stream.wait_event(ready)
            src = to_device(req.source, device=dev)
            if src.is_cuda and isinstance(req.source, torch.Tensor) and req.source.is_cuda:
                src.record_stream(stream)  # the caller's tensor is in use on this stream
            if tuple(src.shape) != self.store.spinor_shape(req.field):
                raise RuntimeError(f"source shape {tuple(src.shape)} does not match field {req.field}")
            srcs.append(src)
        prec = self._prec_for(n)
        xs = [[torch.empty_like(src) for _ in shifts] for src in srcs]
        sh = (ctypes.c_double * n)(*shifts)
        xp = (ctypes.c_void_p * (nrhs * n))(*[x.data_ptr() for row in xs for x in row])
        if nrhs == 1:
            nbytes = int(lib.lym_solver_workspace_bytes(ctypes.byref(d), prec, n))
            ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
            st = (_lib.LymSolveStats * 1)()
            rc = _lib.check(lib.lym_cgm(ctypes.byref(d), prec, n, sh, _lib.ptr(srcs[0]), xp, float(req0.tolerance),
                                        self.max_iter, _lib.ptr(ws), ws.numel(), st,
                                        ctypes.c_void_p(stream.cuda_stream)), "serve")
        else:
            nbytes = int(lib.lym_cgm_batched_workspace_bytes(ctypes.byref(d), prec, n, nrhs))
            ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
            st = (_lib.LymSolveStats * nrhs)()
            bp = (ctypes.c_void_p * nrhs)(*[s.data_ptr() for s in srcs])
            rc = _lib.check(lib.lym_cgm_batched(ctypes.byref(d), prec, nrhs, n, sh, bp, xp, float(req0.tolerance),
                                                self.max_iter, _lib.ptr(ws), ws.numel(), st,
                                                ctypes.c_void_p(stream.cuda_stream)), "serve batched")
        stream.synchronize()
        done = time.perf_counter()
        del rc  # worst status of the group; each request gets its own from stats
        out = []
        for i, (req, _, _, _) in enumerate(items):
            stats = st[i].as_dict(n)
            out.append(SolveResult(rid=req.rid, solutions=xs[i], stats=stats, status=stats["status"],
                                   arrival=req.arrival, start=t0, done=done, footprint=fp,
                                   extra={"batch": nrhs}))
        return out


def latency_summary(results):
    lat = np.array([r.latency for r in results])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "n": int(len(lat))}
