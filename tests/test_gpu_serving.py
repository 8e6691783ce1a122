"""Serving: requests of different volumes, colours, kappas and shift counts
overlap on the device; each result meets 1e-10 (oracle-recomputed) and is
bitwise the same as the same request solved alone."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from oracle import solvers as osv
from oracle import wilson as ow
from tfleonardym_amd import fields


def build_store():
    from tfleonardym_amd.serving import GaugeStore
    store = GaugeStore()
    specs = {"a8su3": ((8, 8, 8, 8), 3), "b4su2": ((4, 4, 4, 4), 2), "c8x16su2": ((8, 4, 4, 16), 2)}
    links = {}
    for i, (k, (g, nc)) in enumerate(specs.items()):
        U = fields.gauge_field(g, nc, seed=300 + i, kind="near_identity", eps=0.25)
        store.add(k, U, g)
        links[k] = (U, g, nc)
    return store, links


def make_requests(links, n=10, seed=5):
    from tfleonardym_amd.serving import SolveRequest
    rng = np.random.default_rng(seed)
    keys = sorted(links)
    reqs = []
    for i in range(n):
        k = keys[i % len(keys)]
        U, g, nc = links[k]
        ns = rng.integers(1, 6)
        shifts = fields.log_shifts(1e-3, 5.0, int(ns)) if ns > 1 else [0.01]
        src = fields.gaussian_spinor(g, 4, nc, seed=1000 + i)
        reqs.append(SolveRequest(field=k, kappa=float(rng.uniform(0.10, 0.125)), shifts=shifts, source=src, rid=i))
    return reqs


def run(store, reqs, slots, precision_mode="mixed"):
    from tfleonardym_amd.serving import SolveRequest, SolveServer
    with SolveServer(store, slots=slots, precision_mode=precision_mode) as srv:
        futs = [srv.submit(SolveRequest(r.field, r.kappa, r.shifts, r.source, rid=r.rid)) for r in reqs]
        out = {f.result(timeout=120).rid: f.result() for f in futs}
        inflight = srv.max_inflight_seen
    return out, inflight


@pytest.mark.parametrize("mode", ["mixed", "half"])
def test_mix_equals_alone_and_meets_tolerance(mode):
    """mode "half": the one-shift requests run with 16-bit Krylov storage."""
    print(f"case serving mode={mode}", flush=True)
    store, links = build_store()
    reqs = make_requests(links)
    mixed, inflight = run(store, reqs, slots=4, precision_mode=mode)
    alone = {}
    for r in reqs:
        res, _ = run(store, [r], slots=1, precision_mode=mode)
        alone.update(res)
    assert inflight >= 2  # requests did overlap
    for r in reqs:
        m, a = mixed[r.rid], alone[r.rid]
        assert m.status == "converged"
        for xm, xa in zip(m.solutions, a.solutions):
            assert torch.equal(xm, xa), r.rid
        U, g, nc = links[r.field]
        oop = ow.WilsonOracle(U, g, r.kappa, True, "fundamental")
        for s, x in zip(r.shifts, m.solutions):
            assert osv.true_residual(oop.apply_square, r.source, x.cpu().numpy(), s) <= 1e-10


def test_over_budget_request_is_rejected():
    from tfleonardym_amd.serving import SolveRequest, SolveServer
    store, links = build_store()
    r = make_requests(links, n=1)[0]
    srv = SolveServer(store, slots=1, hbm_budget_bytes=1024)
    fut = srv.submit(SolveRequest(r.field, r.kappa, r.shifts, r.source))
    with pytest.raises(RuntimeError):
        fut.result(timeout=5)


def test_same_operator_requests_are_merged():
    """Requests on one field with one kappa and shift list, waiting together,
    run as one lym_cgm_batched call; each meets 1e-10 (oracle-recomputed) and
    agrees with the same request solved alone to 1e-8."""
    from tfleonardym_amd.serving import SolveRequest, SolveServer
    print("case serving merge", flush=True)
    store, links = build_store()
    U, g, nc = links["a8su3"]
    shifts = fields.log_shifts(1e-3, 5.0, 4)
    reqs = [SolveRequest("a8su3", 0.124, shifts, fields.gaussian_spinor(g, 4, nc, seed=2000 + i), rid=i)
            for i in range(5)]
    srv = SolveServer(store, slots=1, precision_mode="mixed", max_batch=8)
    futs = [srv.submit(r) for r in reqs]  # queued before the slot starts: one group
    with srv:
        res = [f.result(timeout=120) for f in futs]
    assert srv.batches[5] == 1, dict(srv.batches)
    oop = ow.WilsonOracle(U, g, 0.124, True, "fundamental")
    for r, m in zip(reqs, res):
        assert m.status == "converged" and m.extra["batch"] == 5
        alone, _ = run(store, [r], slots=1)
        for s, x, xa in zip(shifts, m.solutions, alone[r.rid].solutions):
            assert osv.true_residual(oop.apply_square, r.source, x.cpu().numpy(), s) <= 1e-10
            assert rel(x, xa.cpu().numpy()) < 1e-8


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_device_source_waits_for_the_callers_stream():
    """A device source written by work still queued on the caller's stream is
    read only after it: the server's stream waits on an event recorded at
    submit."""
    from tfleonardym_amd.serving import SolveRequest, SolveServer
    store, links = build_store()
    U, g, nc = links["b4su2"]
    b = torch.from_numpy(fields.gaussian_spinor(g, 4, nc, seed=77)).cuda()
    side = torch.cuda.Stream()
    src = torch.zeros_like(b)
    with SolveServer(store, slots=1) as srv:
        with torch.cuda.stream(side):
            torch.cuda._sleep(20_000_000)  # keeps the side stream busy before the copy
            src.copy_(b)
            fut = srv.submit(SolveRequest("b4su2", 0.12, [0.05], src))
        m = fut.result(timeout=60)
    oop = ow.WilsonOracle(U, g, 0.12, True, "fundamental")
    x = m.solutions[0].cpu().numpy()
    assert osv.true_residual(oop.apply_square, b.cpu().numpy(), x, 0.05) <= 1e-10


@pytest.mark.parametrize("mode", ["mixed", "half"])
def test_full_size_fields(mode):
    """The serving configuration's large fields (16^3x32 and 24^3x48 SU(3)),
    kappa in [0.12, 0.135] and up to 16 shifts, run together: every request
    converges and every shift's true relative residual, recomputed here with
    the complex128 operator, is at most 1e-10."""
    from tfleonardym_amd import dirac, lattice
    from tfleonardym_amd.serving import GaugeStore, SolveRequest, SolveServer
    print(f"case serving full size mode={mode}", flush=True)
    rng = np.random.default_rng(91)
    store = GaugeStore()
    ops = {}
    for seed, (k, g) in enumerate((("f16x32", (16, 16, 16, 32)), ("f24x48", (24, 24, 24, 48)))):
        U = fields.gauge_field(g, 3, seed=610 + seed, kind="near_identity", eps=0.25)
        store.add(k, U, g)
        ops[k] = (U, g)
    reqs = []
    for i, (k, ns) in enumerate((("f16x32", 16), ("f24x48", 1), ("f24x48", 12), ("f16x32", 1))):
        g = ops[k][1]
        shifts = fields.log_shifts(1e-4, 10.0, ns) if ns > 1 else [1e-4]
        reqs.append(SolveRequest(k, float(rng.uniform(0.12, 0.135)), shifts,
                                 torch.from_numpy(fields.gaussian_spinor(g, 4, 3, seed=500 + i)), rid=i))
    with SolveServer(store, slots=3, precision_mode=mode) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=300) for f in futs]
    for r, m in zip(reqs, res):
        assert m.status == "converged", (r.rid, m.stats)
        U, g = ops[r.field]
        op = dirac.DiracWilsonOperator(lattice.Configuration(U, g, 3, 4), r.kappa, True,
                                       representation="fundamental")
        b = r.source.cuda()
        bn = torch.linalg.vector_norm(b).item()
        for s, x in zip(r.shifts, m.solutions):
            res_v = b - op.apply(op.apply(x)) - s * x
            assert torch.linalg.vector_norm(res_v).item() / bn <= 1e-10, (r.rid, s)
        del op
    torch.cuda.empty_cache()
