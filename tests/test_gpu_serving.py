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
