"""Admission policy of the solve server (host logic only, no GPU): first fit in
arrival order under a byte budget, and a long waiter blocks later admissions."""
import collections
import threading
import time

from tfleonardym_amd.serving import SolveRequest, SolveServer


def make_server(budget, max_wait_s=10.0):
    s = SolveServer.__new__(SolveServer)
    s.budget = budget
    s.max_wait_s = max_wait_s
    s._pending = collections.deque()
    s._cv = threading.Condition()
    s._inflight_bytes = 0
    return s


def req(rid, arrival):
    return SolveRequest(field="f", kappa=0.1, shifts=[0.1], source=None, rid=rid, arrival=arrival)


def test_first_fit_in_arrival_order():
    s = make_server(100)
    now = time.perf_counter()
    for rid, fp in ((0, 80), (1, 50), (2, 15)):
        s._pending.append((req(rid, now), fp, None))
    r, fp, _ = s._take()
    assert r.rid == 0 and fp == 80
    s._inflight_bytes += fp
    r, fp, _ = s._take()  # 50 does not fit next to 80, 15 does
    assert r.rid == 2
    s._inflight_bytes += fp
    assert s._take() is None
    s._inflight_bytes -= 80
    r, _, _ = s._take()
    assert r.rid == 1


def test_long_waiter_blocks_later_requests():
    s = make_server(100, max_wait_s=0.0)
    old = time.perf_counter() - 1.0
    s._inflight_bytes = 60
    s._pending.append((req(0, old), 50, None))   # does not fit, has waited too long
    s._pending.append((req(1, old + 0.5), 10, None))
    assert s._take() is None                      # 10 would fit, but the old waiter goes first
    s._inflight_bytes = 0
    assert s._take()[0].rid == 0
