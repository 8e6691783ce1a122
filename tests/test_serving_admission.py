"""Admission policy of the solve server (host logic only, no GPU): first fit in
arrival order under a byte budget, a long waiter blocks later admissions, and
waiting requests of one operator and shift list merge into one group."""
import collections
import threading
import time

from tfleonardym_amd.serving import SolveRequest, SolveServer


def make_server(budget, max_wait_s=10.0, max_batch=1, per_rhs=None):
    s = SolveServer.__new__(SolveServer)
    s.budget = budget
    s.max_wait_s = max_wait_s
    s.max_batch = max_batch
    s._pending = collections.deque()
    s._cv = threading.Condition()
    s._inflight_bytes = 0
    if per_rhs is not None:  # group footprint: nrhs x per_rhs (stands in for the liblym query)
        s.footprint = lambda r, nrhs=1: per_rhs * nrhs
    return s


def req(rid, arrival, field="f", kappa=0.1, shifts=(0.1,)):
    return SolveRequest(field=field, kappa=kappa, shifts=list(shifts), source=None, rid=rid, arrival=arrival)


def take1(s):
    """The single request of a group taken with max_batch = 1."""
    items, fp = s._take()
    assert len(items) == 1
    return items[0][0], fp


def test_first_fit_in_arrival_order():
    s = make_server(100)
    now = time.perf_counter()
    for rid, fp in ((0, 80), (1, 50), (2, 15)):
        s._pending.append((req(rid, now), fp, None, None))
    r, fp = take1(s)
    assert r.rid == 0 and fp == 80
    s._inflight_bytes += fp
    r, fp = take1(s)  # 50 does not fit next to 80, 15 does
    assert r.rid == 2
    s._inflight_bytes += fp
    assert s._take() is None
    s._inflight_bytes -= 80
    r, _ = take1(s)
    assert r.rid == 1


def test_long_waiter_blocks_later_requests():
    s = make_server(100, max_wait_s=0.0)
    old = time.perf_counter() - 1.0
    s._inflight_bytes = 60
    s._pending.append((req(0, old), 50, None, None))   # does not fit, has waited too long
    s._pending.append((req(1, old + 0.5), 10, None, None))
    assert s._take() is None                            # 10 would fit, but the old waiter goes first
    s._inflight_bytes = 0
    assert take1(s)[0].rid == 0


def test_same_operator_requests_group_up_to_max_batch():
    s = make_server(1000, max_batch=3, per_rhs=10)
    now = time.perf_counter()
    specs = [("f", 0.1), ("g", 0.1), ("f", 0.1), ("f", 0.12), ("f", 0.1), ("f", 0.1)]
    for rid, (field, kappa) in enumerate(specs):
        s._pending.append((req(rid, now, field, kappa), 10, None, None))
    items, fp = s._take()
    assert [it[0].rid for it in items] == [0, 2, 4] and fp == 30  # same field and kappa, at most 3
    items, fp = s._take()
    assert [it[0].rid for it in items] == [1]                    # another field
    items, fp = s._take()
    assert [it[0].rid for it in items] == [3]                    # another kappa
    items, fp = s._take()
    assert [it[0].rid for it in items] == [5]
    assert s._take() is None


def test_group_stops_at_the_budget_and_shift_lists_must_match():
    s = make_server(25, max_batch=8, per_rhs=10)
    now = time.perf_counter()
    s._pending.append((req(0, now), 10, None, None))
    s._pending.append((req(1, now, shifts=(0.1, 0.2)), 10, None, None))  # other shift list
    s._pending.append((req(2, now), 10, None, None))
    s._pending.append((req(3, now), 10, None, None))
    items, fp = s._take()
    assert [it[0].rid for it in items] == [0, 2] and fp == 20  # a third would need 30 > 25
    s._inflight_bytes += fp
    assert s._take() is None
