"""Admission logic of the request loop, driven by a CPU stand-in executor."""
import threading
import time

import numpy as np
import pytest

from sucemfem.Serving import (Request, Server, box_sizes, footprint_bytes, make_trace,
                              summarize)


def test_box_sizes_match_closed_form_counts():
    assert box_sizes(4, 1)[1:4] == (604, 864, 604)
    _, ne, nf, nd, _ = box_sizes(40, 2)
    assert (ne, nf, nd) == (462520, 777600, 2480240)


def test_trace_is_seeded_and_poisson_like():
    a = make_trace(200, seed=3)
    b = make_trace(200, seed=3)
    assert [(r.kind, r.n, r.order) for r in a] == [(r.kind, r.n, r.order) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert 0.003 < gaps.mean() < 0.008                    # rate 200/s
    kinds = [r.kind for r in a]
    assert 0.55 < kinds.count('eigen') / len(kinds) < 0.85
    for r in a:
        assert 6 <= r.n <= 48 and r.order in (1, 2)
        if r.kind == 'driven':
            assert r.freq > 0
        else:
            assert 0 < r.sigma


class FakeExecutor(object):
    """Sleeps in proportion to the request size and records concurrency."""

    def __init__(self):
        self.lock = threading.Lock()
        self.active = 0
        self.peak = 0
        self.order = []

    def __call__(self, req):
        with self.lock:
            self.active += 1
            self.peak = max(self.peak, self.active)
            self.order.append(req.id)
        time.sleep(0.002 + 1e-9 * req.n ** 3)
        with self.lock:
            self.active -= 1
        return req.id * 2, {"ok": True}


def test_slots_bound_concurrency_and_admission_is_fifo():
    ex = FakeExecutor()
    trace = [Request(i, 'driven', (1, 1, 1), 10, 1, arrival=0.0005 * i) for i in range(40)]
    res, span = Server(ex, slots=3).run(trace)
    assert ex.peak <= 3
    assert ex.order == sorted(ex.order)                   # admitted in arrival order
    assert all(r.status == 'done' and r.value == 2 * r.id for r in res)
    assert all(r.start >= r.arrival - 1e-9 for r in res)
    s = summarize(res, span)
    assert s["done"] == 40 and s["latency_ms_p99"] >= s["latency_ms_p50"] > 0


def test_memory_budget_limits_and_rejects():
    ex = FakeExecutor()
    big = Request(0, 'eigen', (1, 1, 1), 48, 2, arrival=0.0)
    small = [Request(i, 'driven', (1, 1, 1), 8, 1, arrival=0.0) for i in range(1, 6)]
    budget = 3 * footprint_bytes(small[0])
    res, _ = Server(ex, slots=8, hbm_budget=budget).run([big] + small)
    assert res[0].status == 'rejected' and "budget" in res[0].error
    assert all(r.status == 'done' for r in res[1:])
    assert ex.peak <= 3


def test_cancellation_and_failures_are_reported():
    def ex(req):
        if req.id == 2:
            raise RuntimeError("solver failed")
        time.sleep(0.01)
        return None, {}
    srv = Server(ex, slots=1)
    srv.cancel(3)
    trace = [Request(i, 'driven', (1, 1, 1), 6, 1, arrival=0.0) for i in range(5)]
    res, _ = srv.run(trace)
    st = {r.id: r.status for r in res}
    assert st[2] == 'failed' and st[3] == 'cancelled' and st[0] == 'done' and st[4] == 'done'
