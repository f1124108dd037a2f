"""Admission order of the serving loop (torchServe.Server.serve), on the CPU.

The device batch is replaced by a stand-in with the same admission contract:
an add that does not fit the budget returns None, each run advances every
admitted request by the given steps, and a request finishes after its own
number of steps. The reference has no server (each Simulation runs alone,
torchSession.py:578-633); these tests pin the host scheduling only.
"""

import pytest

from torch_amd.torchServe import Request, Server


class _Slots:
    def __init__(self, budget):
        self.budget, self.used, self.next = budget, 0, 1
        self.live = {}

    def add(self, need, steps):
        if self.used + need > self.budget:
            return None
        ident, self.next = self.next, self.next + 1
        self.live[ident] = [need, steps, 0]
        self.used += need
        return ident


class _Batch:
    def __init__(self, budget):
        self._batch = _Slots(budget)
        self._meta = {}
        self.order = []

    def run(self, steps):
        for rec in self._batch.live.values():
            rec[2] = min(rec[1], rec[2] + steps)

    def info(self, ident):
        need, steps, done = self._batch.live[ident]
        return dict(active=done < steps, steps=done, status=0, retries=0)

    def remove(self, ident):
        need = self._batch.live.pop(ident)[0]
        self._batch.used -= need


class _Server(Server):
    def __init__(self, budget, policy, max_skips=16):
        self.policy, self.max_skips = policy, max_skips
        self.batch = _Batch(budget)
        self.round_steps, self.max_steps, self.keep_results = 1, 1000, False
        self.rounds, self.device_ms = 0, 0.0

    def payload(self, need, opts):
        return need, opts["steps"]

    def _admit(self, req, payload):
        ident = self.batch._batch.add(*payload)
        if ident is not None:
            self.batch.order.append(req.ident)
        return ident


def _requests(spec):
    # (cells, steps) per request, all arriving at t = 0
    return [Request(i, 0.0, (lambda c=c, s=s: (c, {"steps": s})), cells=c) for i, (c, s) in enumerate(spec)]


SPEC = [(6, 5), (8, 2), (3, 2), (2, 1), (5, 1)]


def test_fifo_stops_at_the_first_request_that_does_not_fit():
    server = _Server(10, "fifo")
    done = server.serve(_requests(SPEC), prebuild=True)
    assert server.batch.order == [0, 1, 2, 3, 4]
    assert all(r.completed is not None and r.steps == s for r, (_, s) in zip(done, SPEC))


def test_backfill_admits_later_requests_that_fit():
    server = _Server(10, "backfill")
    server.serve(_requests(SPEC), prebuild=True)
    # 8 does not fit beside 6; 3 does, then 2 waits, and 8 gets in last.
    assert server.batch.order[:2] == [0, 2]
    assert sorted(server.batch.order) == [0, 1, 2, 3, 4]


def test_largest_first_orders_waiting_requests_by_cells():
    server = _Server(100, "largest")
    server.serve(_requests(SPEC), prebuild=True)
    assert server.batch.order == [1, 0, 4, 2, 3]


def test_a_request_passed_over_too_often_reserves_the_budget():
    # A long request holds 6 of 10; the 8 at the head keeps failing while
    # short 3s could backfill forever. After max_skips passes nothing more
    # is admitted behind it, so the long request's end lets it in.
    spec = [(6, 12)] + [(8, 1)] + [(3, 2)] * 8
    server = _Server(10, "backfill", max_skips=2)
    reqs = server.serve(_requests(spec), prebuild=True)
    head = reqs[1]
    backfilled_before_head = [i for i in server.batch.order[: server.batch.order.index(1)] if i >= 2]
    assert len(backfilled_before_head) <= 3
    assert head.completed is not None and head.skips > 2


def test_unknown_policy_is_rejected():
    with pytest.raises(ValueError):
        Server.__init__(Server.__new__(Server), policy="random")
