"""Serving loop policy on the CPU: a simulated executor on a virtual clock
exercises admission by footprint, continuous batching and retirement."""
import pytest

from openatom_amd.serve import Scheduler, SessionSpec, poisson_sessions, serve


class VClock:
    def __init__(self):
        self.t = 0.0

    def now(self):
        return self.t

    def sleep_until(self, t):
        self.t = max(self.t, t)


class SimExecutor:
    """Service time proportional to the batch's work; records what ran."""

    def __init__(self, clock, per_item=1e-3):
        self.clock, self.per_item = clock, per_item
        self.batches, self.opened, self.closed = [], [], []

    def footprint(self, n, npw):
        return n * npw

    def open(self, s):
        self.opened.append(s.spec.sid)
        s.state = True

    def ready(self, s):
        return True

    def step(self, batch):
        self.batches.append([s.spec.sid for s in batch])
        self.clock.t += self.per_item * sum(s.spec.n for s in batch) / 128

    def advance(self, s):
        pass

    def close(self, s):
        self.closed.append(s.spec.sid)
        s.state = None


def specs(arrivals, n=128, steps=3):
    return [SessionSpec(i, t, n, 64 * n, steps, i) for i, t in enumerate(arrivals)]


def test_every_request_served_once():
    clk = VClock()
    ex = SimExecutor(clk)
    sp = specs([0.0, 0.0005, 0.01, 0.5], steps=4)
    r = serve(sp, ex, budget_bytes=10**12, clock=clk)
    assert r["sessions_finished"] == 4 and r["requests_completed"] == 16
    for sid in range(4):
        assert sum(b.count(sid) for b in ex.batches) == 4
    assert sorted(ex.closed) == [0, 1, 2, 3]


def test_admission_respects_footprint_budget():
    clk = VClock()
    ex = SimExecutor(clk)
    sp = specs([0.0] * 5, n=128, steps=3)
    fp = 128 * 64 * 128
    r = serve(sp, ex, budget_bytes=2 * fp, clock=clk)
    assert max(len(b) for b in ex.batches) == 2
    assert r["peak_footprint_bytes"] <= 2 * fp
    assert ex.opened == sorted(ex.opened)          # FIFO admission
    assert r["sessions_finished"] == 5


def test_late_arrivals_join_running_batches():
    clk = VClock()
    ex = SimExecutor(clk, per_item=1e-2)
    sp = specs([0.0, 0.015, 0.025], steps=5)
    serve(sp, ex, budget_bytes=10**12, clock=clk)
    sizes = [len(b) for b in ex.batches]
    assert sizes[0] == 1 and max(sizes) == 3       # joined while others were running
    first1 = next(b for b in ex.batches if 1 in b)
    assert first1 == [0, 1]                        # joined session 0's running loop


def test_oversized_session_is_rejected_not_blocking():
    clk = VClock()
    ex = SimExecutor(clk)
    sp = specs([0.0, 0.0], n=128, steps=2) + [SessionSpec(2, 0.0, 1024, 65536, 2, 2)]
    r = serve(sp, ex, budget_bytes=3 * 128 * 64 * 128, clock=clk)
    assert r["sessions_rejected"] == 1 and r["sessions_finished"] == 2


def test_poisson_sessions_rate_and_mix():
    sp = poisson_sessions(2000.0, 50.0, steps=100, seed=1)
    # 20 sessions/s for 50 s
    assert 850 < len(sp) < 1150
    frac = {n: sum(s.n == n for s in sp) / len(sp) for n in (128, 256, 512, 1024)}
    assert abs(frac[128] - 0.4) < 0.06 and abs(frac[1024] - 0.1) < 0.04
    assert all(s.npw == 64 * s.n for s in sp)
    assert all(a.arrival < b.arrival for a, b in zip(sp, sp[1:]))


def test_scheduler_retire_frees_budget():
    s = Scheduler(100, lambda n, npw: 60)
    s.arrive(SessionSpec(0, 0.0, 1, 1, 1, 0))
    s.arrive(SessionSpec(1, 0.0, 1, 1, 1, 1))
    assert [x.spec.sid for x in s.admit(0.0)] == [0]
    assert s.admit(0.1) == []
    s.retire(s.active[0])
    assert [x.spec.sid for x in s.admit(0.2)] == [1]
