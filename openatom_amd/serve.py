"""Continuous-batching request loop for orthonormalisation steps.

Model.  An MD trajectory (a session: one state block of n states and
npw = 64 n plane waves) arrives, stays resident on the device for `steps`
MD steps and leaves.  Each MD step issues one request: an ortho step (S,
S->T, T*psi) of the session's current coefficients; when it completes, the
client perturbs the orthonormalised block by eps (the simulated dynamics:
psi <- T psi + (-1)^k eps P with an envelope-shaped block P drawn once per
session, one fused pass) and issues the next step.  Sessions arrive as a Poisson process; an offered
load of R requests/s is R / steps sessions/s.

Scheduling.
  * Admission by HBM footprint: a session is admitted (its coefficients made
    resident) when the sum of admitted footprints plus its own fits the
    budget; waiting sessions are admitted in arrival order.
  * Continuous batching: every iteration takes the pending request of every
    admitted session and runs them as ONE grouped batch step
    (openatom_amd.ortho.Batch); sessions admitted while a batch runs join at
    the next iteration, finished sessions leave and free their memory.  When
    the membership is unchanged the batch replays its captured hipGraph.
  * Each request's result equals the single-instance step on the same
    coefficients, bit for bit, whatever the mix it ran in.

Reported: step latency (request issued with its coefficients resident ->
result complete) p50 / p99, admission wait (session arrival -> resident)
p50 / p99, completed requests / s.

The policy (Scheduler) is host logic independent of the executor, so it runs
against a simulated executor on the CPU as well as against the GPU.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

SIZES = (128, 256, 512, 1024)
WEIGHTS = (0.4, 0.3, 0.2, 0.1)


@dataclass
class SessionSpec:
    sid: int
    arrival: float   # seconds from the start of the run
    n: int
    npw: int
    steps: int
    seed: int


@dataclass
class Session:
    spec: SessionSpec
    footprint: int
    admitted_at: float = 0.0
    done: int = 0
    issued_at: float = 0.0
    state: object = None           # executor-owned device state
    latencies: list = field(default_factory=list)


def poisson_sessions(rate_req_s: float, duration_s: float, steps: int = 100, seed: int = 0,
                     sizes=SIZES, weights=WEIGHTS, npw_per_state: int = 64) -> list[SessionSpec]:
    """Session arrivals over [0, duration_s) for an offered request rate."""
    rng = np.random.default_rng(seed)
    lam = rate_req_s / steps
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / lam)
        if t >= duration_s:
            break
        n = int(rng.choice(sizes, p=weights))
        out.append(SessionSpec(len(out), t, n, npw_per_state * n, steps,
                               seed * 1_000_003 + len(out)))
    return out


class Scheduler:
    """Admission by footprint (FIFO) and the batch of pending requests."""

    def __init__(self, budget_bytes: int, footprint_fn):
        self.budget = int(budget_bytes)
        self.footprint_fn = footprint_fn
        self.waiting: deque[Session] = deque()
        self.active: dict[int, Session] = {}
        self.used = 0
        self.peak_used = 0
        self.rejected: list[SessionSpec] = []

    def arrive(self, spec: SessionSpec) -> None:
        fp = int(self.footprint_fn(spec.n, spec.npw))
        if fp > self.budget:
            self.rejected.append(spec)   # can never fit, even alone
            return
        self.waiting.append(Session(spec, fp))

    def admit(self, now: float) -> list[Session]:
        out = []
        while self.waiting and self.used + self.waiting[0].footprint <= self.budget:
            s = self.waiting.popleft()
            s.admitted_at = now
            self.used += s.footprint
            self.peak_used = max(self.peak_used, self.used)
            self.active[s.spec.sid] = s
            out.append(s)
        return out

    def batch(self) -> list[Session]:
        return sorted(self.active.values(), key=lambda s: s.spec.sid)

    def retire(self, s: Session) -> None:
        del self.active[s.spec.sid]
        self.used -= s.footprint


class WallClock:
    def __init__(self):
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def sleep_until(self, t: float) -> None:
        d = t - self.now()
        if d > 0:
            time.sleep(d)


def serve(specs: list[SessionSpec], executor, budget_bytes: int, clock=None,
          max_wall_s: float | None = None) -> dict:
    """Run the request loop until every session finished (or max_wall_s)."""
    clock = clock or WallClock()
    sched = Scheduler(budget_bytes, executor.footprint)
    pending = deque(sorted(specs, key=lambda s: s.arrival))
    waits, lat = [], []
    lat_by_n: dict[int, list] = {}
    finished, iterations, batch_sizes = [], 0, []
    t_step = t_advance = 0.0  # loop time in batch steps / client MD steps
    while pending or sched.waiting or sched.active:
        now = clock.now()
        if max_wall_s is not None and now > max_wall_s:
            break
        while pending and pending[0].arrival <= now:
            sched.arrive(pending.popleft())
        for s in sched.admit(now):
            executor.open(s)           # coefficients start moving to the device
            s.issued_at = None
        t_ready = clock.now()
        for s in sched.active.values():
            if s.issued_at is None and executor.ready(s):
                s.issued_at = t_ready  # first request: coefficients resident
                waits.append(s.issued_at - s.spec.arrival)
        batch = [s for s in sched.batch() if s.issued_at is not None]
        if not batch:
            if sched.active:
                clock.sleep_until(clock.now() + 1e-4)  # admitted sessions still loading
            elif pending:
                clock.sleep_until(pending[0].arrival)
            elif sched.waiting:
                break  # nothing active and nothing admissible: cannot progress
            continue
        t0 = clock.now()
        executor.step(batch)
        t_done = clock.now()
        t_step += t_done - t0
        iterations += 1
        batch_sizes.append(len(batch))
        for s in batch:
            s.latencies.append(t_done - s.issued_at)
            lat.append(t_done - s.issued_at)
            lat_by_n.setdefault(s.spec.n, []).append(t_done - s.issued_at)
            s.done += 1
            if s.done >= s.spec.steps:
                executor.close(s)
                sched.retire(s)
                finished.append(s)
            else:
                executor.advance(s)
        t_issue = clock.now()
        t_advance += t_issue - t_done
        for s in batch:
            if s.done < s.spec.steps:
                s.issued_at = t_issue
    elapsed = clock.now()

    def pct(xs, q):
        return float(np.percentile(np.asarray(xs), q)) if xs else None

    return {
        "sessions": len(specs),
        "sessions_finished": len(finished),
        "sessions_rejected": len(sched.rejected),
        "requests_completed": len(lat),
        "elapsed_s": elapsed,
        "throughput_req_s": len(lat) / elapsed if elapsed > 0 else 0.0,
        "latency_p50_ms": None if not lat else 1e3 * pct(lat, 50),
        "latency_p99_ms": None if not lat else 1e3 * pct(lat, 99),
        # per session size: the mix puts n=1024 steps (512x the work of
        # n=128) in the upper percentiles of the pooled latencies
        "latency_by_n_ms": {str(k): {"requests": len(v), "p50": 1e3 * pct(v, 50),
                                     "p99": 1e3 * pct(v, 99)}
                            for k, v in sorted(lat_by_n.items())},
        "admission_wait_p50_ms": None if not waits else 1e3 * pct(waits, 50),
        "admission_wait_p99_ms": None if not waits else 1e3 * pct(waits, 99),
        "iterations": iterations,
        "mean_batch": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
        "batch_step_s": t_step,
        "client_md_step_s": t_advance,
        "peak_footprint_bytes": sched.peak_used,
        "budget_bytes": sched.budget,
    }


class GpuExecutor:
    """Device side of the loop: resident coefficients per session and one
    grouped Batch over all admitted sessions."""

    def __init__(self, handle=None, prec="int8x6a", eps: float = 1e-3, nspin: int = 2,
                 tol: float = 1e-12, maxit: int = 12, use_graph: bool = True,
                 record: bool = False, reserve_bytes: int = 0):
        import torch

        from .ortho import Batch, default_handle, instance_footprint
        self.torch = torch
        self.h = handle or default_handle()
        self.batch = Batch(self.h)
        if reserve_bytes:
            self.batch.reserve(reserve_bytes)
        self.prec, self.eps, self.nspin, self.tol, self.maxit = prec, eps, nspin, tol, maxit
        self.use_graph = use_graph
        self._fp = instance_footprint
        self.members: tuple = ()
        self.replans, self.t_plan, self.t_exec = 0, 0.0, 0.0
        self.stable, self.graph_after = 0, 3
        self.side = torch.cuda.Stream()  # default priority
        self.record = record
        self.records: dict[int, list] = {}

    def tune(self, sessions: int = 24, seed: int = 12345) -> float:
        """Tune the batch's tile table once, before serving, on a synthetic
        mix drawn from the serving size distribution; every later plan takes
        its tile choices from that table.  Returns the seconds spent."""
        from .workload import make_states
        torch = self.torch
        rng = np.random.default_rng(seed)
        t0 = time.perf_counter()
        psis = []
        for i in range(sessions):
            n = int(rng.choice(SIZES, p=WEIGHTS))
            psis.append(make_states(n, 64 * n, eps=1e-3, seed=seed + i, device="cuda",
                                    rng="torch"))
        self.batch.plan(psis, nkx0=1, nspin=self.nspin, tol=self.tol, maxit=self.maxit,
                        prec=self.prec)
        self.batch.tune()
        torch.cuda.synchronize()
        # drop the mix: the next plan (the first real step) replaces the tables
        self.batch._keep = []
        self.batch.psis = self.batch.outs = self.batch.S = self.batch.T = None
        self.members = ()
        del psis
        torch.cuda.empty_cache()
        return time.perf_counter() - t0

    def footprint(self, n: int, npw: int) -> int:
        # psi, psi_out and the MD perturbation block, S and T, plus the batch
        # workspace share
        return 3 * 16 * n * npw + 2 * 8 * n * n + self._fp(n, npw, self.prec)

    def open(self, s: Session) -> None:
        """Make the session's coefficients resident on a side stream (nothing
        in the generator synchronises the host), so the running batches are
        not held up; ready() tells when they are there."""
        from .workload import make_states, md_perturbation
        torch = self.torch
        with torch.cuda.stream(self.side):
            psi = make_states(s.spec.n, s.spec.npw, eps=1e-3, seed=s.spec.seed, device="cuda",
                              rng="torch")
            n = s.spec.n
            s.state = {"psi": psi, "out": torch.empty_like(psi),
                       "P": md_perturbation(n, s.spec.npw, seed=s.spec.seed * 1000 + 1),
                       "S": torch.empty((n, n), dtype=torch.float64, device="cuda"),
                       "T": torch.empty((n, n), dtype=torch.float64, device="cuda"),
                       "ready": torch.cuda.Event()}
            s.state["ready"].record(self.side)

    def ready(self, s: Session) -> bool:
        ev = s.state.get("ready")
        if ev is None:
            return True
        if not ev.query():
            return False
        cur = self.torch.cuda.current_stream()
        cur.wait_event(ev)
        for t in ("psi", "out", "P", "S", "T"):
            s.state[t].record_stream(cur)  # used from now on on the batch stream
        s.state["ready"] = None
        return True

    def step(self, batch: list[Session]) -> None:
        members = tuple(s.spec.sid for s in batch)
        t0 = time.perf_counter()
        self.stable = self.stable + 1 if members == self.members else 0
        if members != self.members:
            self.batch.plan([s.state["psi"] for s in batch], outs=[s.state["out"] for s in batch],
                            nkx0=1, nspin=self.nspin, tol=self.tol, maxit=self.maxit,
                            prec=self.prec, S=[s.state["S"] for s in batch],
                            T=[s.state["T"] for s in batch])
            self.members = members
            self.replans += 1
            t1 = time.perf_counter()
            self.batch.run()
        elif self.use_graph and self.stable >= self.graph_after:
            # a membership that held for a few steps is worth a captured graph
            t1 = time.perf_counter()
            self.batch.launch()
        else:
            t1 = time.perf_counter()
            self.batch.run()
        # wait for this batch only: sessions loading on the side stream go on
        self.torch.cuda.current_stream().synchronize()
        t2 = time.perf_counter()
        self.t_plan += t1 - t0
        self.t_exec += t2 - t1
        if self.record:
            for s in batch:
                self.records.setdefault(s.spec.sid, []).append(s.state["out"].clone())

    def advance(self, s: Session) -> None:
        """The client's MD step: psi <- T psi + (-1)^k eps P, one fused pass."""
        from .workload import md_step_
        st = s.state
        md_step_(st["psi"], st["out"], st["P"], self.eps, s.done)

    def close(self, s: Session) -> None:
        s.state = None
        self.members = ()  # the next plan drops the finished session

    def release(self) -> None:
        self.batch.close()
