"""Continuous-batching request loop over device-resident inputs.

Requests arrive on their own clock (a seeded Poisson process). A dispatcher
admits waiting requests in arrival order whenever a slot is free and the
in-flight byte budget allows; each slot owns a Context (its own HIP stream and
workspaces), so admitted requests run concurrently on the device and a request
that arrives while others are running starts as soon as a slot frees up rather
than waiting for a batch boundary. Routing (learned router or the default
rule) is decided at arrival, batched over the requests that arrived together.

The reference has no serving loop: AutoMatcher::scan (core/src/router.rs:205-225)
handles one call at a time; scan_batch (core/src/batch.rs:170-195) handles a
fixed list. Results of a request do not depend on what else is in flight.
"""
from __future__ import annotations

import queue
import threading
import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .api import Context, GpuMatcher, PatternSet, Tuning


@dataclass
class Job:
    rid: int
    patterns: int          # index into the loop's pattern sets
    ptr: int               # device address of the input
    nbytes: int
    arrival: float         # seconds after the loop starts
    mode: str = "overlapping"
    density: float = 0.0   # planted matches per MiB (a router feature)


@dataclass
class Outcome:
    rid: int
    arrival: float
    admitted: float
    done: float
    count: int
    config: int
    matches: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Slot:
    def __init__(self, index: int, device: int, loop: "RequestLoop"):
        self.index = index
        self.ctx = Context(device)
        self.loop = loop
        self.matchers: Dict[int, GpuMatcher] = {}
        self.inbox: "queue.Queue" = queue.Queue()
        self.thread = threading.Thread(target=self._run, name=f"ws-slot-{index}", daemon=True)
        self.thread.start()

    def matcher(self, i: int) -> GpuMatcher:
        m = self.matchers.get(i)
        if m is None:
            m = self.matchers[i] = GpuMatcher(self.loop.patterns[i], context=self.ctx)
        return m

    def _run(self):
        while True:
            item = self.inbox.get()
            if item is None:
                return
            job, cfg, tuning = item
            err = None
            out = None
            try:
                out = self.matcher(job.patterns).scan_device_array(job.ptr, job.nbytes, job.mode, tuning)
            except Exception as e:  # reported by the dispatcher
                err = e
            self.loop._complete(self, job, cfg, out, err)

    def close(self):
        self.inbox.put(None)
        self.thread.join()
        self.ctx.close()


class RequestLoop:
    """Continuous batching: `slots` concurrent requests, admission bounded by
    free slots and by `budget_bytes` of input in flight (a request larger than
    the budget runs when nothing else is in flight)."""

    def __init__(self, patterns: Sequence[PatternSet], route: Callable[[List[Job]], List[tuple]],
                 slots: int = 4, budget_bytes: int = 1 << 30, device: int = 0, keep_matches: bool = False):
        self.patterns = list(patterns)
        self.route = route          # jobs -> [(config index, Tuning)]
        self.budget = budget_bytes
        self.keep = keep_matches
        self.cv = threading.Condition()
        self.free: deque = deque()
        self.inflight = 0
        self.outcomes: List[Outcome] = []
        self.errors: List[BaseException] = []
        self.t0 = 0.0
        self.slots = [_Slot(i, device, self) for i in range(slots)]
        self.free.extend(self.slots)
        self._admit_time: Dict[int, float] = {}

    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _complete(self, slot, job, cfg, out, err):
        done = self._now()
        with self.cv:
            if err is not None:
                self.errors.append(err)
            else:
                self.outcomes.append(Outcome(job.rid, job.arrival, self._admit_time[job.rid], done, len(out), cfg,
                                             out if self.keep else None))
            self.inflight -= job.nbytes
            self.free.append(slot)
            self.cv.notify()

    def warm(self, jobs: Sequence[Job]) -> None:
        """Compile-once work outside the timed loop: every slot's matcher for
        every pattern set the jobs use."""
        for s in self.slots:
            for i in sorted({j.patterns for j in jobs}):
                s.matcher(i)

    def run(self, jobs: Sequence[Job]) -> List[Outcome]:
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.rid))
        self.outcomes, self.errors = [], []
        waiting: deque = deque()
        routes: Dict[int, tuple] = {}
        i, n, active = 0, len(jobs), 0
        self.t0 = time.perf_counter()
        with self.cv:
            while True:
                now = self._now()
                arrived = []
                while i < n and jobs[i].arrival <= now:
                    arrived.append(jobs[i])
                    i += 1
                if arrived:
                    self.cv.release()
                    try:
                        for j, r in zip(arrived, self.route(arrived)):
                            routes[j.rid] = r
                    finally:
                        self.cv.acquire()
                    waiting.extend(arrived)
                while waiting and self.free and (self.inflight == 0 or
                                                 self.inflight + waiting[0].nbytes <= self.budget):
                    job = waiting.popleft()
                    slot = self.free.popleft()
                    self.inflight += job.nbytes
                    self._admit_time[job.rid] = self._now()
                    cfg, tuning = routes.pop(job.rid)
                    slot.inbox.put((job, cfg, tuning))
                if i >= n and not waiting and len(self.free) == len(self.slots):
                    break
                if self.errors:
                    break
                timeout = None if i >= n else max(0.0, jobs[i].arrival - self._now())
                if timeout is None or timeout > 0:
                    self.cv.wait(timeout)
            # drain what is still running after an error
            while len(self.free) < len(self.slots):
                self.cv.wait()
        if self.errors:
            raise self.errors[0]
        return sorted(self.outcomes, key=lambda o: o.rid)

    def close(self):
        for s in self.slots:
            s.close()


def poisson_arrivals(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    r = np.random.default_rng(seed)
    return np.cumsum(r.exponential(1.0 / rate_per_s, size=n))


def summarize(outcomes: Sequence[Outcome], nbytes: Sequence[int]) -> dict:
    lat = np.array([o.latency for o in outcomes]) * 1e3
    first = min(o.arrival for o in outcomes)
    last = max(o.done for o in outcomes)
    span = last - first
    total = float(np.sum(nbytes))
    return {
        "requests": len(outcomes),
        "bytes": int(total),
        "wall_s": round(span, 3),
        "throughput_gbs": round(total / span / 1e9, 2),
        "requests_per_s": round(len(outcomes) / span, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3),
    }
