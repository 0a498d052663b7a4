"""Serving loop: continuous batching of window requests from many sessions on one device.

Each session keeps its window (and its marginalization prior) resident in a slot of the
library. A request asks for 1-3 LM iterations on its session's window. Requests arrive over
time; a session serves its requests in order. Every loop turn:
  1. moves the requests that have arrived into their sessions' queues;
  2. admits the first queued request of every idle session: if the session's window is not
     resident, it is uploaded when its HBM footprint fits the budget, after evicting resident
     windows of idle sessions (least recently used first); otherwise the request waits;
  3. runs one LM iteration for the running requests of one size bucket (retained dimension)
     in one batched launch sequence (vib_request_iterate), the buckets taking turns;
  4. retires the requests that finished, with their latency (arrival to completion).
Requests that arrive while others are on the device therefore join at the next iteration of
their bucket. The LM work and every per-window decision run in the library.
"""
import collections
import time

import numpy as np

from . import abi


class Session:
    def __init__(self, sid, window):
        self.sid = sid
        self.window = window
        self.slot = None
        self.footprint = abi.window_footprint(window)
        self.nr = window.layout()["nr"]
        self.queue = collections.deque()
        self.active = None
        self.last_used = 0.0
        self.dirty = False  # resident state newer than self.window


class Server:
    def __init__(self, ctx, sessions, hbm_budget_bytes, buckets=(128, 224, 320), precision=abi.VIB_FP64,
                 first_slot=5000, full_iterations=False):
        """full_iterations: every request runs its whole iteration budget, relinearizing each
        time, instead of stopping at convergence (the load of fresh frames on windows that do
        not slide)."""
        self.ctx = ctx
        self.flags = abi.VIB_SOLVE_CONTINUOUS if full_iterations else 0
        self.sessions = sessions
        self.budget = hbm_budget_bytes
        self.buckets = buckets
        self.precision = precision
        self.resident = 0
        self.free_slots = collections.deque(range(first_slot, first_slot + len(sessions)))
        self.turn = 0
        self.uploads = 0
        self.evictions = 0
        self.on_done = None  # callback(request, session) after a request finishes

    def bucket_of(self, s):
        for i, b in enumerate(self.buckets):
            if s.nr <= b:
                return i
        return len(self.buckets) - 1

    def _evict(self, s):
        if s.dirty:
            self.ctx.read_state(s.slot, s.window)
            s.dirty = False
        self.ctx.release(s.slot)
        self.free_slots.append(s.slot)
        s.slot = None
        self.resident -= s.footprint
        self.evictions += 1

    def _make_resident(self, s, now):
        if s.slot is not None:
            return True
        if s.footprint > self.budget:
            raise RuntimeError(f"session {s.sid} needs {s.footprint} bytes, more than the HBM budget")
        idle = sorted((o for o in self.sessions if o.slot is not None and o.active is None and not o.queue),
                      key=lambda o: o.last_used)
        while self.resident + s.footprint > self.budget and idle:
            self._evict(idle.pop(0))
        if self.resident + s.footprint > self.budget:
            return False
        s.slot = self.free_slots.popleft()
        self.ctx.upload(s.slot, s.window)
        self.resident += s.footprint
        self.uploads += 1
        s.last_used = now
        return True

    def preload(self):
        """Makes as many sessions resident as the budget allows (before a timed run)."""
        for s in self.sessions:
            if not self._make_resident(s, 0.0):
                break

    def run(self, requests, clock=time.perf_counter):
        """Serves `requests` (dicts with session, iterations, arrival in seconds from the start;
        sorted by arrival) in real time. Returns the finished requests with `done` (seconds from
        the start), `latency`, `status`, `final_cost`, `iterations_run`."""
        pending = collections.deque(requests)
        t0 = clock()
        finished = []
        active = {}  # slot -> session
        while pending or active or any(s.queue for s in self.sessions):
            now = clock() - t0
            while pending and pending[0]["arrival"] <= now:
                r = pending.popleft()
                self.sessions[r["session"]].queue.append(r)
            for s in self.sessions:
                if s.active is None and s.queue and self._make_resident(s, now):
                    r = s.queue.popleft()
                    opts = abi.options(max_iterations=r["iterations"], precision=self.precision, flags=self.flags)
                    self.ctx.request_begin(s.slot, opts)
                    r["start"] = now
                    s.active = r
                    active[s.slot] = s
            if not active:
                if pending:
                    wait = pending[0]["arrival"] - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                continue
            # the next bucket (in turn) that has running requests
            by_bucket = collections.defaultdict(list)
            for slot, s in active.items():
                by_bucket[self.bucket_of(s)].append(slot)
            keys = sorted(by_bucket)
            b = keys[self.turn % len(keys)]
            self.turn += 1
            slots = by_bucket[b]
            done, reps = self.ctx.request_iterate(slots, self.precision)
            t = clock() - t0
            for slot, d, rep in zip(slots, done, reps):
                if not d:
                    continue
                s = active.pop(slot)
                r = s.active
                s.active = None
                s.dirty = True
                s.last_used = t
                r.update(done=t, latency=t - r["arrival"], status=rep.status, final_cost=rep.final_cost,
                         iterations_run=rep.iterations)
                finished.append(r)
                if self.on_done:
                    self.on_done(r, s)
        return finished


def poisson_requests(n_sessions, rate, duration, rng, iterations=(1, 3)):
    """Requests of the serving workload: Poisson arrivals at `rate` per second over `duration`
    seconds, each for a uniformly chosen session and 1-3 LM iterations."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(dict(id=len(out), session=int(rng.integers(n_sessions)), arrival=t,
                        iterations=int(rng.integers(iterations[0], iterations[1] + 1))))


def summary(finished, wall):
    lat = np.array([r["latency"] for r in finished]) * 1e3
    steps = int(sum(r["iterations_run"] for r in finished))
    return dict(requests=len(finished), window_steps=steps, wall_s=wall, steps_per_s=steps / wall,
                requests_per_s=len(finished) / wall, p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), max_ms=float(lat.max()),
                failed=int(sum(1 for r in finished if r["status"] != 0)))
