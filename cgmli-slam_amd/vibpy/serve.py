"""Driver of the library's request server (vib_server_*, the serving loop in C++).

The scheduling (admission by HBM footprint, size buckets, continuous batching, the per-request
append / LM iterations / slide) runs in the library. This module plays the clients: requests of
many sessions arrive over time (Poisson), each is submitted once its arrival time has passed,
with its session's next keyframe, and the loop turns until every request has finished.
"""
import ctypes as C
import time

import numpy as np

from . import abi


class Server:
    def __init__(self, ctx, windows, hbm_budget_bytes=0, buckets=(128, 224, 320), precision=abi.VIB_FP64,
                 refinement=0, slide_policy=abi.VIB_MARGIN_OLD, full_iterations=False, first_slot=5000, workers=1):
        """windows: the sessions' first windows (session i = windows[i]). full_iterations: every
        request runs its whole iteration budget, relinearizing each time. workers: contexts that
        serve disjoint sessions concurrently (worker 0 is ctx)."""
        self.ctx = ctx
        self.lib = ctx.lib
        o = abi.VibServerOptions()
        o.hbm_budget = int(hbm_budget_bytes)
        o.precision = precision
        o.refinement_steps = refinement
        o.slide_policy = slide_policy
        o.full_iterations = 1 if full_iterations else 0
        o.first_slot = first_slot
        o.workers = int(workers)
        o.num_buckets = len(buckets)
        for i, b in enumerate(buckets):
            o.bucket_limit[i] = int(b)
        self.h = C.c_void_p()
        ctx.check(self.lib.vib_server_create(ctx.h, C.byref(o), C.byref(self.h)), "vib_server_create")
        for i, w in enumerate(windows):
            cs = w.cstruct()
            ctx.check(self.lib.vib_server_add_session(self.h, i, C.byref(cs)), "vib_server_add_session")
        self._out = (abi.VibServed * 512)()

    def close(self):
        if self.h:
            self.lib.vib_server_destroy(self.h)
            self.h = C.c_void_p()

    def preload(self):
        """Makes the sessions' windows resident (session order) until the HBM budget is full."""
        self.ctx.check(self.lib.vib_server_preload(self.h), "vib_server_preload")

    def submit(self, request, session, iterations, keyframe=None):
        k = None if keyframe is None else C.byref(keyframe.cstruct())
        self.ctx.check(self.lib.vib_server_submit(self.h, int(request), int(session), int(iterations), k),
                       "vib_server_submit")

    def turn(self):
        n = C.c_int32()
        self.ctx.check(self.lib.vib_server_turn(self.h, self._out, len(self._out), C.byref(n)), "vib_server_turn")
        return [dict(id=int(r.request), session=int(r.session), status=int(r.status), iterations_run=int(r.iterations),
                     unmatched=int(r.unmatched), final_cost=float(r.final_cost), queued=float(r.queued_seconds),
                     latency=float(r.latency_seconds)) for r in self._out[:n.value]]

    def stats(self):
        st = abi.VibServerStats()
        self.ctx.check(self.lib.vib_server_get_stats(self.h, C.byref(st)), "vib_server_get_stats")
        return {name: getattr(st, name) for name, _ in abi.VibServerStats._fields_}

    def slot(self, session):
        v = C.c_int32()
        self.ctx.check(self.lib.vib_server_session_slot(self.h, int(session), C.byref(v)), "vib_server_session_slot")
        return int(v.value)

    def session_view(self, session):
        """(context, slot) holding the session's window: a non-owning Context of its worker."""
        h, v = C.c_void_p(), C.c_int32()
        self.ctx.check(self.lib.vib_server_session_context(self.h, int(session), C.byref(h), C.byref(v)),
                       "vib_server_session_context")
        return abi.Context.borrowed(h, self.lib), int(v.value)

    def run(self, requests, keyframes=None, on_done=None, clock=time.perf_counter):
        """Serves `requests` (dicts with id, session, iterations, arrival in seconds from the start;
        sorted by arrival) in real time: each is submitted once its arrival has passed, with the
        next keyframe of keyframes[session] when keyframes is given. on_done(result) is called
        right after the turn that finished the request (its session's window is still resident
        and unchanged then). Returns the finished requests."""
        pending = list(requests)
        arrival = {r["id"]: r["arrival"] for r in pending}
        nxt = {}
        t0 = clock()
        finished, i = [], 0
        while len(finished) < len(pending):
            now = clock() - t0
            while i < len(pending) and pending[i]["arrival"] <= now:
                r = pending[i]
                kf = None
                if keyframes is not None:
                    j = nxt.get(r["session"], 0)
                    kf = keyframes[r["session"]][j]
                    nxt[r["session"]] = j + 1
                self.submit(r["id"], r["session"], r["iterations"], kf)
                i += 1
            out = self.turn()
            for res in out:
                res["arrival"] = arrival[res["id"]]
                if on_done:
                    on_done(res)
            finished.extend(out)
            if not out and i < len(pending):
                st = self.stats()
                if st["running"] == 0 and st["queued"] == 0:
                    wait = pending[i]["arrival"] - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
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
                failed=int(sum(1 for r in finished if r["status"] != 0)),
                statuses={abi.STATUS_NAMES.get(k, str(k)): int(v) for k, v in
                          zip(*np.unique([r["status"] for r in finished], return_counts=True))})
