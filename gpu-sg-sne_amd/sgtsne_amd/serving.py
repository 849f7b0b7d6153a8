"""Continuous-batching server of embedding requests on one GPU.

A request is a point set X [n][dim] and an embedding dimension d (2 or 3).
The server keeps one batched engine per d (sgtsne_batch_*), so P, Y and the
grids of every admitted job stay resident on the device.  Its loop:

  1. queue the next `tick` iterations of every engine (asynchronous) that has
     fewer than `depth` ticks in flight (checked without blocking), so the
     device always has work while the host admits and retires;
  2. admissions: with front="async" (the default) each engine has at most
     one front end (kNN, perplexity, P + P^T on the engine's second stream)
     in flight; the loop polls it without blocking, which queues its next
     part once the device has passed the previous one and admits the job
     after the last, then starts the next front end on every engine without
     one: the first queued request for that engine whose HBM footprint fits,
     3-D requests first (their iterations are the longest, so starting them
     early shortens the makespan), then in arrival order.  The host never
     waits for a kNN, so step 1 keeps both engines fed meanwhile.  With
     front="sync" one request per loop pass is admitted by a front end that
     holds the host until its P is built;
  3. retire the jobs whose last iteration the device has passed (checked
     without blocking, so a job still queued holds up neither engine): its Y
     is read, then its slot is freed.  Reads share the front-end stream, so
     an engine with a front end in flight retires on a later pass.  With
     nothing to do (every engine at `depth` queued ticks, no front end
     advanced, nothing admissible, nothing finished) the host sleeps
     `poll_s` before looking again.

A job's result does not depend on what else shares its engine (the engine
guarantees bit-identical trajectories alone and in any mix), so serving
changes latency and throughput, never the embedding.

The reference has no serving loop: its CLI runs one sgtsneCUDA call per
embedding (CU/testing_modules/demo_perplexity.cu:190-198).
"""
import collections
import time

import numpy as np

from . import Batch, Full, tsneparams
from .synth import random_init

_POINT_BYTES = {2: 6 * 2 * 4 + 4, 3: 6 * 4 * 4 + 4}
# the deepest queue of batch iterations the server has been validated with on
# the GPU (iterations per tick, ticks in flight per engine)
MAX_TICK = 10
MAX_DEPTH = 3  # Y, uY, gains, Fattr, Frep, (spare) + perm


def footprint_bytes(n, d, k=90):
    """Device bytes a job holds while admitted: P_sym (<= 2 n k nonzeros,
    8 B each), the per-point state and the grid buffers of its capacity."""
    grid = {2: (512 if n > 20000 else 256), 3: (128 if n > 20000 else 96)}[d]  # the engine's grid cap
    L = 1
    while L < 2 * grid - 1:
        L <<= 1
    G = grid ** d
    H = L // 2 + 1
    grid_b = G * (d + 1) * 8 + G * (d + 2) * 4 + 3 * grid * L ** (d - 1) * 8 + 3 * H ** d * 8
    return 8 * 2 * n * k + _POINT_BYTES[d] * n + grid_b


class Request:
    __slots__ = ("rid", "X", "d", "arrival", "seed", "admitted", "done", "Y", "engine", "job", "t_left",
                 "end_mark")

    def __init__(self, rid, X, d, arrival, seed=0):
        self.rid, self.X, self.d, self.arrival, self.seed = rid, X, d, arrival, seed
        self.admitted = self.done = None
        self.Y = None
        self.engine = self.job = None
        self.t_left = 0
        self.end_mark = None


class Server:
    """hbm_bytes is split between the d=2 and d=3 engines in proportion to
    `share`; each engine admits jobs while their summed footprint fits."""

    def __init__(self, hbm_bytes=48 << 30, max_jobs=256, tick=10, k=90, u=30.0, max_iter=1000,
                 share=(0.5, 0.5), seed=0, depth=3, poll_s=2e-4, engine_factory=None, front="async"):
        if front not in ("async", "sync"):
            raise ValueError(f"front must be 'async' or 'sync', not {front!r}")
        self.k, self.u, self.max_iter, self.tick, self.front = k, u, max_iter, tick, front
        # ticks in flight per engine: an int for both, or {d: depth}
        self.depth = depth if isinstance(depth, dict) else {2: depth, 3: depth}
        if engine_factory is None and (tick > MAX_TICK or max(self.depth.values()) > MAX_DEPTH):
            # a served sweep faulted the GPU at tick 20 / depth 6 and no cause
            # has been shown; deeper queues than the tested setting are refused
            raise ValueError(f"tick <= {MAX_TICK} and depth <= {MAX_DEPTH} ticks in flight are supported "
                             f"(asked for tick {tick}, depth {depth})")
        self.poll_s = poll_s
        self.engines = {}
        self.budget = {}
        self.used = {}
        for d, frac in zip((2, 3), share):
            b = int(hbm_bytes * frac)
            nnz = min(int(0.6 * b) // 8, 2_000_000_000)   # P_sym: 8 B per nonzero
            pts = max(1, nnz // k)                         # >= the points of any mix (>= k nonzeros per row)
            grid_budget = b  # the footprint check below is the binding limit
            params = tsneparams(d=d, maxIter=max_iter, seed=seed)
            if engine_factory is None:
                self.engines[d] = Batch(params, max_jobs, pts, nnz, grid_budget)
            else:  # any object with Batch's add_points/step/mark/passed/wait/job_left/job_status/get_Y/remove/close
                self.engines[d] = engine_factory(params, max_jobs, pts, nnz, grid_budget)
            self.budget[d] = b
            self.used[d] = 0
        self.queue = collections.deque()
        self.active = {2: {}, 3: {}}
        self.inflight = {2: collections.deque(), 3: collections.deque()}  # marks of queued ticks
        self.fronts = {2: None, 3: None}  # (request, footprint, ticket) of each engine's front end in flight
        self.finished = []

    def close(self):
        for b in self.engines.values():
            b.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, req):
        self.queue.append(req)

    def _register(self, r, fp, job, now_fn):
        eng = self.engines[r.d]
        r.engine, r.job, r.admitted = r.d, job, now_fn()
        r.t_left = eng.job_left(job)  # one admission iteration plus max_iter
        r.X = None
        self.used[r.d] += fp
        self.active[r.d][job] = (r, fp)

    def _admit(self, now_fn, limit=None):
        # first fit, 3-D requests first, then in arrival order: a large request
        # at the head does not block smaller ones behind it.  At most `limit`
        # admissions per call: each runs a front end on the host's time, and
        # the loop tops the engines up between them.
        admitted = 0
        keep = collections.deque()
        order = sorted(self.queue, key=lambda q: (-q.d, q.arrival, q.rid))
        self.queue = collections.deque()
        for i, r in enumerate(order):
            if limit is not None and admitted >= limit:
                keep.extend(order[i:])
                break
            fp = footprint_bytes(r.X.shape[0], r.d, self.k)
            if fp > self.budget[r.d]:
                r.done, r.Y = now_fn(), None  # can never fit: rejected
                self.finished.append(r)
                continue
            if self.used[r.d] + fp > self.budget[r.d]:
                keep.append(r)
                continue
            try:
                # the start depends on the request alone, so its result does too
                y0 = random_init(r.X.shape[0], r.d, seed=r.seed)
                job = self.engines[r.d].add_points(r.X, k=self.k, u=self.u, y_in=y0, max_iter=self.max_iter)
            except Full:
                if not self.active[r.d]:
                    r.done, r.Y = now_fn(), None  # does not fit even an empty engine: rejected
                    self.finished.append(r)
                else:
                    keep.append(r)
                continue
            self._register(r, fp, job, now_fn)
            admitted += 1
        self.queue = collections.deque(sorted(keep, key=lambda q: (q.arrival, q.rid)))
        return admitted

    def _start_fronts(self, now_fn):
        """Starts a front end on every engine without one in flight: its
        first queued request that fits (3-D engine first, then arrival
        order), whose footprint is reserved at once.  Returns how many."""
        started = 0
        keep = collections.deque()
        for r in sorted(self.queue, key=lambda q: (-q.d, q.arrival, q.rid)):
            if self.fronts[r.d] is not None:
                keep.append(r)
                continue
            fp = footprint_bytes(r.X.shape[0], r.d, self.k)
            if fp > self.budget[r.d]:
                r.done, r.Y = now_fn(), None  # can never fit: rejected
                self.finished.append(r)
                continue
            if self.used[r.d] + fp > self.budget[r.d]:
                keep.append(r)
                continue
            y0 = random_init(r.X.shape[0], r.d, seed=r.seed)  # the start depends on the request alone
            t = self.engines[r.d].add_points_async(r.X, k=self.k, u=self.u, y_in=y0, max_iter=self.max_iter)
            self.used[r.d] += fp
            self.fronts[r.d] = (r, fp, t)
            started += 1
        self.queue = collections.deque(sorted(keep, key=lambda q: (q.arrival, q.rid)))
        return started

    def _poll_fronts(self, now_fn):
        """Advances every front end in flight without blocking; registers
        the jobs admitted.  Returns whether any front end moved on."""
        moved = False
        for d, f in self.fronts.items():
            if f is None:
                continue
            r, fp, t = f
            try:
                job = self.engines[d].poll_points(t)
            except Full:
                self.fronts[d] = None
                self.used[d] -= fp
                if not self.active[d]:
                    r.done, r.Y = now_fn(), None  # does not fit even an empty engine: rejected
                    self.finished.append(r)
                else:
                    self.queue.append(r)  # retried once jobs retire
                moved = True
                continue
            if job is None:
                continue
            self.fronts[d] = None
            self.used[d] -= fp  # _register counts it again
            self._register(r, fp, job, now_fn)
            moved = True
        return moved

    def _advance(self):
        """Queues a tick on every engine whose jobs have iterations left;
        returns whether any tick was queued."""
        stepped = False
        for d, act in self.active.items():
            live = [r for r, _ in act.values() if r.t_left > 0]
            if not live:
                continue
            eng, marks = self.engines[d], self.inflight[d]
            while marks and eng.passed(marks[0]):
                marks.popleft()
            if len(marks) >= self.depth[d]:
                continue  # enough queued: this engine is topped up on a later pass
            eng.step(self.tick)
            m = eng.mark()
            marks.append(m)
            for r in live:
                r.t_left -= self.tick
                if r.t_left <= 0:
                    r.end_mark = m  # its last iteration is queued before this mark
            stepped = True
        return stepped

    def _retire(self, now_fn):
        """Retires the jobs whose last iteration the device has passed (a
        non-blocking check, so a job still queued holds up neither engine);
        returns how many."""
        # an engine with a front end in flight reads on a later pass: the
        # reads would queue behind the front end's part on the same stream
        due = [(d, job) for d, act in self.active.items() if self.fronts[d] is None
               for job, (r, _) in act.items() if r.t_left <= 0]
        ready = [(d, job) for d, job in due if self.engines[d].passed(self.active[d][job][0].end_mark)]
        for d, job in ready:
            r, fp = self.active[d].pop(job)
            b = self.engines[d]
            st, it = b.job_status(job)
            r.Y = b.get_Y(job) if st >= 0 else None
            r.done = now_fn()
            b.remove(job)
            self.used[d] -= fp
            self.finished.append(r)
        return len(ready)

    def busy(self):
        return (bool(self.queue) or any(self.active[d] for d in self.active)
                or any(f is not None for f in self.fronts.values()))

    def run(self, arrivals, make_X, clock=None):
        """Serves requests arriving at the given times.  arrivals: list of
        (time_s, n, d, seed); make_X(n, seed) builds the points.  Returns the
        finished requests (arrival, admitted, done in seconds of `clock`)."""
        t0 = time.perf_counter()
        now = clock or (lambda: time.perf_counter() - t0)
        pending = collections.deque(sorted(arrivals))
        rid = 0
        while pending or self.busy():
            while pending and pending[0][0] <= now():
                t, n, d, seed = pending.popleft()
                self.submit(Request(rid, make_X(n, seed), d, t, seed))
                rid += 1
            if not self.busy():
                time.sleep(max(0.0, min(0.05, pending[0][0] - now())))
                continue
            stepped = self._advance()  # queued on the GPU; returns at once
            if self.front == "async":
                admitted = self._poll_fronts(now)  # never waits for the device
                retired = self._retire(now)
                admitted = self._start_fronts(now) or admitted
            else:
                admitted = self._admit(now, limit=1)  # a front end overlaps the queued iterations
                retired = self._retire(now)
            if not (stepped or admitted or retired):
                time.sleep(self.poll_s)  # every engine has `depth` ticks queued
        return sorted(self.finished, key=lambda r: r.rid)


def one_at_a_time(arrivals, make_X, k=90, u=30.0, max_iter=1000):
    """The plain method, one request after another in arrival order (exact
    kNN, perplexity and sgtsneCUDA-style run per request)."""
    from . import knn_exact, perplexityEqualization, sgtsne, LAYOUT_AOS

    t0 = time.perf_counter()
    out = []
    for i, (t, n, d, seed) in enumerate(sorted(arrivals)):
        X = make_X(n, seed)
        wait = t - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        I, D = knn_exact(X, k)
        P = perplexityEqualization(I, D, n, k, u)
        Y = sgtsne(P, tsneparams(d=d, maxIter=max_iter, layout=LAYOUT_AOS), y_in=random_init(n, d, seed=seed))
        r = Request(i, None, d, t, seed)
        r.done = time.perf_counter() - t0
        r.Y = Y
        out.append(r)
    return out


def latency_stats(reqs):
    lat = np.array([r.done - r.arrival for r in reqs])
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return {"requests": len(reqs), "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4), "mean_s": round(float(lat.mean()), 4),
            "makespan_s": round(float(span), 3), "embeds_per_s": round(len(reqs) / span, 4)}
