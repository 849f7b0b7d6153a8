"""Continuous-batching server of embedding requests on one GPU.

A request is a point set X [n][dim] and an embedding dimension d (2 or 3).
The server keeps one batched engine per d (sgtsne_batch_*), so P, Y and the
grids of every admitted job stay resident on the device.  Its loop:

  1. queue the next `tick` iterations of every engine (asynchronous), keeping
     at most `depth` ticks in flight per engine: the loop waits for an
     engine's oldest queued tick, never for its whole queue, so the device
     always has work while the host admits and retires;
  2. while the GPU runs them, admit queued requests whose HBM footprint fits,
     3-D requests first (their iterations are the longest, so starting them
     early shortens the makespan), then in arrival order: the front end (kNN,
     perplexity, P + P^T) runs on the engine's second stream, overlapping the
     queued iterations;
  3. retire the jobs that reached max_iter: a finished job's Y is read as soon
     as its own last iteration is done, then its slot is freed.

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

_POINT_BYTES = {2: 6 * 2 * 4 + 4, 3: 6 * 4 * 4 + 4}  # Y, uY, gains, Fattr, Frep, (spare) + perm


def footprint_bytes(n, d, k=90):
    """Device bytes a job holds while admitted: P_sym (<= 2 n k nonzeros,
    8 B each), the per-point state and the grid buffers of its capacity."""
    grid = {2: (512 if n > 20000 else 256), 3: (128 if n > 20000 else 64)}[d]
    L = 1
    while L < 2 * grid - 1:
        L <<= 1
    G = grid ** d
    H = L // 2 + 1
    grid_b = G * (d + 1) * 8 + G * (d + 2) * 4 + 3 * grid * L ** (d - 1) * 8 + 3 * H ** d * 8
    return 8 * 2 * n * k + _POINT_BYTES[d] * n + grid_b


class Request:
    __slots__ = ("rid", "X", "d", "arrival", "seed", "admitted", "done", "Y", "engine", "job", "t_left")

    def __init__(self, rid, X, d, arrival, seed=0):
        self.rid, self.X, self.d, self.arrival, self.seed = rid, X, d, arrival, seed
        self.admitted = self.done = None
        self.Y = None
        self.engine = self.job = None
        self.t_left = 0


class Server:
    """hbm_bytes is split between the d=2 and d=3 engines in proportion to
    `share`; each engine admits jobs while their summed footprint fits."""

    def __init__(self, hbm_bytes=48 << 30, max_jobs=256, tick=10, k=90, u=30.0, max_iter=1000,
                 share=(0.5, 0.5), seed=0, depth=3):
        self.k, self.u, self.max_iter, self.tick, self.depth = k, u, max_iter, tick, depth
        self.engines = {}
        self.budget = {}
        self.used = {}
        for d, frac in zip((2, 3), share):
            b = int(hbm_bytes * frac)
            nnz = min(int(0.6 * b) // 8, 2_000_000_000)   # P_sym: 8 B per nonzero
            pts = max(1, nnz // k)                         # >= the points of any mix (>= k nonzeros per row)
            grid_budget = b  # the footprint check below is the binding limit
            self.engines[d] = Batch(tsneparams(d=d, maxIter=max_iter, seed=seed), max_jobs, pts, nnz,
                                    grid_budget)
            self.budget[d] = b
            self.used[d] = 0
        self.queue = collections.deque()
        self.active = {2: {}, 3: {}}
        self.inflight = {2: collections.deque(), 3: collections.deque()}  # marks of queued ticks
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

    def _admit(self, now_fn):
        # first fit, 3-D requests first, then in arrival order: a large request
        # at the head does not block smaller ones behind it
        keep = collections.deque()
        order = sorted(self.queue, key=lambda q: (-q.d, q.arrival, q.rid))
        self.queue = collections.deque()
        for r in order:
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
            r.engine, r.job, r.admitted = r.d, job, now_fn()
            r.t_left = self.max_iter + 1  # one admission iteration, then max_iter
            r.X = None
            self.used[r.d] += fp
            self.active[r.d][job] = (r, fp)
        self.queue = collections.deque(sorted(keep, key=lambda q: (q.arrival, q.rid)))

    def _advance(self):
        for d, act in self.active.items():
            if act:
                eng, marks = self.engines[d], self.inflight[d]
                while len(marks) >= self.depth:
                    eng.wait(marks.popleft())
                eng.step(self.tick)
                marks.append(eng.mark())
                for r, _ in act.values():
                    r.t_left -= self.tick

    def _retire(self, now_fn):
        for d, act in self.active.items():
            for job in [j for j, (r, _) in act.items() if r.t_left <= 0]:
                r, fp = act.pop(job)
                b = self.engines[d]
                st, it = b.job_status(job)
                r.Y = b.get_Y(job) if st >= 0 else None
                r.done = now_fn()
                b.remove(job)
                self.used[d] -= fp
                self.finished.append(r)

    def busy(self):
        return bool(self.queue) or any(self.active[d] for d in self.active)

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
            self._advance()       # queued on the GPU; returns at once
            self._admit(now)      # front ends overlap the queued iterations
            self._retire(now)
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
