"""The serving loop's host logic (sgtsne_amd.serving.Server) on the CPU,
against a stand-in engine that models a device queue: queued batch
iterations run a few at a time, each time the host looks at the device.

Checks: every request finishes; a job is retired only after the device has
run all its iterations (one admission iteration + max_iter); no engine ever
has more than `depth` ticks queued; 3-D requests are admitted ahead of 2-D
ones waiting with them; a request whose footprint can never fit is rejected
at once; at most one admission per loop pass."""
import threading
import time

import numpy as np
import pytest

from sgtsne_amd import serving


class FakeEngine:
    """Batch's serving surface over a modelled queue of batch iterations."""

    def __init__(self, d, log, speed, front_end_s=0.0, front_polls=3):
        self.d, self.log, self.speed, self.front_end_s = d, log, speed, front_end_s
        self.front_polls = front_polls  # polls until an asynchronous front end admits its job
        self.front = None
        self.lock = threading.Lock()  # the library's batch mutex
        self.host_iter = 0     # batch iterations queued
        self.dev_iter = 0      # batch iterations the device has run
        self.mark_at = []      # host_iter at each mark
        self.jobs = {}
        self.next_job = 0
        self.max_queued = 0

    def _run(self):
        self.dev_iter = min(self.host_iter, self.dev_iter + self.speed)

    def add_points(self, X, k, u, y_in, max_iter):
        time.sleep(self.front_end_s)  # the front end holds no lock
        with self.lock:
            self._run()
            job = self.next_job
            self.next_job += 1
            self.jobs[job] = (self.host_iter, max_iter, X.shape[0])
            self.log.append(("admit", self.d, X.shape[0]))
            return job

    def add_points_async(self, X, k, u, y_in, max_iter):
        assert self.front is None, "two front ends in flight on one engine"
        self.front = [X, max_iter, self.front_polls, len(self.log)]
        self.log.append(("start", self.d, X.shape[0]))
        return len(self.log)

    def poll_points(self, ticket, wait=False):
        f = self.front
        assert f is not None and ticket == f[3] + 1
        f[2] -= 1
        if f[2] > 0 and not wait:
            return None
        self.front = None
        return self.add_points(f[0], None, None, None, f[1])

    def step(self, iters):
        with self.lock:
            self.host_iter += iters
            self.max_queued = max(self.max_queued, self.host_iter - self.dev_iter)

    def mark(self):
        with self.lock:
            self.mark_at.append(self.host_iter)
            return len(self.mark_at) - 1

    def passed(self, m):
        with self.lock:
            self._run()
            return self.dev_iter >= self.mark_at[m]

    def wait(self, m):
        with self.lock:
            self.dev_iter = max(self.dev_iter, self.mark_at[m])

    def job_left(self, job):
        with self.lock:
            start, max_iter, _ = self.jobs[job]
            return max(0, max_iter + 1 - (self.host_iter - start))

    def job_status(self, job):
        assert self.front is None, "a read queued behind a front end in flight"
        with self.lock:
            start, max_iter, _ = self.jobs[job]
            done = self.dev_iter - start
            assert done >= max_iter + 1, "retired before its last iteration ran"
            return 1, max_iter

    def get_Y(self, job):
        return np.zeros((self.jobs[job][2], self.d), np.float32)

    def remove(self, job):
        with self.lock:
            del self.jobs[job]
            self.log.append(("remove", self.d))

    def close(self):
        pass


def _server(log, speed=7, front_end_s=0.0, front_polls=3, **kw):
    engines = {}

    def factory(params, max_jobs, pts, nnz, grid_budget):
        e = FakeEngine(params.d, log, speed, front_end_s, front_polls)
        engines[params.d] = e
        return e

    args = dict(hbm_bytes=8 << 30, tick=10, k=30, max_iter=60, depth=3, poll_s=0.0, engine_factory=factory)
    args.update(kw)
    return serving.Server(**args), engines


def test_server_finishes_every_request_after_its_iterations():
    log = []
    srv, eng = _server(log, front_end_s=0.001)
    arrivals = [(0.0005 * i, 500 + 37 * i, 2 + (i % 2), i) for i in range(12)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 12 and all(r.Y is not None for r in got)
    assert [r.rid for r in got] == list(range(12))
    for e in eng.values():
        assert e.max_queued <= 3 * 10
        assert not e.jobs
    assert all(r.admitted >= r.arrival and r.done >= r.admitted for r in got)


@pytest.mark.parametrize("front", ["async", "sync"])
def test_server_both_front_end_modes_finish(front):
    log = []
    srv, eng = _server(log, front=front, front_end_s=0.0005)
    arrivals = [(0.0003 * i, 400 + 53 * i, 2 + (i % 2), i) for i in range(10)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 10 and all(r.Y is not None for r in got)
    assert sum(1 for x in log if x[0] == "admit") == 10
    assert sum(1 for x in log if x[0] == "start") == (10 if front == "async" else 0)


def test_server_async_front_ends_keep_engines_fed():
    # while one engine's front end is in flight, the loop keeps queuing ticks
    # for the jobs already admitted (each poll returns at once), holds at most
    # one front end per engine and starts the 3-D one first
    log = []
    srv, eng = _server(log, speed=2, front_polls=40, max_iter=200)
    steps = []
    for e in eng.values():
        orig = e.step

        def step(iters, e=e, orig=orig):
            steps.append((e.d, len(log)))
            orig(iters)

        e.step = step
    arrivals = [(0.0, 600, 2, 0), (0.0, 700, 3, 1), (0.0, 500, 2, 2), (0.0, 800, 3, 3)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 4 and all(r.Y is not None for r in got)
    starts = [i for i, x in enumerate(log) if x[0] == "start"]
    assert log[starts[0]][1] == 3  # the 3-D engine's front end goes first
    # ticks queued on the 2-D engine while the second 2-D front end was in flight
    s2 = [i for i in starts if log[i][1] == 2][1]
    a2 = log.index(("admit", 2, log[s2][2]))
    assert any(d == 2 and s2 < at <= a2 for d, at in steps)
    # at most one front end per engine: each start is followed by its
    # engine's admission before that engine's next start
    for d in (2, 3):
        seq = [x[0] for x in log if x[0] in ("start", "admit") and x[1] == d]
        assert seq == ["start", "admit"] * 2


def test_server_async_front_end_full_is_retried_or_rejected():
    # an admission that finds the engine full (sgtsne_batch_add_poll returns
    # SGTSNE_ERR_FULL after the front end) puts the request back in the queue
    # while the engine has jobs, and rejects it when the engine is empty
    log = []
    srv, eng = _server(log, max_iter=30, front_polls=2)
    e2 = eng[2]
    orig = e2.poll_points
    state = {"full_once": True}

    def poll(ticket, wait=False):
        if state["full_once"] and e2.jobs:
            state["full_once"] = False
            e2.front = None
            raise serving.Full("arena full")
        return orig(ticket, wait)

    e2.poll_points = poll
    arrivals = [(0.0, 500, 2, 0), (0.0, 600, 2, 1)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 2 and all(r.Y is not None for r in got)
    assert not state["full_once"]  # the second request was refused once, then admitted
    assert sum(1 for x in log if x == ("start", 2, 600)) == 2

    log = []
    srv, eng = _server(log, max_iter=30)

    def always_full(ticket, wait=False):
        eng[3].front = None
        raise serving.Full("never fits")

    eng[3].poll_points = always_full
    with srv:
        got = srv.run([(0.0, 700, 3, 5)], lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 1 and got[0].Y is None  # an empty engine that cannot take it: rejected


def test_server_admits_3d_first_one_per_pass_and_rejects_oversize():
    log = []
    srv, eng = _server(log, hbm_bytes=1 << 30)
    too_big = 2_000_000  # its footprint exceeds the engine's whole share
    assert serving.footprint_bytes(too_big, 2, k=30) > srv.budget[2]
    for rid, (n, d) in enumerate([(400, 2), (300, 2), (too_big, 2), (500, 3), (600, 3)]):
        srv.submit(serving.Request(rid, np.zeros((n, 4), np.float32), d, 0.0, rid))
    assert srv._admit(lambda: 0.0, limit=1) == 1
    assert log == [("admit", 3, 500)]
    assert srv._admit(lambda: 0.0, limit=1) == 1
    assert log[-1] == ("admit", 3, 600)
    srv._admit(lambda: 0.0)  # no limit: the rest that fit, in arrival order
    assert [x for x in log if x[0] == "admit"][2:] == [("admit", 2, 400), ("admit", 2, 300)]
    rejected = [r for r in srv.finished if r.Y is None]
    assert len(rejected) == 1 and rejected[0].rid == 2
    assert not srv.queue
    srv.close()


@pytest.mark.parametrize("speed", [1, 50])
def test_server_depth_bounds_queued_ticks(speed):
    # a slow device (1 iteration per look) and a fast one: never more than
    # `depth` ticks queued, and the fast one never leaves a finished job waiting
    log = []
    srv, eng = _server(log, speed=speed, depth=2, max_iter=40)
    arrivals = [(0.0, 800, 2, 0), (0.0, 900, 3, 1), (0.0002, 700, 2, 2)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 3
    for e in eng.values():
        assert e.max_queued <= 2 * 10


def test_server_depth_per_engine():
    # depth may differ per engine: {d: ticks in flight}
    log = []
    srv, eng = _server(log, speed=1, depth={2: 1, 3: 2}, max_iter=40)
    arrivals = [(0.0, 800, 2, 0), (0.0, 900, 3, 1)]
    with srv:
        got = srv.run(arrivals, lambda n, s: np.zeros((n, 4), np.float32))
    assert len(got) == 2
    assert eng[2].max_queued <= 1 * 10 and eng[3].max_queued <= 2 * 10


def test_server_refuses_unvalidated_queue_depths():
    # on real engines the server refuses queues deeper than the setting it was
    # validated with (a tick 20 / depth 6 sweep faulted the GPU, cause unknown)
    from sgtsne_amd import serving

    for kw in ({"tick": 20}, {"depth": 6}, {"depth": {2: 3, 3: 4}}):
        with pytest.raises(ValueError, match="supported"):
            serving.Server(**kw)
