"""The batched engine (sgtsne_batch_*), through the C-ABI.

Bars: each job's per-iteration terms within 1e-4 relative of the FP64 oracle
(same bar as the single-job path); a job's trajectory bit-identical whether
it runs alone or among other jobs (admitted earlier or later, removed on the
way), with or without graph replay; final KL of a 1000-iteration job within
1% of the FP64 loop."""
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _mixture(n, dim, clusters, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 4, size=(clusters, dim))
    return (c[rng.integers(0, clusters, n)] + rng.standard_normal((n, dim))).astype(np.float32)


def _job(sg, O, n, seed, k=45, u=15.0, dim=20):
    X = _mixture(n, dim, 6, seed)
    I, D = O.knn_exact(X, k)
    So = O.symmetrize(O.perplexity_equalize(I, D, u))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    return S, So


def _clustered(n, d, rng, spread=30.0, clusters=12):
    centres = rng.uniform(-spread, spread, size=(clusters, d))
    lab = rng.integers(0, clusters, n)
    return (centres[lab] + rng.standard_normal((n, d)) * rng.uniform(0.3, 2.0, clusters)[lab, None]).astype(np.float32)


def _batch(sg, d=2, max_jobs=16, points=200_000, nnz=20_000_000, **kw):
    return sg.Batch(sg.tsneparams(d=d, **kw), max_jobs, points, nnz)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_batch_terms_vs_oracle(sg, O, d):
    # three jobs of uneven size in one batch; the terms of each at a formed
    # (clustered) embedding and at an early one, against the FP64 oracle
    sizes = (700, 3100, 5200)
    rng = np.random.default_rng(20 + d)
    h = {1: 0.5, 2: 0.7, 3: 1.2}[d]
    with _batch(sg, d=d) as B:
        jobs = []
        for t, n in enumerate(sizes):
            S, So = _job(sg, O, n, 100 * d + t)
            Y = _clustered(n, d, rng) if t != 1 else (rng.standard_normal((n, d)) * 2.0).astype(np.float32)
            jobs.append((B.add(S, symmetrize=False, y_in=Y), So))
        B.step(1)  # admission: centring and the first grid
        for j, So in jobs:
            Y = B.get_Y(j).astype(np.float64)
            Fa, Fr, Z, ng = B.terms(j)
            Fao = O.attractive(So, Y)
            Fro, Zo, ngo = O.repulsive_interp(Y.copy(), h, O.ZETA_DIRECT)
            assert ng == ngo
            assert rel_l2(Fa, Fao) < TOL
            assert rel_l2(Fr, Fro) < TOL
            assert abs(Z - Zo) / Zo < TOL


# every shape the batch autotuner times (unroll, rows per work-group, points
# of Y in LDS; batch.hip tune()); unroll 8 and 16 are the masked-step forms
_BATCH_CANDIDATES = [(1, 0, 0), (2, 0, 0), (4, 0, 0), (8, 0, 0), (16, 0, 0), (2, 32, 0), (2, 64, 0), (1, 64, 0),
                     (2, 128, 0), (1, 128, 0), (2, 256, 0), (1, 256, 4096), (2, 512, 4096), (1, 512, 8192),
                     (2, 1024, 8192), (16, 64, 0), (16, 256, 0), (8, 256, 0)]


@pytest.mark.parametrize("d", [1, 2, 3])
def test_batch_attractive_candidates_vs_oracle(sg, O, d):
    # each candidate pinned in turn: every job's Fattr against the FP64 oracle,
    # and the same bits for every candidate (the lanes per row are fixed)
    sizes = (1301, 4099, 2203)
    rng = np.random.default_rng(60 + d)
    with _batch(sg, d=d) as B:
        jobs = []
        for t, n in enumerate(sizes):
            S, So = _job(sg, O, n, 300 * d + t)
            jobs.append((B.add(S, symmetrize=False, y_in=_clustered(n, d, rng)), So))
        B.step(1)
        Yc = {j: B.get_Y(j).astype(np.float64) for j, _ in jobs}
        ref = {j: O.attractive(So, Yc[j]) for j, So in jobs}
        first = None
        for cand in _BATCH_CANDIDATES:
            u, rpb, win = cand
            if win * (4 if d == 3 else d) * 4 > 65536:
                with pytest.raises(sg.SgtsneError):
                    B.set_attractive(u, rpb, win)
                continue
            B.set_attractive(u, rpb, win)
            t = B.tuning()
            assert (t["unroll"], t["rows_per_block"], t["window"]) == cand
            got = {}
            for j, _ in jobs:
                Fa, _, _, _ = B.terms(j)
                assert rel_l2(Fa, ref[j]) < TOL, (cand, j)
                got[j] = Fa
            if first is None:
                first = got
            for j in got:
                assert np.array_equal(got[j], first[j]), (cand, j)


def test_batch_algebraic_zeta(sg, O):
    S, So = _job(sg, O, 2000, 5)
    Y = _clustered(2000, 2, np.random.default_rng(3))
    with _batch(sg, zeta_mode=sg.ZETA_ALGEBRAIC) as B:
        j = B.add(S, symmetrize=False, y_in=Y)
        B.step(1)
        Yc = B.get_Y(j).astype(np.float64)
        _, Fr, Z, _ = B.terms(j)
    Fro, Zo, _ = O.repulsive_interp(Yc.copy(), 0.7, O.ZETA_ALGEBRAIC)
    assert rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL


def test_batch_iterations_track_oracle(sg, O):
    # ten iterations of a batched job against the FP64 loop from the same start
    n = 2500
    S, So = _job(sg, O, n, 7)
    y0 = (np.random.default_rng(4).standard_normal((n, 2)) * 1e-2).astype(np.float32)
    with _batch(sg) as B:
        other = B.add(*_job(sg, O, 1800, 8)[:1], symmetrize=False)
        j = B.add(S, symmetrize=False, y_in=y0)
        B.step(1)
        Yc = B.get_Y(j).astype(np.float64)
        B.step(10)
        Y10 = B.get_Y(j)
        st, it = B.job_status(j)
        assert st == sg.JOB_RUNNING and it == 10
    Yo = Yc.copy()
    O.run(So, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, 10, zmode=O.ZETA_DIRECT)
    assert rel_l2(Y10, Yo) < 1e-3


def _run_alone(sg, S, y0, iters, graph=True):
    with _batch(sg) as B:
        B.set_graph(graph)
        j = B.add(S, symmetrize=False, y_in=y0)
        B.step(iters + 1)
        return B.get_Y(j)


def test_batch_alone_vs_mix_bitexact(sg, O):
    # job A's trajectory (through the re-orders at iterations 50 and 150) is
    # the same bits alone and in a mix where other jobs arrive before and after
    # it and one of them leaves (forcing a compaction that moves A)
    SA, _ = _job(sg, O, 3000, 31)
    SB, _ = _job(sg, O, 5000, 32)
    SC, _ = _job(sg, O, 1200, 33)
    y0 = (np.random.default_rng(5).standard_normal((3000, 2)) * 1e-4).astype(np.float32)
    iters = 170
    Y_alone = _run_alone(sg, SA, y0, iters)
    with _batch(sg, max_jobs=4) as B:
        b = B.add(SB, symmetrize=False)
        B.step(37)
        a = B.add(SA, symmetrize=False, y_in=y0)
        c = B.add(SC, symmetrize=False)
        B.step(60)
        B.remove(b)  # A moves down over B's rows
        B.step(iters + 1 - 60)
        st, it = B.job_status(a)
        assert st == sg.JOB_RUNNING and it == iters
        Y_mix = B.get_Y(a)
        assert B.info()["jobs"] == 2
    assert np.array_equal(Y_alone, Y_mix)


def test_batch_graph_replay_equals_launches(sg, O):
    S, _ = _job(sg, O, 2200, 41)
    y0 = (np.random.default_rng(6).standard_normal((2200, 2)) * 1e-4).astype(np.float32)
    Yg = _run_alone(sg, S, y0, 80, graph=True)
    Yl = _run_alone(sg, S, y0, 80, graph=False)
    assert np.array_equal(Yg, Yl)


def test_batch_final_kl(sg, O):
    # 1000 iterations of a batched job (alongside two others): final KL within
    # 1% of the FP64 loop from the same start
    n = 2500
    X = _mixture(n, 50, 10, 12)
    I, D = sg.knn_exact(X, 90)
    P = sg.perplexityEqualization(I, D, n, 90, 30.0)
    So = O.symmetrize(O.CSR(n, P.rowptr.astype(np.int64), P.col, P.val.astype(np.float64)))
    y0 = (1e-4 * np.random.default_rng(1).standard_normal((n, 2))).astype(np.float32)
    with _batch(sg) as B:
        B.add(_job(sg, O, 900, 51)[0], symmetrize=False)
        j = B.add(P, symmetrize=True, y_in=y0)
        B.add(_job(sg, O, 4000, 52)[0], symmetrize=False, max_iter=400)
        B.step(1001)
        st, it = B.job_status(j)
        assert st == sg.JOB_DONE and it == 1000
        Y = B.get_Y(j).astype(np.float64)
    Yo = y0.astype(np.float64).copy()
    O.run(So, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, 1000, zmode=O.ZETA_DIRECT)
    _, Zg, _ = O.repulsive_interp(Y.copy(), 0.7, O.ZETA_DIRECT)
    _, Zo, _ = O.repulsive_interp(Yo.copy(), 0.7, O.ZETA_DIRECT)
    klg = O.kl(So, Y, Zg)
    klo = O.kl(So, Yo, Zo)
    assert abs(klg - klo) / klo < 0.01, (klg, klo)


def test_batch_add_points_front_end(sg, O):
    # the device front end (exact kNN, perplexity, P + P^T) gives the same
    # P_sym as the host-visible path: the attractive terms agree
    n = 1500
    X = _mixture(n, 30, 5, 61)
    I, D = O.knn_exact(X, 60)
    So = O.symmetrize(O.perplexity_equalize(I, D, 20.0))
    Y = _clustered(n, 2, np.random.default_rng(7))
    with _batch(sg) as B:
        j = B.add_points(X, k=60, u=20.0, y_in=Y)
        B.step(1)
        Yc = B.get_Y(j).astype(np.float64)
        Fa, _, _, _ = B.terms(j)
    assert rel_l2(Fa, O.attractive(So, Yc)) < TOL


def test_batch_admission_limits(sg, O):
    S, _ = _job(sg, O, 1000, 71)
    with sg.Batch(sg.tsneparams(), 2, 2500, 10_000_000) as B:
        a = B.add(S, symmetrize=False)
        B.add(S, symmetrize=False)
        with pytest.raises(sg.Full):
            B.add(S, symmetrize=False)  # no slot
        B.remove(a)
        with pytest.raises(sg.Full):
            B.add(_job(sg, O, 1600, 72)[0], symmetrize=False)  # 1000 + 1600 > 2500 points
        B.add(S, symmetrize=False)
        B.step(3)
        assert B.info()["jobs"] == 2


def test_batch_numeric_error_is_reported(sg, O):
    S, _ = _job(sg, O, 800, 81)
    Y = np.zeros((800, 2), np.float32)
    Y[3, 1] = np.nan
    with _batch(sg) as B:
        bad = B.add(S, symmetrize=False, y_in=Y)
        good = B.add(S, symmetrize=False)
        B.step(5)
        assert B.job_status(bad)[0] == -4
        st, it = B.job_status(good)
        assert st == sg.JOB_RUNNING and it == 4
        assert np.all(np.isfinite(B.get_Y(good)))


def test_batch_marks_wait_and_passed(sg, O):
    # marks are points on the iteration stream: once the latest one is waited
    # for, every earlier one (including those whose event slot was reused by
    # a later mark) reads as passed; an unknown mark is an argument error
    S, _ = _job(sg, O, 1500, 51)
    with _batch(sg, max_jobs=2) as B:
        B.add(S, symmetrize=False)
        marks = []
        for _ in range(40):
            B.step(2)
            marks.append(B.mark())
        assert marks == list(range(marks[0], marks[0] + 40))
        B.wait(marks[-1])
        assert all(B.passed(m) for m in marks)
        B.step(5)
        m = B.mark()
        B.sync()
        assert B.passed(m)
        with pytest.raises(sg.SgtsneError):
            B.passed(m + 1)


def test_batch_job_left(sg, O):
    # job_left is the host's count of batch iterations a job still needs
    from sgtsne_amd import synth

    X, _ = synth.gaussian_job(1800, seed=61)
    y0 = synth.random_init(1800, 2, seed=61)
    with _batch(sg, max_jobs=4) as B:
        a = B.add_points(X, k=30, u=10.0, y_in=y0, max_iter=40)
        assert B.job_left(a) == 41
        B.step(15)
        assert B.job_left(a) == 26
        B.step(30)
        assert B.job_left(a) == 0
        st, it = B.job_status(a)
        assert st == sg.JOB_DONE and it == 40


def test_batch_async_front_end_same_bits(sg, O):
    # the front end in parts (add_points_async + non-blocking polls) while
    # iterations of another job are queued between the polls, on one host
    # thread: the admitted job runs the same bits as when admitted by
    # add_points alone; several front ends in flight at once do too
    from sgtsne_amd import synth

    X, _ = synth.gaussian_job(1800, seed=61)
    y0 = synth.random_init(1800, 2, seed=61)
    with _batch(sg, max_jobs=4) as B:
        a = B.add_points(X, k=30, u=10.0, y_in=y0, max_iter=40)
        B.step(41)
        Y_alone = B.get_Y(a)
    S, _ = _job(sg, O, 2500, 62)
    with _batch(sg, max_jobs=12) as B:
        other = B.add(S, symmetrize=False, max_iter=400)
        B.step(20)
        t = B.add_points_async(X, k=30, u=10.0, y_in=y0, max_iter=40)
        with pytest.raises(sg.SgtsneError):
            B.poll_points(t + 1)  # no such ticket
        steps, polls, a = 0, 0, None
        while a is None and polls < 5000:
            a = B.poll_points(t)
            polls += 1
            if a is None and steps < 300:
                B.step(1)
                steps += 1
        assert a is not None, polls
        assert polls >= 3  # each of the three parts is taken on by a poll of its own
        left = B.job_left(a)
        assert left == 41
        B.step(left)
        st, it = B.job_status(a)
        assert st == sg.JOB_DONE and it == 40
        assert np.array_equal(B.get_Y(a), Y_alone)
        assert B.job_left(other) == max(0, 401 - 20 - steps - left)
        # eight front ends in flight (a ninth is refused), polled in turn:
        # every job admitted runs the same bits as the one admitted alone
        ts = [B.add_points_async(X, k=30, u=10.0, y_in=y0, max_iter=40) for _ in range(8)]
        with pytest.raises(sg.SgtsneError):
            B.add_points_async(X, k=30, u=10.0, y_in=y0, max_iter=40)
        jobs = {}
        t_end = time.perf_counter() + 60.0
        while len(jobs) < len(ts) and time.perf_counter() < t_end:
            for t in ts:
                if t not in jobs:
                    j = B.poll_points(t)
                    if j is not None:
                        jobs[t] = j
        assert len(jobs) == len(ts)
        assert B.poll_points(B.add_points_async(X, k=30, u=10.0, y_in=y0, max_iter=40), wait=True) is not None
        B.step(41)
        for j in jobs.values():
            assert np.array_equal(B.get_Y(j), Y_alone)


_THREADED_FRONT_END_FAULT = (
    "a front end on a second host thread overlapping steps on the first (the served loop with a worker-thread "
    "admission) faulted the GPU once; no cause was found in the code, so batches are called from one thread "
    "and this case stays off the GPU")


@pytest.mark.skip(reason=_THREADED_FRONT_END_FAULT)
def test_batch_threaded_front_end(sg, O):
    # a front end on a second thread (add_points) overlapping steps on this
    # one; the job it admits runs the same bits as when admitted alone
    import threading

    from sgtsne_amd import synth

    X, _ = synth.gaussian_job(1800, seed=61)
    y0 = synth.random_init(1800, 2, seed=61)
    with _batch(sg, max_jobs=4) as B:
        a = B.add_points(X, k=30, u=10.0, y_in=y0, max_iter=40)
        B.step(41)
        Y_alone = B.get_Y(a)
    S, _ = _job(sg, O, 2500, 62)
    with _batch(sg, max_jobs=4) as B:
        other = B.add(S, symmetrize=False, max_iter=400)
        box = {}
        t = threading.Thread(target=lambda: box.setdefault("job", B.add_points(X, k=30, u=10.0, y_in=y0,
                                                                               max_iter=40)))
        t.start()
        steps = 0
        while t.is_alive() and steps < 300:
            B.step(1)
            steps += 1
        t.join()
        a = box["job"]
        left = B.job_left(a)
        assert 0 <= left <= 41
        B.step(left)
        st, it = B.job_status(a)
        assert st == sg.JOB_DONE and it == 40
        assert np.array_equal(B.get_Y(a), Y_alone)
        assert B.job_left(other) == max(0, 401 - steps - left)


def test_batch_staging_ring_wrap_keeps_uploads(sg, O):
    # a job whose Y fills the 16 MiB pinned staging ring exactly (2^21 points
    # in 2-D), admitted while another job's iterations are queued: its job
    # record then wraps onto the start of the ring, where Y's copy may still
    # be waiting.  The ring commits and waits before reusing that memory, so
    # the Y read back equals the caller's bit for bit.
    S_small, _ = _job(sg, O, 3000, 41)
    n = 1 << 21
    idx = np.arange(n)
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    col = np.stack([(idx - 1) % n, (idx + 1) % n], 1).reshape(-1).astype(np.int32)
    val = np.full(2 * n, 0.5 / n, np.float32)
    y0 = np.random.default_rng(9).standard_normal((n, 2)).astype(np.float32)
    with _batch(sg, max_jobs=4, points=n + 8192, nnz=2 * n + 2_000_000, layout=sg.LAYOUT_AOS) as B:
        B.add(S_small, symmetrize=False)
        B.step(60)  # queued: the big job's copies wait behind these
        big = B.add(sg.CSR(n, rp, col, val), symmetrize=False, y_in=y0, max_iter=2)
        assert np.array_equal(B.get_Y(big), y0)
