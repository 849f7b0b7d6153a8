"""The continuous-batching server (sgtsne_amd.serving) on the GPU.

Bar: every request's embedding is bit-identical whether it is served alone
or in a mix of requests of both embedding dimensions arriving while others
run; the footprint admission never starves a request that fits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make_X(n, seed):
    from sgtsne_amd import synth

    return synth.gaussian_job(n, dim=20, seed=seed)[0]


def _server(sg, **kw):
    from sgtsne_amd import serving

    args = dict(hbm_bytes=4 << 30, max_jobs=8, tick=10, k=30, u=10.0, max_iter=160)
    args.update(kw)
    return serving.Server(**args)


def test_serving_alone_vs_mix_bitexact(sg):
    arrivals = [(0.0, 3000, 2, 11), (0.0, 1500, 3, 12), (0.02, 5000, 2, 13), (0.03, 2500, 2, 14),
                (0.05, 4000, 3, 15), (0.3, 1200, 2, 16)]
    with _server(sg) as srv:
        mix = srv.run(arrivals, _make_X)
    assert len(mix) == len(arrivals)
    for r in mix:
        assert r.Y is not None and np.all(np.isfinite(r.Y)) and r.done >= r.arrival
    for i in (2, 4):  # one 2-D and one 3-D request that arrived while others ran
        t, n, d, seed = arrivals[i]
        with _server(sg) as srv:
            alone = srv.run([(0.0, n, d, seed)], _make_X)
        assert alone[0].Y.shape == (n, d)
        assert np.array_equal(alone[0].Y, mix[i].Y), (n, d)


def test_serving_admission_by_footprint(sg):
    from sgtsne_amd import serving

    # a budget that holds about two of these jobs at once per engine: the
    # rest wait in the queue and are admitted as slots free up
    fp = serving.footprint_bytes(4000, 2, k=30)
    arrivals = [(0.0, 4000, 2, 100 + i) for i in range(5)]
    with _server(sg, hbm_bytes=int(2 * 2.5 * fp / 0.5) // 2, share=(0.5, 0.5), max_iter=60) as srv:
        got = srv.run(arrivals, _make_X)
    assert len(got) == 5 and all(r.Y is not None for r in got)
    admitted = sorted(r.admitted for r in got)
    assert admitted[-1] > admitted[0]  # later ones waited for room
    st = serving.latency_stats(got)
    assert st["p99_s"] >= st["p50_s"] > 0
