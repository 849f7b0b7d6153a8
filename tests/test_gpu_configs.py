"""Parity at the benchmark configurations, through the C-ABI.

  C2  784-D sparse clusters (sparse784 generator at N = 5000): kNN bit-exact,
      P, and the terms of every storage format (COO, hybrid ELL + COO, CSR)
      against the FP64 oracle
  C3  the power-law 50-D mixture at its full N = 1.3M: kNN rows bit-exact on a
      sample of query points, P + P^T sums to 1, and the terms at d = 2 and
      d = 3 (3-D grid FFT) against the FP64 oracle from the same P_sym and Y
  C5  served requests of the serving distribution (50-D, k = 90, u = 30,
      d in {2, 3}): each served result's final KL within 1% of the FP64 loop
      from the same start

Tolerances as everywhere: indices bit-exact, each FP32 term within 1e-4
relative (L2) of the oracle, final KL within 1%."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _clustered(n, d, rng, spread=30.0, clusters=12):
    centres = rng.uniform(-spread, spread, size=(clusters, d))
    lab = rng.integers(0, clusters, n)
    return (centres[lab] + rng.standard_normal((n, d)) * rng.uniform(0.3, 2.0, clusters)[lab, None]).astype(np.float32)


def _check_terms(O, ctx, So, Y, h, tag):
    if Y is not None:  # else the context's current iterate
        ctx.set_Y(Y)
    Yc = ctx.get_Y()
    Fa, Fr, Z, ng = ctx.terms()
    Yd = Yc.astype(np.float64)
    Fao = O.attractive(So, Yd)
    Fro, Zo, ngo = O.repulsive_interp(Yd.copy(), h, O.ZETA_DIRECT)
    errs = (rel_l2(Fa, Fao), rel_l2(Fr, Fro), abs(Z - Zo) / Zo)
    print(tag, "ng", ng, "errors attractive / repulsive / Z", errs, flush=True)
    assert ng == ngo, tag
    assert errs[0] < TOL and errs[1] < TOL and errs[2] < TOL, (tag, errs)


# ---------------------------------------------------------------- C2
def test_c2_sparse784_formats(sg, O):
    from sgtsne_amd import synth

    n, k, u = 5000, 90, 30.0
    X, _ = synth.sparse_clusters(n=n)
    I, D = sg.knn_exact(X, k)
    Io, Do = O.knn_exact(X, k)
    assert np.array_equal(I, Io) and np.array_equal(D, Do)
    P = sg.perplexityEqualization(I, D, n, k, u)
    Po = O.perplexity_equalize(Io, Do, u)
    assert np.array_equal(P.col, Po.col) and np.allclose(P.val, Po.val, rtol=1e-6, atol=1e-12)
    So = O.symmetrize(Po)
    rng = np.random.default_rng(2)
    cases = [(rng.standard_normal((n, 2)) * 1e-2).astype(np.float32), _clustered(n, 2, rng)]
    for fmt in ("COO", "HYBRID", "CSR_SEG", "CSR"):
        with sg.Context(sg.tsneparams(format=getattr(sg, "FORMAT_" + fmt))) as ctx:
            ctx.upload_P(P, symmetrize=True)
            for t, Y in enumerate(cases):
                _check_terms(O, ctx, So, Y, 0.7, f"C2 {fmt} case {t}")


# ---------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3(sg):
    from sgtsne_amd import synth

    print("C3: generating 1.3M x 50-D", flush=True)
    X, _ = synth.powerlaw_mixture()
    n = X.shape[0]
    print("C3: kNN k=90 on the GPU", flush=True)
    I, D, fb = sg.knn(X, 90, sg.KNN_AUTO)
    print("C3: kNN done, rows recomputed exactly:", fb, flush=True)
    P = sg.perplexityEqualization(I, D, n, 90, 30.0)
    return X, I, D, P


def test_c3_knn_rows_sample(sg, O, c3):
    X, I, D, _ = c3
    rows = np.sort(np.random.default_rng(5).choice(X.shape[0], 48, replace=False)).astype(np.int32)
    Io, Do = O.knn_rows(X, 90, rows)
    assert np.array_equal(I[rows], Io) and np.array_equal(D[rows], Do)


@pytest.mark.parametrize("d", [2, 3])
def test_c3_terms_full_size(sg, O, c3, d):
    # the terms at a spread-out start and at the run's own iterates (after the
    # exaggeration phase and at the end of a 1000-iteration run), each from
    # the same P_sym and Y as the FP64 oracle
    from sgtsne_amd import synth

    X, _, _, P = c3
    n = X.shape[0]
    h = {2: 0.7, 3: 1.2}[d]
    rng = np.random.default_rng(30 + d)
    with sg.Context(sg.tsneparams(d=d)) as ctx:
        ctx.upload_P(P, symmetrize=True)
        S = ctx.get_P()
        assert abs(S.val.sum(dtype=np.float64) - 1.0) < 1e-4
        So = O.CSR(n, S.rowptr.astype(np.int64), S.col, S.val.astype(np.float64))
        del S
        _check_terms(O, ctx, So, (rng.standard_normal((n, d)) * 3.0).astype(np.float32), h, f"C3 d={d} spread start")
        ctx.set_Y(synth.random_init(n, d, seed=1))
        ctx.step(300)
        _check_terms(O, ctx, So, None, h, f"C3 d={d} iteration 300")
        ctx.step(700)
        _check_terms(O, ctx, So, None, h, f"C3 d={d} iteration 1000")


# ---------------------------------------------------------------- C5
def test_c5_served_results_match_oracle_kl(sg, O):
    # requests of the serving distribution (50-D Gaussian-mixture jobs, k = 90,
    # u = 30, d in {2, 3}, 1000 iterations) served as a mix; every result's
    # final KL against the FP64 loop from the request's own start
    from sgtsne_amd import serving, synth

    arrivals = [(0.0, 2600, 2, 501), (0.0, 1100, 3, 502), (0.05, 1800, 2, 503), (0.1, 900, 3, 504)]

    def make_X(n, seed):
        return synth.gaussian_job(n, dim=50, seed=seed)[0]

    with serving.Server(hbm_bytes=8 << 30, max_jobs=8, tick=10, k=90, u=30.0, max_iter=1000) as srv:
        got = srv.run(arrivals, make_X)
    assert len(got) == len(arrivals)
    for r, (t, n, d, seed) in zip(got, sorted(arrivals)):
        assert r.Y is not None and r.Y.shape == (n, d) and np.all(np.isfinite(r.Y))
        X = make_X(n, seed)
        I, D = O.knn_exact(X, 90)
        S = O.symmetrize(O.perplexity_equalize(I, D, 30.0))
        Yo = synth.random_init(n, d, seed=seed).astype(np.float64)
        h = {2: 0.7, 3: 1.2}[d]
        O.run(S, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, 1000, h=h, zmode=O.ZETA_DIRECT)
        _, Zo, _ = O.repulsive_interp(Yo.copy(), h, O.ZETA_DIRECT)
        Yg = r.Y.astype(np.float64)
        _, Zg, _ = O.repulsive_interp(Yg.copy(), h, O.ZETA_DIRECT)
        klo, klg = O.kl(S, Yo, Zo), O.kl(S, Yg, Zg)
        print(f"C5 n={n} d={d}: served KL {klg:.6f}, oracle {klo:.6f}", flush=True)
        assert abs(klg - klo) / klo < 0.01, (n, d, klg, klo)
