"""GPU parity of the HIP path against the FP64 oracle, through the C-ABI.

Tolerances: integer / index outputs bit-exact; FP32 terms within 1e-4
relative (L2 norm) of the FP64 oracle from identical P and Y; Z within 1e-4
relative; final KL within 1%."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-4


def _mixture(n, dim, clusters, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 4, size=(clusters, dim))
    return (c[rng.integers(0, clusters, n)] + rng.standard_normal((n, dim))).astype(np.float32)


def _oracle_csr(O, P):
    return O.CSR(P.n, P.rowptr.astype(np.int64), P.col, P.val.astype(np.float64))


@pytest.fixture(scope="module")
def mnist():
    return np.load(os.path.join(GOLD, "mnist500.npz")), np.load(os.path.join(GOLD, "mnist500_oracle.npz"))


# ---------------------------------------------------------------- kNN
def test_knn_bitexact_golden(sg, mnist):
    f, g = mnist
    I, D = sg.knn_exact(f["X"], 90)
    assert np.array_equal(I, g["I"])
    assert np.array_equal(D, g["D"])


@pytest.mark.parametrize("n,dim,k", [(1000, 50, 90), (777, 7, 30), (130, 3, 129)])
def test_knn_bitexact_random(sg, O, n, dim, k):
    X = _mixture(n, dim, 5, n)
    I, D = sg.knn_exact(X, k)
    Io, Do = O.knn_exact(X, k)
    assert np.array_equal(I, Io) and np.array_equal(D, Do)


def test_knn_ties(sg, O):
    X = np.repeat(np.arange(40, dtype=np.float32)[:, None] % 4, 3, axis=1)  # many duplicates
    I, D = sg.knn_exact(X, 15)
    Io, Do = O.knn_exact(X, 15)
    assert np.array_equal(I, Io) and np.array_equal(D, Do)


@pytest.mark.parametrize("method", ["EXACT_FP64", "FILTERED"])
@pytest.mark.parametrize("n,dim,k", [(1000, 50, 90), (777, 7, 30), (130, 3, 113), (5000, 50, 90), (3000, 128, 45)])
def test_knn_methods_bitexact(sg, O, method, n, dim, k):
    # both methods give the oracle's rows bit for bit; (130, 3, 113): every
    # other point is a candidate, so the filter's certificate holds trivially
    X = _mixture(n, dim, 5, n + dim)
    I, D, fb = sg.knn(X, k, getattr(sg, "KNN_" + method))
    Io, Do = O.knn_exact(X, k)
    assert np.array_equal(I, Io) and np.array_equal(D, Do)
    assert (fb == -1) == (method == "EXACT_FP64")


def test_knn_filtered_falls_back_on_wide_ties(sg, O):
    # integer lattice: around an interior point the 33rd to 56th neighbours
    # all lie at squared distance 5, a tie wider than the 16 spare candidates,
    # so those rows cannot be certified and are recomputed exactly; every row
    # still equals the oracle's
    g = np.arange(8, dtype=np.float32)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    I, D, fb = sg.knn(X, 33, sg.KNN_FILTERED)
    Io, Do = O.knn_exact(X, 33)
    assert fb > 0
    assert np.array_equal(I, Io) and np.array_equal(D, Do)


def test_knn_filtered_offset_data(sg, O):
    # a large common offset is removed by the centring before FP32 rounding
    X = _mixture(2000, 20, 4, 9) + np.float32(3.0e3)
    I, D, fb = sg.knn(X, 40, sg.KNN_FILTERED)
    Io, Do = O.knn_exact(X, 40)
    assert np.array_equal(I, Io) and np.array_equal(D, Do)


def test_knn_methods_agree_large(sg):
    X = _mixture(20000, 50, 12, 3)
    Ie, De, _ = sg.knn(X, 90, sg.KNN_EXACT_FP64)
    If, Df, fb = sg.knn(X, 90, sg.KNN_FILTERED)
    assert np.array_equal(Ie, If) and np.array_equal(De, Df)
    assert 0 <= fb < 200


@pytest.mark.parametrize("data", ["gaussian", "powerlaw"])
def test_knn_filtered_pruned_clusters(sg, data):
    # clustered inputs at a size where the filter orders the points by 64
    # cells and skips the tiles of far cells: every row still equals the
    # exact FP64 kernel's (itself pinned to the oracle above)
    from sgtsne_amd import synth

    X = synth.gaussian_job(40000, seed=77)[0] if data == "gaussian" else synth.powerlaw_mixture(n=40000)[0]
    Ie, De, _ = sg.knn(X, 90, sg.KNN_EXACT_FP64)
    If, Df, fb = sg.knn(X, 90, sg.KNN_FILTERED)
    assert np.array_equal(Ie, If) and np.array_equal(De, Df)
    assert 0 <= fb < 400


# ---------------------------------------------------------------- P construction
def test_perplexity_golden(sg, mnist):
    f, g = mnist
    P = sg.perplexityEqualization(g["I"], g["D"], 500, 90, 30.0)
    assert np.array_equal(P.rowptr, g["P_rowptr"]) and np.array_equal(P.col, g["P_col"])
    assert np.abs(P.val - g["P_val"]).max() <= 2e-7 * np.abs(g["P_val"]).max()


def test_perplexity_random(sg, O):
    X = _mixture(3000, 20, 8, 1)
    I, D = O.knn_exact(X, 45)
    P = sg.perplexityEqualization(I, D, 3000, 45, 15.0)
    Po = O.perplexity_equalize(I, D, 15.0)
    assert np.array_equal(P.col, Po.col)
    assert np.allclose(P.val, Po.val, rtol=1e-6, atol=1e-12)
    assert np.allclose(np.add.reduceat(P.val.astype(np.float64), P.rowptr[:-1]), 1.0, atol=1e-5)


def test_symmetrize_golden(sg, mnist):
    f, g = mnist
    P = sg.CSR(500, g["P_rowptr"], g["P_col"], g["P_val"].astype(np.float32))
    with sg.Context(sg.tsneparams()) as ctx:
        ctx.upload_P(P, symmetrize=True)
        S = ctx.get_P()
    assert np.array_equal(S.rowptr, g["S_rowptr"]) and np.array_equal(S.col, g["S_col"])
    assert np.allclose(S.val, g["S_val"], rtol=1e-6, atol=0)


def test_lambda_rescale(sg, O):
    X = _mixture(800, 10, 4, 2)
    I, D = O.knn_exact(X, 20)
    Po = O.perplexity_equalize(I, D, 7.0)
    P = sg.CSR(800, Po.rowptr, Po.col, Po.val.astype(np.float32))
    with sg.Context(sg.tsneparams(**{"lambda": 3.0})) as ctx:
        ctx.upload_P(P, symmetrize=True)
        S = ctx.get_P()
    Pl = _oracle_csr(O, P)
    O.lambda_rescale(Pl, 3.0)
    So = O.symmetrize(Pl)
    assert np.array_equal(S.col, So.col)
    assert np.allclose(S.val, So.val, rtol=2e-5, atol=0)


# ---------------------------------------------------------------- FFT convolution
@pytest.mark.parametrize("d,ng", [(1, 16), (1, 117), (1, 512), (2, 16), (2, 52), (2, 66), (2, 130),
                                  (3, 16), (3, 28)])
def test_grid_conv_vs_oracle(sg, O, d, ng):
    rng = np.random.default_rng(ng + d)
    V = rng.uniform(0, 1, size=(d + 1,) + (ng,) * d)
    for h in (0.05, 0.7):
        Phi = sg.grid_conv(V.astype(np.float32), h, zdirect=True)
        if ng ** d <= 4096:
            ref = O.conv_nopad_direct(V, h)
            ref1 = O.conv_nopad_direct(V[:1], h, kpow=1)
        else:
            ref = O.conv_nopad_fft(V, h)
            ref1 = O.conv_nopad_fft(V[:1], h, kpow=1)
        assert rel_l2(Phi[:d + 1], ref) < 1e-5
        assert rel_l2(Phi[d + 1:], ref1) < 1e-5
        Phi2 = sg.grid_conv(V.astype(np.float32), h, zdirect=False)
        assert Phi2.shape[0] == d + 1 and rel_l2(Phi2, ref) < 1e-5


_ROCFFT_SNIPPET = r"""
import sys, numpy as np, torch
f = np.load(sys.argv[1]); V = f["V"]; h = float(f["h"]); d = V.ndim - 1; ng = V.shape[1]
L = 2 * ng
idx = torch.arange(L, device="cuda", dtype=torch.float32)
m = torch.minimum(idx, L - idx)
r2 = sum(g ** 2 for g in torch.meshgrid(*([m] * d), indexing="ij"))
K = (1.0 + h * h * r2) ** -2
Vt = torch.zeros((d + 1,) + (L,) * d, device="cuda")
Vt[(slice(None),) + (slice(0, ng),) * d] = torch.from_numpy(V).cuda()
dims = tuple(range(1, d + 1))
out = torch.fft.ifftn(torch.fft.fftn(Vt, dim=dims) * torch.fft.fftn(K), dim=dims).real
np.save(sys.argv[2], out[(slice(None),) + (slice(0, ng),) * d].cpu().numpy())
"""


@pytest.mark.parametrize("d,ng", [(1, 200), (2, 64), (2, 90), (3, 20)])
def test_grid_conv_vs_rocfft(sg, tmp_path, d, ng):
    # the same zero-padded convolution through torch.fft (rocFFT) in FP32, run
    # in its own process: torch carries its own HIP runtime
    import subprocess
    import sys

    rng = np.random.default_rng(3 * ng + d)
    V = rng.uniform(0, 1, size=(d + 1,) + (ng,) * d).astype(np.float32)
    h = 0.3
    np.savez(tmp_path / "in.npz", V=V, h=h)
    r = subprocess.run([sys.executable, "-c", _ROCFFT_SNIPPET, str(tmp_path / "in.npz"), str(tmp_path / "out.npy")],
                       capture_output=True, text=True, timeout=280)
    if r.returncode != 0 and "No module named" in r.stderr:
        pytest.skip("torch is not importable")
    assert r.returncode == 0, r.stderr[-2000:]
    ref = np.load(tmp_path / "out.npy")
    Phi = sg.grid_conv(V, h)
    assert rel_l2(Phi, ref) < 1e-5


# ---------------------------------------------------------------- per-iteration terms
def _terms_oracle(O, S, Y, h, zmode=None):
    # the GPU's default Z is the direct convolution with 1/(1 + r^2)
    zmode = O.ZETA_DIRECT if zmode is None else zmode
    Yd = Y.astype(np.float64)
    Fa = O.attractive(S, Yd)
    Fr, Z, ng = O.repulsive_interp(Yd.copy(), h, zmode)
    return Fa, Fr, Z, ng


@pytest.mark.parametrize("fmt", ["CSR_SEG", "CSR", "COO", "HYBRID"])
@pytest.mark.parametrize("zeta", ["DIRECT", "ALGEBRAIC"])
def test_terms_golden(sg, mnist, fmt, zeta):
    _check_terms_golden(sg, mnist, fmt, zeta, reorder=2)


@pytest.mark.parametrize("fmt", ["CSR_SEG", "CSR", "COO", "HYBRID"])
@pytest.mark.parametrize("zeta", ["DIRECT", "ALGEBRAIC"])
def test_terms_golden_caller_order(sg, mnist, fmt, zeta):
    # the golden terms with P kept in the caller's row order (no locality order)
    _check_terms_golden(sg, mnist, fmt, zeta, reorder=0)


def _check_terms_golden(sg, mnist, fmt, zeta, reorder):
    f, g = mnist
    S = sg.CSR(500, g["S_rowptr"], g["S_col"], g["S_val"].astype(np.float32))
    sfx = "_direct" if zeta == "DIRECT" else ""
    prm = sg.tsneparams(format=getattr(sg, "FORMAT_" + fmt), zeta_mode=getattr(sg, "ZETA_" + zeta), reorder=reorder)
    with sg.Context(prm) as ctx:
        ctx.upload_P(S, symmetrize=False)
        for tag in ("y0", "ys"):
            ctx.set_Y(g[f"{tag}_Y"])
            Fa, Fr, Z, ng = ctx.terms()
            assert ng == g[f"{tag}_ng"]
            assert rel_l2(Fa, g[f"{tag}_Fattr"]) < TOL
            assert rel_l2(Fr, g[f"{tag}_Frep{sfx}"]) < TOL
            assert abs(Z - g[f"{tag}_Z{sfx}"]) / g[f"{tag}_Z{sfx}"] < TOL


def _clustered_embedding(n, d, rng, spread=30.0, clusters=12):
    # what t-SNE produces: dense unit-scale blobs scattered over a wide extent
    centres = rng.uniform(-spread, spread, size=(clusters, d))
    lab = rng.integers(0, clusters, n)
    return (centres[lab] + rng.standard_normal((n, d)) * rng.uniform(0.3, 2.0, clusters)[lab, None]).astype(np.float32)


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("fmt", ["CSR_SEG", "COO", "HYBRID"])
def test_terms_random(sg, O, d, fmt):
    n = 4000
    X = _mixture(n, 30, 6, 10 + d)
    I, D = O.knn_exact(X, 60)
    So = O.symmetrize(O.perplexity_equalize(I, D, 20.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(d)
    h = {1: 0.5, 2: 0.7, 3: 1.2}[d]
    cases = [(rng.standard_normal((n, d)) * 1e-3).astype(np.float32),   # the start of a run
             (rng.standard_normal((n, d)) * 3.0).astype(np.float32),    # early exaggeration
             _clustered_embedding(n, d, rng)]                           # a formed embedding
    with sg.Context(sg.tsneparams(d=d, format=getattr(sg, "FORMAT_" + fmt))) as ctx:
        ctx.upload_P(S, symmetrize=False)
        for Y in cases:
            ctx.set_Y(Y)
            Yc = ctx.get_Y()  # centred copy: the oracle sees the same Y
            Fa, Fr, Z, ng = ctx.terms()
            Fao, Fro, Zo, ngo = _terms_oracle(O, So, Yc, h)
            assert ng == ngo
            assert rel_l2(Fa, Fao) < TOL
            assert rel_l2(Fr, Fro) < TOL
            assert abs(Z - Zo) / Zo < TOL


@pytest.mark.parametrize("d", [1, 2, 3])
def test_terms_fixed_grid(sg, O, d):
    # the CUDA path's default grid rule (CU/gradient_descend.cu:139-140): the
    # grid side is getBestGridSize(params.ng) + 2 whatever the extent
    n = 4000
    X = _mixture(n, 30, 6, 50 + d)
    I, D = O.knn_exact(X, 60)
    So = O.symmetrize(O.perplexity_equalize(I, D, 20.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(70 + d)
    ngf = {1: 50, 2: 50, 3: 30}[d]
    cases = [(rng.standard_normal((n, d)) * 1e-3).astype(np.float32), _clustered_embedding(n, d, rng)]
    with sg.Context(sg.tsneparams(d=d, grid_mode=sg.GRID_FIXED, ng=ngf)) as ctx:
        ctx.upload_P(S, symmetrize=False)
        for Y in cases:
            ctx.set_Y(Y)
            Yc = ctx.get_Y()
            Fa, Fr, Z, ng = ctx.terms()
            Fro, Zo, ngo = O.repulsive_interp(Yc.astype(np.float64), 0.7, O.ZETA_DIRECT, fixed_ng=ngf)
            assert ng == ngo == O.best_grid_size(ngf) + 2
            assert rel_l2(Fa, O.attractive(So, Yc.astype(np.float64))) < TOL
            assert rel_l2(Fr, Fro) < TOL
            assert abs(Z - Zo) / Zo < TOL


@pytest.mark.parametrize("spread", ["AUTO", "ATOMIC", "MFMA"])
def test_terms_spread_modes_3d(sg, O, spread):
    # 3-D: the matrix-core form on grids of side 16, and the fallback above
    n = 6001
    X = _mixture(n, 20, 5, 79)
    I, D = O.knn_exact(X, 30)
    So = O.symmetrize(O.perplexity_equalize(I, D, 10.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(80)
    cases = [(rng.standard_normal((n, 3)) * 1e-4).astype(np.float32),   # ng 16
             (rng.standard_normal((n, 3)) * 2.0).astype(np.float32),    # ng 16
             _clustered_embedding(n, 3, rng, spread=20.0)]              # ng > 16
    sides = []
    with sg.Context(sg.tsneparams(d=3, spread=getattr(sg, "SPREAD_" + spread))) as ctx:
        ctx.upload_P(S, symmetrize=False)
        for Y in cases:
            ctx.set_Y(Y)
            Yc = ctx.get_Y()
            _, Fr, Z, ng = ctx.terms()
            Fro, Zo, ngo = O.repulsive_interp(Yc.astype(np.float64), 1.2, O.ZETA_DIRECT)
            assert ng == ngo
            assert rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL, (spread, ng)
            sides.append(ng)
    assert sides[0] == 16 and sides[1] == 16 and sides[2] > 16, sides


@pytest.mark.parametrize("spread", ["AUTO", "ATOMIC", "MFMA"])
def test_terms_spread_modes(sg, O, spread):
    # every spread path against the oracle on 2-D grids of side 16, 20-32
    # (the matrix-core form: one and two 16-node tiles per axis), larger
    # (where the matrix-core request falls back to an LDS copy of the grid)
    # and beyond an LDS copy (fixed-point windows, or atomics when asked);
    # the start of a run crowds every point into a few cells, the worst case
    # for atomics
    n = 6001
    X = _mixture(n, 20, 5, 77)
    I, D = O.knn_exact(X, 30)
    So = O.symmetrize(O.perplexity_equalize(I, D, 10.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(78)
    cases = [(rng.standard_normal((n, 2)) * 1e-4).astype(np.float32),   # ng 16
             (rng.standard_normal((n, 2)) * 2.2).astype(np.float32),    # ng 20-32
             _clustered_embedding(n, 2, rng, spread=4.5),               # ng 20-32
             _clustered_embedding(n, 2, rng),                           # ng > 32
             _clustered_embedding(n, 2, rng, spread=80.0)]              # too large for an LDS copy
    sides = []
    with sg.Context(sg.tsneparams(spread=getattr(sg, "SPREAD_" + spread))) as ctx:
        ctx.upload_P(S, symmetrize=False)
        for Y in cases:
            ctx.set_Y(Y)
            Yc = ctx.get_Y()
            _, Fr, Z, ng = ctx.terms()
            Fro, Zo, ngo = O.repulsive_interp(Yc.astype(np.float64), 0.7, O.ZETA_DIRECT)
            assert ng == ngo
            assert rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL, (spread, ng)
            sides.append(ng)
    assert sides[0] == 16 and 16 < sides[1] <= 32 and 16 < sides[2] <= 32 and sides[3] > 32 and sides[4] > 74, sides


@pytest.mark.parametrize("d", [2, 3])
def test_large_grid_spread_is_deterministic(sg, O, d):
    # grids no LDS copy holds go through fixed-point windows: the terms are
    # the oracle's and the same bits on every evaluation, whatever order the
    # work-groups' atomics land in
    n = 8001
    X = _mixture(n, 20, 5, 81 + d)
    I, D = O.knn_exact(X, 30)
    So = O.symmetrize(O.perplexity_equalize(I, D, 10.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(90 + d)
    Y = _clustered_embedding(n, d, rng, spread={2: 80.0, 3: 25.0}[d])
    h = {2: 0.7, 3: 1.2}[d]
    with sg.Context(sg.tsneparams(d=d)) as ctx:
        ctx.upload_P(S, symmetrize=False)
        ctx.set_Y(Y)
        runs = [ctx.terms() for _ in range(3)]
        Yc = ctx.get_Y()
    _, Fr, Z, ng = runs[0]
    assert ng > (74 if d == 2 else 16), ng
    for _, Fr2, Z2, _ in runs[1:]:
        assert np.array_equal(Fr, Fr2) and Z == Z2
    Fro, Zo, ngo = O.repulsive_interp(Yc.astype(np.float64), h, O.ZETA_DIRECT)
    assert ng == ngo and rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL


@pytest.mark.parametrize("d", [1, 2, 3])
def test_terms_reference_acceptance_setup(sg, O, d):
    # the reference's repulsion self-test inputs (HY/test_modules.cpp:122-156):
    # Y ~ U[0,1]^d, h in {0.05, 0.08, 0.13}; FP32 GPU against the FP64 oracle
    n = 3000
    rng = np.random.default_rng(40 + d)
    Y = rng.uniform(0, 1, size=(n, d)).astype(np.float32)
    S = sg.CSR(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    for h in (0.05, 0.08, 0.13):
        for zeta, zmode in (("DIRECT", O.ZETA_DIRECT), ("ALGEBRAIC", O.ZETA_ALGEBRAIC)):
            with sg.Context(sg.tsneparams(d=d, h=h, zeta_mode=getattr(sg, "ZETA_" + zeta))) as ctx:
                ctx.upload_P(S, symmetrize=False)
                ctx.set_Y(Y)
                Yc = ctx.get_Y().astype(np.float64)
                _, Fr, Z, ng = ctx.terms()
                Fro, Zo, ngo = O.repulsive_interp(Yc.copy(), h, zmode)
                Fe, Ze = O.repulsive_exact(Yc)
                assert ng == ngo
                assert rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL, zeta
                # and the reference's own acceptance bar against the exact sum
                assert np.abs(Fr - Fe).max() <= 1e-7 and abs(Z - Ze) / Ze <= 1e-4, zeta


def test_terms_ragged_rows(sg, O):
    # rows from 0 to 300 entries, including empty rows and a hub
    rng = np.random.default_rng(9)
    n = 2000
    lens = rng.integers(0, 40, n)
    lens[:5] = 0
    lens[7] = 300
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(5, n, rows.size)  # rows 0..4 stay empty after P + P^T
    A = {}
    for r, c in zip(rows, cols):
        if r != c:
            A[(r, c)] = rng.uniform(0.1, 1)
            A[(c, r)] = A[(r, c)]
    keys = sorted(A)
    rp = np.zeros(n + 1, np.int64)
    for r, _ in keys:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    col = np.array([c for _, c in keys], np.int32)
    val = np.array([A[k] for k in keys])
    val /= val.sum()
    So = O.CSR(n, rp, col, val)
    for fmt in ("CSR_SEG", "CSR", "COO", "HYBRID"):
        with sg.Context(sg.tsneparams(format=getattr(sg, "FORMAT_" + fmt))) as ctx:
            ctx.upload_P(sg.CSR(n, rp, col, val.astype(np.float32)), symmetrize=False)
            Y = (rng.standard_normal((n, 2)) * 5).astype(np.float32)
            ctx.set_Y(Y)
            Yc = ctx.get_Y()
            Fa, Fr, Z, ng = ctx.terms()
            Fao, Fro, Zo, _ = _terms_oracle(O, So, Yc, 0.7)
            assert rel_l2(Fa, Fao) < TOL, fmt
            assert np.all(Fa[:5] == 0)


def test_exact_repulsion(sg, O):
    n = 3000
    rng = np.random.default_rng(11)
    Y = (rng.standard_normal((n, 2)) * 10).astype(np.float32)
    S = sg.CSR(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    with sg.Context(sg.tsneparams()) as ctx:
        ctx.upload_P(S, symmetrize=False)
        ctx.set_Y(Y)
        Yc = ctx.get_Y()
        Fe, Ze = ctx.repulsive_exact()
    Fo, Zo = O.repulsive_exact(Yc.astype(np.float64))
    assert rel_l2(Fe, Fo) < 1e-5 and abs(Ze - Zo) / Zo < 1e-6


def test_kl_golden(sg, mnist):
    # the golden KL uses the reference's algebraic Z
    f, g = mnist
    S = sg.CSR(500, g["S_rowptr"], g["S_col"], g["S_val"].astype(np.float32))
    with sg.Context(sg.tsneparams(zeta_mode=sg.ZETA_ALGEBRAIC)) as ctx:
        ctx.upload_P(S, symmetrize=False)
        ctx.set_Y(g["ys_Y"])
        kl = ctx.kl(1.0)
    assert abs(kl - g["ys_kl"]) / abs(g["ys_kl"]) < TOL


# ---------------------------------------------------------------- iterations
def test_iterations_golden(sg, O, mnist):
    # from the reference's own start (initial2.mat): 10 iterations track the
    # FP64 loop closely; after 100 (trajectories of an exaggerated descent
    # diverge) the KL agrees within 1%
    f, g = mnist
    S = sg.CSR(500, g["S_rowptr"], g["S_col"], g["S_val"].astype(np.float32))
    So = O.CSR(500, g["S_rowptr"].astype(np.int64), g["S_col"], g["S_val"])
    with sg.Context(sg.tsneparams()) as ctx:
        ctx.upload_P(S, symmetrize=False)
        ctx.set_Y(f["Y0"])
        Y0c = ctx.get_Y().astype(np.float64)
        ctx.step(10)
        Y10 = ctx.get_Y()
        ctx.step(90)
        kl = ctx.kl(1.0)
    Yo = Y0c.copy()
    uY, gn = np.zeros_like(Yo), np.ones_like(Yo)
    O.run(So, Yo, uY, gn, 0, 10, zmode=O.ZETA_DIRECT)
    assert rel_l2(Y10, Yo) < 1e-3
    Yo = g["run100_Y"]
    _, Zo, _ = O.repulsive_interp(Yo.copy(), 0.7, O.ZETA_DIRECT)
    klo = O.kl(So, Yo, Zo, 1.0)
    assert abs(kl - klo) / klo < 0.01


def test_full_run_final_kl(sg, O):
    # the drop-in entry, 1000 iterations, final KL within 1% of the FP64 loop
    n = 2500
    X = _mixture(n, 50, 10, 12)
    I, D = sg.knn_exact(X, 90)
    P = sg.perplexityEqualization(I, D, n, 90, 30.0)
    y0 = (1e-4 * np.random.default_rng(1).standard_normal((n, 2))).astype(np.float32)
    prm = sg.tsneparams(layout=sg.LAYOUT_AOS)
    ti = np.zeros(7 * prm.maxIter + 1)
    Y = sg.sgtsne(P, prm, y_in=y0, time_info=ti)
    assert np.all(np.isfinite(Y)) and ti[7 * prm.maxIter] > 0 and ti[0] > 0
    So = O.symmetrize(_oracle_csr(O, P))
    Yo = y0.astype(np.float64).copy()
    O.run(So, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, 1000, zmode=O.ZETA_DIRECT)
    _, Zg, _ = O.repulsive_interp(Y.astype(np.float64).copy(), 0.7, O.ZETA_DIRECT)
    _, Zo, _ = O.repulsive_interp(Yo.copy(), 0.7, O.ZETA_DIRECT)
    klg = O.kl(So, Y.astype(np.float64), Zg)
    klo = O.kl(So, Yo, Zo)
    assert abs(klg - klo) / klo < 0.01, (klg, klo)


def test_soa_layout_and_errors(sg):
    n = 600
    X = _mixture(n, 10, 3, 13)
    I, D = sg.knn_exact(X, 30)
    P = sg.perplexityEqualization(I, D, n, 30, 10.0)
    prm = sg.tsneparams(maxIter=20, layout=sg.LAYOUT_SOA, d=3)
    Y = sg.sgtsne(P, prm)
    assert Y.shape == (3, n) and np.all(np.isfinite(Y))
    # a non-finite embedding is reported, not propagated
    with sg.Context(sg.tsneparams()) as ctx:
        ctx.upload_P(P, symmetrize=True)
        Yb = np.zeros((n, 2), np.float32)
        Yb[5, 0] = np.nan
        ctx.set_Y(Yb)
        with pytest.raises(sg.SgtsneError, match="-4"):
            ctx.step(1)


def test_reorder_modes_agree(sg, O):
    # the device order (as given / breadth-first / re-sorted along Y at
    # iteration 50) changes only the summation order.  After the re-order the
    # caller-order bookkeeping is exact: the terms at the returned Y match the
    # oracle's; and the final quality (KL) agrees across the three modes.
    n = 3000
    X = _mixture(n, 20, 6, 14)
    I, D = O.knn_exact(X, 45)
    P = sg.perplexityEqualization(I, D, n, 45, 15.0)
    So = O.symmetrize(_oracle_csr(O, P))
    y0 = (1e-4 * np.random.default_rng(2).standard_normal((n, 2))).astype(np.float32)
    kls = {}
    for mode in (0, 1, 2):
        with sg.Context(sg.tsneparams(reorder=mode)) as ctx:
            ctx.upload_P(P, symmetrize=True)
            ctx.set_Y(y0)
            ctx.step(60)
            Y = ctx.get_Y()
            Fa, Fr, Z, ng = ctx.terms()
            Fao, Fro, Zo, ngo = _terms_oracle(O, So, Y, 0.7)
            assert ng == ngo
            assert rel_l2(Fa, Fao) < TOL and rel_l2(Fr, Fro) < TOL and abs(Z - Zo) / Zo < TOL, mode
            ctx.step(240)
            kls[mode] = ctx.kl(1.0)
    for mode in (1, 2):
        assert abs(kls[mode] - kls[0]) / kls[0] < 0.02, kls


# every shape of the default CSR kernel: same result as the oracle, and the
# same bits for every shape with the same lanes per row
# (lanes per row, unroll, rows per work-group, points of Y in LDS); unroll 8
# and 16 are the masked-step forms, whose last step of a row is partly masked
_SHAPES = [(16, 0, 0, 0), (32, 0, 0, 0), (64, 0, 0, 0), (16, 4, 0, 0), (16, 2, 0, 0), (16, 1, 64, 0),
           (16, 2, 128, 0), (16, 1, 256, 4096), (16, 2, 512, 4096), (16, 2, 1024, 8192), (32, 2, 256, 0),
           (32, 1, 512, 4096), (16, 0, 0, -1), (32, 0, 0, -1), (64, 0, 0, -1),
           (16, 8, 0, 0), (32, 8, 0, 0), (64, 8, 0, 0), (16, 16, 0, 0), (32, 16, 0, 0),
           (32, 2, 0, 0), (32, 4, 0, 0), (64, 2, 0, 0), (32, 1, 64, 0), (32, 2, 64, 0), (32, 2, 1024, 8192), (64, 1, 256, 0),
           (16, 1, 512, 8192), (32, 16, 64, 0), (32, 16, 256, 0), (16, 16, 128, 0), (32, 8, 256, 0),
           # chunk-cached: 127 chunks of 64 points (8128 < n: the rest from Y) or,
           # in 3-D, 8192 cached points; ranges not dividing n
           (32, 16, 2048, -512), (16, 16, 512, -64), (64, 8, 1024, -128), (32, 8, 4096, -1024),
           (32, 16, 1000, -64)]


@pytest.mark.parametrize("d", [1, 2, 3])
def test_attractive_kernel_shapes(sg, O, d):
    # n is not a multiple of any work-group's rows, and rows have 30-90
    # entries: short rows end inside a masked step, long ones take several
    n = 12007
    X = _mixture(n, 20, 8, 90 + d)
    I, D = O.knn_exact(X, 30)
    So = O.symmetrize(O.perplexity_equalize(I, D, 10.0))
    S = sg.CSR(n, So.rowptr, So.col, So.val.astype(np.float32))
    rng = np.random.default_rng(d)
    Y = _clustered_embedding(n, d, rng)
    got = {}
    with sg.Context(sg.tsneparams(d=d, reorder=1)) as ctx:
        ctx.upload_P(S, symmetrize=False)
        ctx.set_Y(Y)
        Yc = ctx.get_Y()
        Fao = O.attractive(So, Yc.astype(np.float64))
        for shp in _SHAPES:
            g, u, rpb, win = shp
            if win * (4 if d == 3 else d) * 4 > 65536:
                with pytest.raises(sg.SgtsneError):
                    ctx.set_attractive(*shp)
                continue
            ctx.set_attractive(*shp)
            assert ctx.tuning()["window"] == win
            Fa, _, _, _ = ctx.terms()
            assert rel_l2(Fa, Fao) < TOL, shp
            got[shp] = Fa
    for shp, Fa in got.items():
        ref = got[(shp[0], 0, 0, 0)]
        assert np.array_equal(Fa, ref), shp
