"""Pins of the FP64 oracle (CPU).  Each check is independent of the oracle's
own formula: closed forms, brute force, textbook routines (numpy.fft, dense
linear algebra), invariants, and the reference's own acceptance test
(HY/test_modules.cpp:122-156)."""
import numpy as np
import pytest

from conftest import rel_l2

GRID_LIST = [8, 9, 10, 11, 12, 13, 14, 15, 16, 20, 25, 26, 28, 32, 33, 35, 36, 39, 40, 42, 44, 45, 48,
             49, 50, 52, 54, 55, 56, 60, 63, 64, 65, 66, 70, 72, 75, 77, 78, 80, 84, 88, 90, 91, 96, 98, 99,
             100, 104, 105, 108, 110, 112, 117, 120, 125, 126, 130, 132, 135, 140, 144, 147, 150, 154, 156,
             160, 165, 168, 175, 176, 180, 182, 189, 192, 195, 196, 198, 200, 208, 210, 216, 220, 224, 225,
             231, 234, 240, 245, 250, 252, 260, 264, 270, 273, 275, 280, 288, 294, 297, 300, 308, 312, 315,
             320, 325, 330, 336, 343, 350, 351, 352, 360, 364, 375, 378, 385, 390, 392, 396, 400, 416, 420,
             432, 440, 441, 448, 450, 455, 462, 468, 480, 490, 495, 500, 504, 512]


def test_grid_size_table(O):
    # HY/qq.cpp:108-130: smallest listed size >= nGrid + 2, minus 2; capped at 510
    for g in range(1, 700):
        want = next((s - 2 for s in GRID_LIST if g + 2 <= s), 510)
        assert O.best_grid_size(g) == want
    assert O.best_grid_size(14) == 14 and O.best_grid_size(50) == 50 and O.best_grid_size(51) == 52


@pytest.mark.parametrize("n", [8, 12, 13, 25, 52, 66, 77, 91, 117, 143, 512, 1024])
def test_fft_matches_numpy(O, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.fft(x, -1) - np.fft.fft(x)).max() < 1e-12 * n
    assert np.abs(O.fft(x, +1) - np.fft.ifft(x) * n).max() < 1e-12 * n


@pytest.mark.parametrize("d,ng", [(1, 30), (1, 52), (2, 16), (2, 20), (3, 10)])
def test_conv_even_odd_equals_toeplitz(O, d, ng):
    # the 2^d-term circulant construction equals the direct non-periodic sum
    rng = np.random.default_rng(d * 100 + ng)
    V = rng.standard_normal((d + 1,) + (ng,) * d)
    for h in (0.05, 0.4, 2.0):
        a = O.conv_nopad_fft(V, h)
        b = O.conv_nopad_direct(V, h)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("d,ng", [(1, 40), (2, 18), (3, 9)])
def test_conv_cauchy_kernel_even_odd_equals_toeplitz(O, d, ng):
    # the same construction with the kernel 1/(1 + h^2 r^2) (direct Z)
    rng = np.random.default_rng(7 * d + ng)
    V = rng.standard_normal((2,) + (ng,) * d)
    a = O.conv_nopad_fft(V, 0.3, kpow=1)
    b = O.conv_nopad_direct(V, 0.3, kpow=1)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_conv_direct_is_toeplitz_product(O):
    # 1-D: the direct sum is the dense Toeplitz matrix times V
    rng = np.random.default_rng(3)
    ng, h = 23, 0.3
    V = rng.standard_normal((2, ng))
    i = np.arange(ng)
    T = (1.0 + h * h * (i[:, None] - i[None, :]) ** 2) ** -2
    assert np.allclose(O.conv_nopad_direct(V, h), V @ T.T, rtol=1e-13, atol=0)
    T1 = 1.0 / (1.0 + h * h * (i[:, None] - i[None, :]) ** 2)
    assert np.allclose(O.conv_nopad_direct(V, h, kpow=1), V @ T1.T, rtol=1e-13, atol=0)


@pytest.mark.parametrize("delta", [0.0, 0.1, 0.37, 0.5, 0.81, 0.999])
def test_lagrange_weights_reproduce_cubics(O, delta):
    # nodes f..f+3 sit at offsets -(1+delta), -delta, 1-delta, 2-delta from the
    # point: cubic Lagrange interpolation reproduces 1, t, t^2, t^3 exactly
    w = O.lagrange_weights(delta)
    t = np.array([-(1 + delta), -delta, 1 - delta, 2 - delta])
    assert abs(w.sum() - 1) < 1e-14
    for p in (1, 2, 3):
        assert abs((w * t ** p).sum()) < 1e-13
    if delta == 0.0:
        assert np.allclose(w, [0, 1, 0, 0], atol=1e-15)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_s2g_g2s_adjoint(O, d):
    rng = np.random.default_rng(d)
    n, ng, m = 300, 12, d + 1
    y = rng.uniform(0, ng - 3 - 1e-9, size=(n, d))
    q = rng.standard_normal((n, m))
    V = O.s2g(y, q, ng)
    Phi = rng.standard_normal((m,) + (ng,) * d)
    lhs = (V * Phi).sum()
    rhs = (q * O.g2s(y, Phi)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_s2g_brute_force_tiny(O, d):
    # one point: the grid holds the outer product of the 1-D weights
    y = np.array([[1.3, 2.6, 0.2][:d]])
    q = np.array([[1.0] + [2.0] * d])
    ng = 7
    V = O.s2g(y, q, ng)
    ref = np.zeros((d + 1,) + (ng,) * d)
    ws = [O.lagrange_weights(v - np.floor(v)) for v in y[0]]
    fs = [int(np.floor(v)) for v in y[0]]
    for idx in np.ndindex(*(4,) * d):
        w = np.prod([ws[a][idx[a]] for a in range(d)])
        node = tuple(fs[a] + idx[a] for a in range(d))[::-1]  # x fastest
        for c in range(d + 1):
            ref[(c,) + node] += q[0, c] * w
    assert np.allclose(V, ref, atol=1e-15)


def _exact_np(Y):
    diff = Y[:, None, :] - Y[None, :, :]
    w = 1.0 / (1.0 + (diff ** 2).sum(-1))
    np.fill_diagonal(w, 0.0)
    Z = w.sum()
    F = (diff * (w ** 2)[:, :, None]).sum(1) / Z
    return F, Z


@pytest.mark.parametrize("d", [1, 2, 3])
def test_exact_repulsion_brute_force(O, d):
    rng = np.random.default_rng(10 + d)
    Y = rng.standard_normal((400, d)) * 3
    F, Z = O.repulsive_exact(Y)
    Fn, Zn = _exact_np(Y)
    assert abs(Z - Zn) / Zn < 1e-12 and rel_l2(F, Fn) < 1e-12


@pytest.mark.parametrize("d", [1, 2, 3])
def test_z_identity(O, d):
    # sum_i [(1 + 2|y_i|^2) Phi0_i - 2 y_i . Phi_i] - N = sum_{i != j} 1/(1 + d_ij^2)
    rng = np.random.default_rng(20 + d)
    Y = rng.standard_normal((300, d)) * 2 + 5
    Phi = O.phi_exact(Y)
    z = (1 + 2 * (Y ** 2).sum(1)) * Phi[:, 0] - 2 * (Y * Phi[:, 1:]).sum(1)
    _, Zn = _exact_np(Y)
    assert abs((z.sum() - len(Y)) - Zn) / Zn < 1e-10


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("n", [1000, 2000])
def test_interp_repulsion_reference_acceptance(O, d, n):
    # HY/test_modules.cpp:122-156: Y ~ U[0,1]^d, h in {0.05, 0.08, 0.13},
    # max |dFrep| <= 1e-7 and |dZ|/Z <= 1e-4 against the exact sum
    rng = np.random.default_rng(n + d)
    Y = rng.uniform(0, 1, size=(n, d))
    Fe, Ze = O.repulsive_exact(Y)
    for h in (0.05, 0.08, 0.13):
        for zmode in (O.ZETA_ALGEBRAIC, O.ZETA_DIRECT):
            Fi, Zi, ng = O.repulsive_interp(Y.copy(), h, zmode)
            assert np.abs(Fi - Fe).max() <= 1e-7
            assert abs(Zi - Ze) / Ze <= 1e-4


def test_direct_z_matches_exact_sum_clustered(O):
    # the direct Z approximates sum_{i != j} 1/(1 + d^2) on a clustered
    # embedding as well as the algebraic one does
    rng = np.random.default_rng(91)
    c = rng.uniform(-20, 20, size=(8, 2))
    Y = c[rng.integers(0, 8, 1500)] + rng.standard_normal((1500, 2))
    _, Ze = O.repulsive_exact(Y)
    _, Za, _ = O.repulsive_interp(Y.copy(), 0.3, O.ZETA_ALGEBRAIC)
    _, Zd, _ = O.repulsive_interp(Y.copy(), 0.3, O.ZETA_DIRECT)
    assert abs(Zd - Ze) / Ze < 1e-3 and abs(Za - Ze) / Ze < 1e-3


def test_interp_repulsion_grid_and_shift(O):
    # the grid follows ceil(maxy/h) through the table and the input is shifted
    # to min 0 in place (HY/qq.cpp:147-166)
    rng = np.random.default_rng(5)
    Y = rng.uniform(-3, 7, size=(500, 2))
    Yc = Y.copy()
    _, _, ng = O.repulsive_interp(Yc, 0.7)
    maxy = (Y - Y.min(0)).max()
    assert ng == O.best_grid_size(max(int(np.ceil(maxy / 0.7)), 14)) + 2
    assert np.allclose(Yc, Y - Y.min(0))


@pytest.mark.parametrize("d", [1, 2, 3])
def test_attractive_dense(O, d):
    rng = np.random.default_rng(30 + d)
    n = 200
    A = (rng.uniform(size=(n, n)) < 0.05) * rng.uniform(size=(n, n))
    rp = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int64)
    col = np.nonzero(A)[1].astype(np.int32)
    P = O.CSR(n, rp, col, A[A != 0])
    Y = rng.standard_normal((n, d))
    diff = Y[:, None, :] - Y[None, :, :]
    q = 1.0 / (1.0 + (diff ** 2).sum(-1))
    F = ((A * q)[:, :, None] * diff).sum(1)
    assert rel_l2(O.attractive(P, Y), F) < 1e-13


def test_knn_brute_force(O):
    rng = np.random.default_rng(40)
    X = rng.standard_normal((400, 17)).astype(np.float32)
    k = 12
    I, D = O.knn_exact(X, k)
    Xd = X.astype(np.float64)
    Dn = ((Xd[:, None, :] - Xd[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(Dn, np.inf)
    order = np.argsort(Dn, axis=1, kind="stable")[:, :k]
    assert np.array_equal(I[:, 0], np.arange(400)) and np.all(D[:, 0] == 0)
    assert np.array_equal(I[:, 1:], order)
    assert np.allclose(D[:, 1:], np.take_along_axis(Dn, order, 1), rtol=1e-13)


def test_knn_rows_equal_full_rows(O):
    rng = np.random.default_rng(31)
    X = rng.standard_normal((700, 12)).astype(np.float32)
    I, D = O.knn_exact(X, 25)
    rows = rng.choice(700, 40, replace=False)
    Ir, Dr = O.knn_rows(X, 25, rows)
    assert np.array_equal(Ir, I[rows]) and np.array_equal(Dr, D[rows])


def test_knn_ties_lower_index_first(O):
    # duplicated points: equal distances are ordered by index
    X = np.zeros((6, 3), np.float32)
    X[3:] = 1.0
    I, D = O.knn_exact(X, 4)
    assert list(I[0]) == [0, 1, 2, 3, 4]
    assert list(I[4]) == [4, 3, 5, 0, 1]


def _entropy(p):
    p = p[p > 0]
    return -(p * np.log(p)).sum()


def test_perplexity_rows(O):
    rng = np.random.default_rng(50)
    X = rng.standard_normal((300, 10)).astype(np.float32)
    k, u = 30, 10.0
    I, D = O.knn_exact(X, k)
    P = O.perplexity_equalize(I, D, u)
    V = P.val.reshape(300, k)
    assert np.allclose(V.sum(1), 1.0, atol=1e-12)
    H = np.array([_entropy(r) for r in V])
    assert np.abs(H - np.log(u)).max() < 1e-5
    # the bisection solves H(beta) = log u; compare with a root bracket
    from scipy.optimize import brentq
    for i in range(0, 300, 37):
        d = D[i, 1:]

        def Hb(b):
            p = np.exp(-b * (d - d.min()))
            p /= p.sum()
            return _entropy(p) - np.log(u)

        beta = brentq(Hb, 1e-8, 1e4, xtol=1e-14)
        p = np.exp(-beta * (d - d.min()))
        p /= p.sum()
        assert np.abs(p - V[i]).max() < 1e-4
    assert np.array_equal(P.col.reshape(300, k), I[:, 1:])


def test_perplexity_equal_distances(O):
    # all distances equal: the row is uniform for every beta and its entropy is
    # log k, so u = k is met at once and p = 1/k
    d = np.concatenate([[0.0], np.full(20, 3.7)])
    val, beta, it = O.equalize_vertex(d, 20.0)
    assert it == 1 and beta == 1.0
    assert np.allclose(val, 1 / 20, rtol=1e-14)
    # u < k cannot be met by a uniform row: whatever beta the search ends on,
    # the row stays uniform
    val, beta, it = O.equalize_vertex(d, 5.0)
    assert np.ptp(val) == 0.0


def test_symmetrize(O):
    rng = np.random.default_rng(60)
    X = rng.standard_normal((250, 8)).astype(np.float32)
    I, D = O.knn_exact(X, 15)
    P = O.perplexity_equalize(I, D, 5.0)
    S = O.symmetrize(P)
    A = P.dense()
    B = (A + A.T) / (A + A.T).sum()
    assert np.allclose(S.dense(), B, rtol=1e-13, atol=0)
    assert S.nnz == np.count_nonzero(A + A.T)
    for i in range(250):
        c = S.col[S.rowptr[i]:S.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)


def test_make_stochastic(O):
    rng = np.random.default_rng(61)
    n = 50
    rp = np.arange(0, 5 * n + 1, 5, dtype=np.int64)
    P = O.CSR(n, rp, rng.integers(0, n, 5 * n), rng.uniform(0.1, 1, 5 * n))
    O.make_stochastic(P)
    assert np.allclose(np.add.reduceat(P.val, rp[:-1]), 1.0)


def test_lambda_rescale_solves_equation(O):
    # rows after rescaling: sum_j P_ij^sigma = lambda for the bisection's sigma,
    # then normalised to 1 (HY/graph_rescaling.cpp:19-104)
    rng = np.random.default_rng(62)
    n, k, lam = 40, 12, 3.0
    rp = np.arange(0, k * n + 1, k, dtype=np.int64)
    v = rng.uniform(0.01, 1, size=k * n)
    v /= np.repeat(np.add.reduceat(v, rp[:-1]), k)
    P = O.CSR(n, rp, rng.integers(0, n, k * n), v.copy())
    O.lambda_rescale(P, lam)
    from scipy.optimize import brentq
    for i in range(n):
        p0 = v[i * k:(i + 1) * k]
        sig = brentq(lambda s: (p0 ** s).sum() - lam, 1e-6, 1e3, xtol=1e-15)
        want = p0 ** sig / (p0 ** sig).sum()
        assert np.abs(P.val[i * k:(i + 1) * k] - want).max() < 1e-4


def test_update_hand_cases(O):
    # dy = alpha Fa - Fr; gains +0.2 when sign(dy) != sign(uy) else *0.8 (floor
    # 0.01); uy = mom uy - eta g dy; y += uy; then Y is mean-centred
    Fa = np.array([[1.0], [0.0], [0.0]])
    Fr = np.array([[0.0], [0.5], [0.0]])
    Y = np.array([[0.0], [1.0], [2.0]])
    uY = np.array([[0.0], [-0.1], [0.3]])
    g = np.array([[1.0], [1.0], [0.011]])
    O.update(Fa, Fr, 2.0, 0.5, 10.0, Y, uY, g)
    # point 0: dy=2, sign(dy)=1 != sign(0)=0 -> g=1.2, uy=-24, y=-24
    # point 1: dy=-0.5, sign -1 == sign(-0.1) -> g=0.8, uy=-0.05+4=3.95, y=4.95
    # point 2: dy=0, sign 0 != 1 -> g=0.211, uy=0.15, y=2.15
    assert np.allclose(g[:, 0], [1.2, 0.8, 0.211])
    assert np.allclose(uY[:, 0], [-24.0, 3.95, 0.15])
    raw = np.array([-24.0, 4.95, 2.15])
    assert np.allclose(Y[:, 0], raw - raw.mean())
    g2 = np.array([[0.0105]])
    O.update(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, 0.5, 1.0, np.zeros((1, 1)), np.zeros((1, 1)), g2)
    assert g2[0, 0] == 0.01


def test_kl_dense(O):
    rng = np.random.default_rng(70)
    n = 150
    A = (rng.uniform(size=(n, n)) < 0.08) * rng.uniform(size=(n, n))
    A = A + A.T
    A /= A.sum()
    rp = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int64)
    P = O.CSR(n, rp, np.nonzero(A)[1].astype(np.int32), A[A != 0])
    Y = rng.standard_normal((n, 2))
    _, Z = _exact_np(Y)
    diff = Y[:, None, :] - Y[None, :, :]
    Q = 1.0 / (1.0 + (diff ** 2).sum(-1)) / Z
    m = A > 0
    want = (A[m] * np.log(A[m] / Q[m])).sum()
    assert abs(O.kl(P, Y, Z, 1.0) - want) < 1e-9 * abs(want)


def test_run_decreases_kl(O):
    rng = np.random.default_rng(80)
    centres = rng.normal(0, 6, size=(3, 10))
    X = (centres[rng.integers(0, 3, 400)] + rng.standard_normal((400, 10))).astype(np.float32)
    I, D = O.knn_exact(X, 30)
    S = O.symmetrize(O.perplexity_equalize(I, D, 10.0))
    Y = 1e-4 * rng.standard_normal((400, 2))
    uY, g = np.zeros_like(Y), np.ones_like(Y)
    Zs, ngs = O.run(S, Y, uY, g, 0, 300)
    _, Z = _exact_np(Y)
    kl_end = O.kl(S, Y, Z)
    Y0 = 1e-4 * rng.standard_normal((400, 2))
    _, Z0 = _exact_np(Y0)
    assert kl_end < 0.5 * O.kl(S, Y0, Z0)
    assert np.allclose(Y.mean(0), 0, atol=1e-9)


def test_golden_fixture_regression(O):
    # the committed oracle vectors still reproduce (guards accidental edits)
    import os
    here = os.path.join(os.path.dirname(__file__), "golden")
    f = np.load(os.path.join(here, "mnist500.npz"))
    g = np.load(os.path.join(here, "mnist500_oracle.npz"))
    I, D = O.knn_exact(f["X"], 90)
    assert np.array_equal(I, g["I"]) and np.array_equal(D, g["D"])
    P = O.perplexity_equalize(I, D, 30.0)
    S = O.symmetrize(P)
    assert np.array_equal(S.col, g["S_col"]) and np.allclose(S.val, g["S_val"], rtol=1e-12)
    Y = g["ys_Y"].astype(np.float64)
    Fr, Z, ng = O.repulsive_interp(Y.copy(), 0.7)
    assert ng == g["ys_ng"] and abs(Z - g["ys_Z"]) <= 1e-9 * Z


def _schedule_case(O, n=300, seed=81):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0, 6, size=(3, 10))
    X = (centres[rng.integers(0, 3, n)] + rng.standard_normal((n, 10))).astype(np.float32)
    I, D = O.knn_exact(X, 20)
    return O.symmetrize(O.perplexity_equalize(I, D, 7.0)), rng


def _hand_steps(O, S, Y, uY, g, it0, it1, alpha0, early, h=0.7, alpha_at=None, mom_at=250):
    # the reference's control flow (HY/gradient_descendHYB.cpp:203-222,
    # CU/gradient_descend.cu:215-236): gradient and update with the current
    # alpha and momentum, THEN "if (iter == stop_lying_iter) alpha = 1" and
    # "if (iter == mom_switch_iter) momentum = final_momentum"
    alpha_at = early if alpha_at is None else alpha_at
    alpha, mom = alpha0, 0.5
    for it in range(it1):
        if it >= it0:
            Fr, _, _ = O.repulsive_interp(Y, h, O.ZETA_ALGEBRAIC)
            Fa = O.attractive(S, Y)
            O.update(Fa, Fr, alpha, mom, 200.0, Y, uY, g)
        if it == alpha_at:
            alpha = 1.0
        if it == mom_at:
            mom = 0.8


def test_run_schedule_alpha_switch(O):
    # or_run's exaggeration switch against the hand-stepped reference loop;
    # moving the switch by one iteration either way is detected
    S, rng = _schedule_case(O)
    y0 = rng.standard_normal((S.n, 2))
    runs = {}
    for tag, at in (("ref", 3), ("early", 2), ("late", 4)):
        Y, uY, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
        _hand_steps(O, S, Y, uY, g, 0, 7, 12.0, 3, alpha_at=at)
        runs[tag] = Y
    Y, uY, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    O.run(S, Y, uY, g, 0, 7, alpha0=12.0, early_iter=3)
    assert rel_l2(Y, runs["ref"]) < 1e-13
    assert rel_l2(runs["early"], runs["ref"]) > 1e-6
    assert rel_l2(runs["late"], runs["ref"]) > 1e-6


def test_run_schedule_momentum_switch(O):
    # the momentum switch at iteration 250, over the window 247..253, from a
    # state with nonzero velocity; a switch one iteration off is detected
    S, rng = _schedule_case(O, seed=82)
    Y = 1e-2 * rng.standard_normal((S.n, 2))
    uY, g = np.zeros_like(Y), np.ones_like(Y)
    O.run(S, Y, uY, g, 0, 247, alpha0=4.0, early_iter=20)
    assert np.abs(uY).max() > 0
    state = (Y.copy(), uY.copy(), g.copy())
    runs = {}
    for tag, at in (("ref", 250), ("early", 249), ("late", 251)):
        Yh, uh, gh = (a.copy() for a in state)
        _hand_steps(O, S, Yh, uh, gh, 247, 254, 4.0, 20, mom_at=at)
        runs[tag] = Yh
    Yo, uo, go = (a.copy() for a in state)
    O.run(S, Yo, uo, go, 247, 254, alpha0=4.0, early_iter=20)
    assert rel_l2(Yo, runs["ref"]) < 1e-13
    assert rel_l2(runs["early"], runs["ref"]) > 1e-7
    assert rel_l2(runs["late"], runs["ref"]) > 1e-7


@pytest.mark.parametrize("d", [1, 2, 3])
def test_fixed_grid_mode(O, d):
    # CU/gradient_descend.cu:139-140: nGrid = getBestGridSize(params.ng), so
    # params.ng = 50 gives a 52-node grid whatever the extent; at the same
    # nGrid the h rule gives the same numbers, and on the reference's
    # acceptance inputs (Y ~ U[0,1]^d) it meets the bar against the exact sum
    rng = np.random.default_rng(90 + d)
    n = 1500
    Y = rng.uniform(0, 1, size=(n, d))
    ng_fixed = {1: 50, 2: 50, 3: 20}[d]
    F, Z, ng = O.repulsive_interp(Y.copy(), 0.7, O.ZETA_ALGEBRAIC, fixed_ng=ng_fixed)
    want_grid = next(s for s in GRID_LIST if ng_fixed + 2 <= s)
    assert ng == want_grid
    maxy = (Y - Y.min(0)).max()
    h_same = maxy / (want_grid - 2 - 0.5)  # ceil(maxy / h) = nGrid
    F2, Z2, ng2 = O.repulsive_interp(Y.copy(), h_same, O.ZETA_ALGEBRAIC)
    assert ng2 == ng and np.array_equal(F, F2) and Z == Z2
    Fe, Ze = _exact_np(Y)
    assert np.abs(F - Fe).max() <= 1e-7 and abs(Z - Ze) / Ze <= 1e-4
    # a wide embedding keeps the fixed grid (the h rule would grow it)
    Yw = Y * (30.0 if d == 3 else 60.0)
    _, _, ngw = O.repulsive_interp(Yw.copy(), 0.7, O.ZETA_ALGEBRAIC, fixed_ng=ng_fixed)
    _, _, ngh = O.repulsive_interp(Yw.copy(), 0.7, O.ZETA_ALGEBRAIC)
    assert ngw == want_grid and ngh > want_grid
