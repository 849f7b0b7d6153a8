"""The two command-line programs end to end on the GPU.

Each writes its embedding file the way the reference does; the embedding read
back from that file must reach a final KL within 1% of the FP64 oracle's loop
run from the same P and the same start (the reference's random start, since no
yin2m.txt is present), the north-star bar for a full run."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu

BIN = os.path.join(PKG, "bin")


def _mixture(n, dim, clusters, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 4, size=(clusters, dim))
    return (c[rng.integers(0, clusters, n)] + rng.standard_normal((n, dim))).astype(np.float32)


def _final_kl(O, S, Y):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    _, Z, _ = O.repulsive_interp(Y.copy(), 0.7, O.ZETA_DIRECT)
    return O.kl(S, Y, Z)


def _oracle_kl(O, S, y0, iters):
    Y = y0.astype(np.float64).copy()
    O.run(S, Y, np.zeros_like(Y), np.ones_like(Y), 0, iters, zmode=O.ZETA_DIRECT)
    return _final_kl(O, S, Y)


def _run(args, cwd):
    print("running", args, flush=True)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def test_demo_perplexity_end_to_end(sg, O, tmp_path):
    n, dim, u = 2000, 20, 10.0
    X = _mixture(n, dim, 5, 17)
    (tmp_path / "x.txt").write_text(f"{n} {dim}\n" + "\n".join(" ".join("%.9g" % v for v in r) for r in X) + "\n")
    (tmp_path / "timeInfo.txt").write_text("")  # present: the report is appended
    _run([os.path.join(BIN, "sgtsne_demo_perplexity"), "-u", "10", "-t", "4", "x.txt"], tmp_path)
    lines = (tmp_path / "sgtsneEmbedding.txt").read_text().splitlines()
    assert len(lines) == n and all(line.endswith(" ") for line in lines)
    Y = np.array([[float(v) for v in line.split()] for line in lines])
    assert Y.shape == (n, 2) and np.all(np.isfinite(Y))
    t = (tmp_path / "timeInfo.txt").read_text().splitlines()
    assert t[0] == "N d gds matrix total: " and t[1].startswith(f"{n} 2 ") and len(t) == 2 + 1000
    # the oracle from the same P (exact kNN, float distances divided by their
    # maximum, as the program does) and the reference's start
    nn = 30
    I, D = O.knn_exact(X, nn)
    Df = D.astype(np.float32)
    Df /= Df.max()
    S = O.symmetrize(O.perplexity_equalize(I, Df.astype(np.float64), u))
    y0 = sg.random_start(n, 2, sg.LAYOUT_AOS)
    klo = _oracle_kl(O, S, y0, 1000)
    # the file holds 6 significant digits: ample for the KL
    klg = _final_kl(O, S, Y)
    assert abs(klg - klo) / klo < 0.01, (klg, klo)


def test_demo_stochastic_matrix_end_to_end(sg, O, tmp_path):
    n, dim, k, u = 1800, 15, 30, 10.0
    X = _mixture(n, dim, 4, 23)
    I, D = O.knn_exact(X, k)
    P = O.perplexity_equalize(I, D, u)  # row-stochastic CSR = column-stochastic CSC of P^T
    rows = np.repeat(np.arange(n), np.diff(P.rowptr))
    # file triplets (row, col) of the column-stochastic matrix: column j = row j of P
    body = "".join("%d %d %.17g\n" % (c + 1, r + 1, v) for r, c, v in zip(rows, P.col, P.val))
    (tmp_path / "p.mtx").write_text(f"%%MatrixMarket matrix coordinate real general\n{n} {n} {P.nnz}\n" + body)
    _run([os.path.join(BIN, "sgtsne_demo_stochastic_matrix"), "-m", "1000", "p.mtx"], tmp_path)
    lines = (tmp_path / "embedding.txt").read_text().splitlines()
    Y = np.array([[float(v) for v in line.split()] for line in lines])
    assert Y.shape == (n, 2) and np.all(np.isfinite(Y))
    # the oracle: the same column normalisation (makeStochastic), P + P^T
    val = P.val.copy()
    O.make_stochastic(O.CSR(n, P.rowptr, P.col, val))
    S = O.symmetrize(O.CSR(n, P.rowptr, P.col, val))
    y0 = sg.random_start(n, 2, sg.LAYOUT_AOS)
    klo = _oracle_kl(O, S, y0, 1000)
    klg = _final_kl(O, S, Y)
    assert abs(klg - klo) / klo < 0.01, (klg, klo)
