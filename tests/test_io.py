"""File formats and command-line programs of the reference (CPU only).

The writers are checked byte for byte against what std::ofstream writes, the
way the reference's extractEmbeddingText(T) and ExtractTimeInfo use it
(CU/utils.cu:72-138, HY/utils.cpp:180-206): a small C++ program compiled here
writes the same values with operator<<, and the files must be identical.  The
readers are checked against numpy parses of files written with known values,
and the programs' option parsing through their dry-run mode."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

BIN = os.path.join(PKG, "bin")

_OSTREAM_PROG = r'''
#include <fstream>
#include <vector>
#include <cstdio>
// argv: in.bin n d kind out ; kind: f32soa | f32aos | f64aos | time
int main(int argc, char **argv) {
  FILE *f = std::fopen(argv[1], "rb");
  int n = std::atoi(argv[2]), d = std::atoi(argv[3]);
  std::string kind = argv[4];
  if (kind == "f32soa" || kind == "f32aos") {
    std::vector<float> y((size_t)n * d);
    if (std::fread(y.data(), 4, y.size(), f) != y.size()) return 2;
    std::ofstream o(argv[5]);
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < d; j++) o << (kind == "f32soa" ? y[i + j * n] : y[i * d + j]) << " ";
      o << std::endl;
    }
  } else if (kind == "f64aos") {
    std::vector<double> y((size_t)n * d);
    if (std::fread(y.data(), 8, y.size(), f) != y.size()) return 2;
    std::ofstream o(argv[5]);
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < d; j++) o << y[i * d + j] << " ";
      o << std::endl;
    }
  } else {  // time: n d iters, 7*iters+3 doubles, appended as ExtractTimeInfo does
    int iters = std::atoi(argv[6]);
    std::vector<double> t((size_t)7 * iters + 3);
    if (std::fread(t.data(), 8, t.size(), f) != t.size()) return 2;
    std::ofstream o(argv[5], std::ios::app);
    o << "N d gds matrix total: \n";
    o << n << " " << d << " " << t[7 * iters] << " " << t[7 * iters + 1] << " " << t[7 * iters + 2] << "\n";
    o << "Fattr S2G G2G G2S Phi2F Nuconv Pre";
    for (int i = 0; i < iters; i++) {
      for (int j = 0; j < 7; j++) o << t[i * 7 + j] << " ";
      o << "\n";
    }
  }
  std::fclose(f);
  return 0;
}
'''


@pytest.fixture(scope="module")
def ostream_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("ostream")
    (d / "w.cpp").write_text(_OSTREAM_PROG)
    subprocess.run(["g++", "-std=c++17", "-O1", str(d / "w.cpp"), "-o", str(d / "w")], check=True)
    return d / "w"


def _awkward_values(rng, count, dtype):
    # every magnitude class the %g format switches between, signs, zero,
    # integers, values that round at the 6th digit
    v = rng.standard_normal(count) * 10.0 ** rng.integers(-38 if dtype == np.float32 else -300, 30, count)
    v[:8] = [0.0, -0.0, 1.0, -123456.5, 999999.5, 1e-5, 0.0001, 1234567.0]
    return v.astype(dtype)


@pytest.mark.parametrize("layout", ["soa", "aos"])
def test_embedding_text_matches_ofstream(sg, ostream_exe, tmp_path, layout):
    rng = np.random.default_rng(1)
    n, d = 300, 3
    y = _awkward_values(rng, n * d, np.float32)
    y.tofile(tmp_path / "y.bin")
    subprocess.run([str(ostream_exe), str(tmp_path / "y.bin"), str(n), str(d), "f32" + layout,
                    str(tmp_path / "want.txt")], check=True)
    Y = y.reshape(d, n) if layout == "soa" else y.reshape(n, d)
    sg.write_embedding_text(tmp_path / "got.txt", Y, sg.LAYOUT_SOA if layout == "soa" else sg.LAYOUT_AOS)
    assert (tmp_path / "got.txt").read_bytes() == (tmp_path / "want.txt").read_bytes()


def test_embedding_text_f64_and_bin(sg, ostream_exe, tmp_path):
    rng = np.random.default_rng(2)
    n, d = 200, 2
    y = _awkward_values(rng, n * d, np.float64)
    y.tofile(tmp_path / "y.bin")
    subprocess.run([str(ostream_exe), str(tmp_path / "y.bin"), str(n), str(d), "f64aos", str(tmp_path / "want.txt")],
                   check=True)
    sg.write_embedding_text(tmp_path / "got.txt", y.reshape(n, d))
    assert (tmp_path / "got.txt").read_bytes() == (tmp_path / "want.txt").read_bytes()
    sg.write_embedding_bin(tmp_path / "e.bin", y.reshape(n, d))
    assert np.array_equal(np.fromfile(tmp_path / "e.bin", np.float64), y)


def test_time_info_appends_only_to_an_existing_file(sg, ostream_exe, tmp_path):
    rng = np.random.default_rng(3)
    iters, n, d = 20, 1000, 2
    t = rng.uniform(0, 5, 7 * iters + 3)
    t.tofile(tmp_path / "t.bin")
    p = tmp_path / "timeInfo.txt"
    sg.write_time_info(p, t, n, d, iters)
    assert p.exists() and p.read_bytes() == b""  # a fresh file stays empty, as in the reference
    sg.write_time_info(p, t, n, d, iters)
    want = tmp_path / "want.txt"
    want.write_bytes(b"")
    subprocess.run([str(ostream_exe), str(tmp_path / "t.bin"), str(n), str(d), "time", str(want), str(iters)],
                   check=True)
    assert p.read_bytes() == want.read_bytes()


def test_read_X_text(sg, tmp_path):
    rng = np.random.default_rng(4)
    X = (rng.standard_normal((57, 9)) * 1e3).astype(np.float32)
    lines = ["57 9"] + ["\t".join("%.9g" % v for v in row) + "  " for row in X]
    (tmp_path / "x.txt").write_text("\n".join(lines) + "\n")
    got = sg.read_X_text(tmp_path / "x.txt")
    assert got.dtype == np.float32 and np.array_equal(got, X)
    # a positive n reads only the first n points (the reference's -n)
    assert np.array_equal(sg.read_X_text(tmp_path / "x.txt", n=10), X[:10])
    with pytest.raises(sg.SgtsneError):
        sg.read_X_text(tmp_path / "x.txt", n=58)
    (tmp_path / "short.txt").write_text("3 2\n1 2\n3 4\n5\n")
    with pytest.raises(sg.SgtsneError, match="missing"):
        sg.read_X_text(tmp_path / "short.txt")
    with pytest.raises(sg.SgtsneError):
        sg.read_X_text(tmp_path / "absent.txt")


def test_read_P_mtx(sg, tmp_path):
    # 1-based COO in arbitrary order, comment lines first; the result is CSC
    # with each column's entries in file order (a stable counting sort)
    rng = np.random.default_rng(5)
    m = n = 40
    nnz = 300
    r = rng.integers(0, m, nnz)
    c = rng.integers(0, n, nnz)
    v = rng.uniform(-1, 1, nnz)
    body = "".join("%d %d %.17g\n" % (a + 1, b + 1, x) for a, b, x in zip(r, c, v))
    (tmp_path / "p.mtx").write_text("%%MatrixMarket matrix coordinate real general\n% a comment\n"
                                    f"{m} {n} {nnz}\n" + body)
    mm, nn, colptr, row, val = sg.read_P_mtx(tmp_path / "p.mtx")
    order = np.argsort(c, kind="stable")
    assert (mm, nn) == (m, n)
    assert np.array_equal(colptr, np.concatenate([[0], np.cumsum(np.bincount(c, minlength=n))]))
    assert np.array_equal(row, r[order]) and np.array_equal(val, v[order])
    # the same matrix as scipy's Matrix Market reader (duplicates summed)
    import scipy.io
    from scipy.sparse import csc_matrix
    A = scipy.io.mmread(str(tmp_path / "p.mtx")).toarray()
    B = csc_matrix((val, row, colptr), shape=(m, n)).toarray()
    assert np.allclose(A, B, rtol=0, atol=1e-15)
    (tmp_path / "bad.mtx").write_text("3 3 1\n4 1 0.5\n")
    with pytest.raises(sg.SgtsneError, match="out of range"):
        sg.read_P_mtx(tmp_path / "bad.mtx")


def test_simulation_data(sg):
    # CU/utils.cu:310-332: group j of n/c consecutive points has coordinate a
    # from N(20, 2) when a >= j, else N(-20, 2); deterministic
    n, dim, c = 4000, 12, 4
    X = sg.simulation_data(n, dim, c).astype(np.float64)
    assert np.array_equal(X, sg.simulation_data(n, dim, c))
    for j in range(c):
        blk = X[j * n // c:(j + 1) * n // c]
        for a in range(dim):
            mu = 20.0 if a >= j else -20.0
            assert abs(blk[:, a].mean() - mu) < 0.25 and abs(blk[:, a].std() - 2.0) < 0.2


def test_random_start(sg):
    # CU/utils.cu:185-192: 1e-4 N(0, 1), drawn in [d][N] order; the AoS form
    # is the same points
    y = sg.random_start(20000, 2, sg.LAYOUT_SOA)
    assert np.array_equal(sg.random_start(20000, 2, sg.LAYOUT_AOS), y.T)
    assert abs(y.mean()) < 3e-6 and abs(y.std() - 1e-4) < 3e-6
    # one default-seeded stream: every call starts the same sequence
    assert np.array_equal(sg.random_start(5, 2, sg.LAYOUT_SOA).reshape(-1), y.reshape(-1)[:10])


def _demo(name, args, cwd, env_extra=None):
    env = dict(os.environ, SGTSNE_DEMO_DRY_RUN="1", **(env_extra or {}))
    return subprocess.run([os.path.join(BIN, name)] + args, cwd=cwd, env=env, capture_output=True, text=True,
                          timeout=60)


def test_demo_perplexity_options(sg, tmp_path):
    # the flag set of CU/testing_modules/demo_perplexity.cu:55-97
    X = np.arange(12, dtype=np.float32).reshape(4, 3)
    (tmp_path / "x.txt").write_text("4 3\n" + "\n".join(" ".join(str(v) for v in r) for r in X) + "\n")
    r = _demo("sgtsne_demo_perplexity", ["-d", "3", "-a", "8", "-m", "700", "-e", "100", "-h", "0.9", "-p", "4",
                                         "-u", "10", "-f", "1", "-g", "64", "-t", "3", "-n", "3", "x.txt"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert ("d=3 alpha=8 maxIter=700 earlyIter=100 h=0.9 np=4 u=10 ComputeError=1 ng=64 sim=0 n=3 format=3"
            in r.stdout)
    assert "points=3 dim=3 neighbours=30 file=x.txt" in r.stdout
    assert "sum=36" in r.stdout  # the first three points only
    # defaults (CU/types.hpp:55-72), perplexity 30 -> 90 neighbours
    r = _demo("sgtsne_demo_perplexity", ["x.txt"], tmp_path)
    assert "d=2 alpha=12 maxIter=1000 earlyIter=250 h=-1 np=0 u=30 ComputeError=-1 ng=50 sim=0 n=4 format=2" in r.stdout
    assert "neighbours=90" in r.stdout
    # synthetic data (-s clusters): 50-D
    r = _demo("sgtsne_demo_perplexity", ["-s", "3", "-n", "90", "unused"], tmp_path)
    assert r.returncode == 0 and "points=90 dim=50" in r.stdout
    # errors as in the reference
    r = _demo("sgtsne_demo_perplexity", ["-d", "2", "-m"], tmp_path)  # the last argument is an option
    assert r.returncode == 1 and "No filename provided!" in r.stderr
    r = _demo("sgtsne_demo_perplexity", ["-d", "2"], tmp_path)  # "2" is taken as the file, as in the reference
    assert r.returncode == 1 and "cannot open 2" in r.stderr
    r = _demo("sgtsne_demo_perplexity", ["-z", "x.txt"], tmp_path)
    assert r.returncode == 0 and "Unknown option: 'z'!" in r.stderr


def test_demo_stochastic_matrix_options(sg, tmp_path):
    # the flag set of HY/demo_stochastic_matrix.cpp:45-71
    (tmp_path / "p.mtx").write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 2 0.5\n2 1 1\n"
                                    "3 1 1\n1 3 0.5\n")
    r = _demo("sgtsne_demo_stochastic_matrix", ["-l", "0.5", "-d", "1", "-a", "4", "-m", "300", "-e", "50",
                                                "-h", "0.3", "-p", "2", "p.mtx"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "lambda=0.5 d=1 alpha=4 maxIter=300 earlyIter=50 h=0.3 np=2 n=3" in r.stdout
    assert "m=3 n=3 nnz=4 sum=3" in r.stdout
    r = _demo("sgtsne_demo_stochastic_matrix", [], tmp_path)
    assert r.returncode == 1 and "No filename provided!" in r.stderr
