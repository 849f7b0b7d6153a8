"""The C-ABI library (CPU-side checks: no kernel is launched)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "sgtsne.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char \*)\s*(sgtsne_\w+)\(", src, re.M)))


def test_header_declares_binding_list(sg):
    assert set(declared_symbols()) == set(sg.EXPORTED)


def test_library_exports_every_declared_symbol(sg):
    out = subprocess.run(["nm", "-D", "--defined-only", sg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (sgtsne_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = sg.lib()
    for s in declared_symbols():
        assert getattr(lib, s) is not None


def test_params_default_matches_reference(sg):
    p = sg.tsneparams()
    # CU/types.hpp:55-72 defaults (format: this engine's default is its own CSR kernel)
    assert (p.d, p.lambda_, p.alpha, p.maxIter, p.earlyIter, p.h, p.ng) == (2, 1.0, 12.0, 1000, 250, -1.0, 50)
    assert p.format == sg.FORMAT_CSR_SEG and p.bs == 0 and p.grid_mode == sg.GRID_H


def test_argument_errors_without_device(sg):
    lib = sg.lib()
    # NULL / bad sizes are rejected before any device work
    I = np.zeros((4, 3), np.int32)
    D = np.zeros((4, 3))
    rp = np.zeros(5, np.int32)
    col = np.zeros(8, np.int32)
    val = np.zeros(8, np.float32)
    assert lib.sgtsne_perplexity_equalize(I, D, 0, 2, 30.0, rp, col, val) == -1
    X = np.zeros((3, 2), np.float32)
    assert lib.sgtsne_knn_exact(X, 3, 2, 5, np.zeros((3, 6), np.int32), np.zeros((3, 6))) == -1
    V = np.zeros(4, np.float32)
    assert lib.sgtsne_grid_conv(4, 8, 1.0, 0, V, V) == -1
    p = sg.tsneparams(d=4)
    h = C.c_void_p()
    assert lib.sgtsne_ctx_create(C.byref(h), C.byref(p)) == -1
    assert b"d must be" in lib.sgtsne_last_error()
    p = sg.tsneparams(format=1)
    assert lib.sgtsne_ctx_create(C.byref(h), C.byref(p)) == -1
    # the asynchronous front end of a batch: a NULL batch is an argument error
    t = C.c_int64(-1)
    job = C.c_int(7)
    assert lib.sgtsne_batch_add_points_async(None, X, 3, 2, 1, 10.0, None, 0, 0, C.byref(t)) == -1
    assert t.value == -1
    assert lib.sgtsne_batch_add_poll(None, 1, 0, C.byref(job)) == -1


def test_product_never_uses_the_oracle():
    # the product sources and library neither import nor link the oracle
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt, f
    so = os.path.join(PKG, "lib", "libsgtsne_hip.so")
    if os.path.exists(so):
        out = subprocess.run(["ldd", so], capture_output=True, text=True).stdout
        assert "oracle" not in out


def test_cpp_mirror_header_compiles_and_links(sg, tmp_path):
    # include/sgtsne.hpp (the reference-shaped C++ entry points) builds and
    # links against the library; the program makes no call that needs a GPU.
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "sgtsne.hpp"\n'
        "int main(int argc, char **) {\n"
        "  tsneparams p; sparse_matrix<float> P;\n"
        "  if (argc > 99) {\n"
        "    float *y = sgtsneCUDA<float>(&P, p, nullptr, nullptr); free(y);\n"
        "    auto Q = perplexityEqualization<float>(nullptr, nullptr, 0, 0, 30.f); (void)Q;\n"
        "    sgtsne_hyb::sparse_matrix H{}; sgtsne_hyb::tsneparams hp;\n"
        "    double *yy = sgtsne_hyb::sgtsne(H, hp); free(yy);\n"
        "  }\n"
        "  return (p.d == 2 && p.maxIter == 1000 && p.earlyIter == 250 && p.ng == 50) ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "use"
    libdir = os.path.dirname(sg.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lsgtsne_hip", "-o", str(exe)], check=True)
    env = dict(os.environ, LD_LIBRARY_PATH=libdir)
    assert subprocess.run([str(exe)], env=env).returncode == 0


def _bandwidth(rp, col, perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    return int(np.abs(inv[rows] - inv[col]).max())


def test_locality_order_is_rcm(sg):
    # host-only entry: a permutation whose bandwidth matches scipy's reverse
    # Cuthill-McKee on a kNN graph of points in the plane given in random order
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    rng = np.random.default_rng(0)
    n, k = 3000, 8
    pts = rng.uniform(0, 1, size=(n, 2))
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    nbr = np.argsort(d2, axis=1)[:, 1:k + 1]
    A = csr_matrix((np.ones(n * k), nbr.ravel(), np.arange(0, n * k + 1, k)), shape=(n, n))
    A = (A + A.T).tocsr()
    A.sort_indices()
    rp, col = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    P = sg.CSR(n, rp, col, np.ones(col.size, np.float32))
    perm = sg.locality_order(P)
    assert np.array_equal(np.sort(perm), np.arange(n))
    ours = _bandwidth(rp, col, perm)
    ref = _bandwidth(rp, col, reverse_cuthill_mckee(A, symmetric_mode=True).astype(np.int64))
    rand = _bandwidth(rp, col, np.arange(n))
    assert ours <= 1.25 * ref and ours < rand / 5, (ours, ref, rand)


def _rcm_sequential(n, rp, col):
    # the order the library computes, one vertex at a time: per component, a
    # pseudo-peripheral root (George-Liu: up to 4 re-sweeps from the
    # lowest-degree vertex of the last level while the depth grows), a
    # Cuthill-McKee sweep taking unplaced neighbours by (degree, index), reversed
    deg = np.diff(rp)
    placed = np.zeros(n, bool)
    perm = []

    def levels(root):
        seen = {root}
        front, depth = [root], 0
        while True:
            nxt = []
            for v in front:
                for u in col[rp[v]:rp[v + 1]]:
                    if u not in seen:
                        seen.add(u)
                        nxt.append(int(u))
            if not nxt:
                return depth, front
            depth, front = depth + 1, nxt

    for cand in sorted(range(n), key=lambda i: (deg[i], i)):
        if placed[cand]:
            continue
        root = cand
        depth, last = levels(root)
        for _ in range(4):
            best = min(last, key=lambda v: (deg[v], v))
            d2, last2 = levels(best)
            if d2 <= depth:
                break
            depth, root, last = d2, best, last2
        head = len(perm)
        perm.append(root)
        placed[root] = True
        while head < len(perm):
            v = perm[head]
            nb = []
            for u in col[rp[v]:rp[v + 1]]:
                if not placed[u]:
                    placed[u] = True
                    nb.append((int(deg[u]), int(u)))
            perm.extend(u for _, u in sorted(nb))
            head += 1
    return np.array(perm[::-1], np.int64)


def test_locality_order_matches_sequential_sweep(sg):
    # levels of thousands of vertices take the multi-threaded sweep; the
    # order must be the one-vertex-at-a-time order bit for bit
    from scipy.sparse import csr_matrix

    rng = np.random.default_rng(3)
    n, k = 30000, 10
    A = csr_matrix((np.ones(n * k), rng.integers(0, n, n * k), np.arange(0, n * k + 1, k)), shape=(n, n))
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    A.sort_indices()
    rp, col = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    perm = sg.locality_order(sg.CSR(n, rp, col, np.ones(col.size, np.float32)))
    assert np.array_equal(perm, _rcm_sequential(n, rp, col))


def test_locality_order_components_and_isolated(sg):
    # two components and isolated vertices are all placed exactly once
    rp = np.array([0, 1, 2, 2, 3, 4, 4], np.int32)
    col = np.array([1, 0, 4, 3], np.int32)
    perm = sg.locality_order(sg.CSR(6, rp, col, np.ones(4, np.float32)))
    assert np.array_equal(np.sort(perm), np.arange(6))


_HYB_PROG = r'''
#include "sgtsne.hpp"
#include <cstdio>
#include <vector>
// reads n, nnz, CSC offsets (uint32), row ids (uint32), values (double) from
// argv[1]; mode "stoch": writes makeStochastic's values and count to argv[2];
// mode "run": writes sgtsne_hyb::sgtsne's Y (d = 2, 45 iterations, fixed
// start) to argv[2]
int main(int argc, char **argv) {
  FILE *f = std::fopen(argv[1], "rb");
  int n, nnz;
  if (std::fread(&n, 4, 1, f) != 1 || std::fread(&nnz, 4, 1, f) != 1) return 2;
  std::vector<uint32_t> col(n + 1), row(nnz);
  std::vector<double> val(nnz);
  if (std::fread(col.data(), 4, n + 1, f) != (size_t)n + 1) return 2;
  if (std::fread(row.data(), 4, nnz, f) != (size_t)nnz) return 2;
  if (std::fread(val.data(), 8, nnz, f) != (size_t)nnz) return 2;
  std::fclose(f);
  sgtsne_hyb::sparse_matrix P{n, n, nnz, row.data(), col.data(), val.data()};
  FILE *o = std::fopen(argv[2], "wb");
  if (std::string(argv[3]) == "stoch") {
    uint32_t k = sgtsne_hyb::makeStochastic(P);
    std::fwrite(&k, 4, 1, o);
    std::fwrite(val.data(), 8, nnz, o);
  } else {
    sgtsne_hyb::tsneparams p;
    p.maxIter = 45;  // below iteration 50, where a timed choice of point order could differ between runs
    p.earlyIter = 20;
    std::vector<double> y0(2 * (size_t)n);
    for (int i = 0; i < 2 * n; i++) y0[i] = 1e-3 * ((i * 7919) % 1000 - 500) / 500.0;
    double *y = sgtsne_hyb::sgtsne(P, p, y0.data());
    std::fwrite(y, 8, 2 * (size_t)n, o);
    std::free(y);
  }
  std::fclose(o);
  return 0;
}
'''


def _hyb_exe(sg, tmp_path):
    src = tmp_path / "hyb.cpp"
    src.write_text(_HYB_PROG)
    exe = tmp_path / "hyb"
    libdir = os.path.dirname(sg.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                    "-lsgtsne_hip", "-o", str(exe)], check=True)
    return exe, dict(os.environ, LD_LIBRARY_PATH=libdir)


def _write_csc(path, n, colptr, rows, vals):
    with open(path, "wb") as f:
        np.array([n, rows.size], np.int32).tofile(f)
        colptr.astype(np.uint32).tofile(f)
        rows.astype(np.uint32).tofile(f)
        vals.astype(np.float64).tofile(f)


def _weighted_graph(n, k, seed):
    # a general weighted kNN-like CSC (not stochastic): column j holds k
    # entries of positive weight, plus an all-zero column and a stochastic one
    rng = np.random.default_rng(seed)
    colptr = np.arange(0, n * k + 1, k)
    rows = np.concatenate([np.sort(rng.choice(np.delete(np.arange(n), j), k, replace=False)) for j in range(n)])
    vals = rng.uniform(0.1, 3.0, n * k)
    vals[k * 3:k * 4] = 0.0                                        # zero-sum column 3: left alone
    vals[k * 5:k * 6] = 1.0 / k                                    # column 5 already stochastic
    return colptr, rows, vals


def test_hyb_make_stochastic_matches_oracle(sg, O, tmp_path):
    # sgtsne_hyb::makeStochastic (HY/sparsematrix.cpp:27-67) against the
    # oracle's restatement, on the CSC offsets as the oracle's row pointers
    n, k = 200, 12
    colptr, rows, vals = _weighted_graph(n, k, 3)
    _write_csc(tmp_path / "P.bin", n, colptr, rows, vals)
    exe, env = _hyb_exe(sg, tmp_path)
    subprocess.run([str(exe), str(tmp_path / "P.bin"), str(tmp_path / "out.bin"), "stoch"], env=env, check=True)
    out = np.fromfile(tmp_path / "out.bin", np.uint8)
    cnt = int(out[:4].view(np.uint32)[0])
    got = out[4:].view(np.float64)
    want = vals.copy()
    cnt_o = O.lib().or_make_stochastic(n, colptr.astype(np.int64), want)
    assert cnt == cnt_o >= 1
    assert np.array_equal(got, want)
    sums = np.add.reduceat(got, colptr[:-1])
    assert np.allclose(np.delete(sums, 3), 1.0, atol=1e-12) and sums[3] == 0.0


@pytest.mark.gpu
def test_hyb_entry_normalises_like_the_reference(sg, tmp_path):
    # the HYB drop-in on a non-stochastic P gives the same embedding as on the
    # same P normalised beforehand (HY/sgtsne.cpp:60), bit for bit.  The
    # column sums are sequential FP64 sums, as in the reference's loop
    # (np.cumsum adds in order), so the normalised columns sum to 1 within
    # 1e-12 and the entry leaves them as they are
    n, k = 1500, 20
    colptr, rows, vals = _weighted_graph(n, k, 4)
    vals[k * 3:k * 4] = 0.5  # every column has a positive sum here
    norm = vals / np.repeat(np.cumsum(vals.reshape(n, k), axis=1)[:, -1], k)
    exe, env = _hyb_exe(sg, tmp_path)
    ys = []
    for tag, v in (("raw", vals), ("norm", norm)):
        _write_csc(tmp_path / f"{tag}.bin", n, colptr, rows, v)
        subprocess.run([str(exe), str(tmp_path / f"{tag}.bin"), str(tmp_path / f"{tag}.y"), "run"], env=env,
                       check=True, timeout=120)
        ys.append(np.fromfile(tmp_path / f"{tag}.y", np.float64))
    assert np.all(np.isfinite(ys[0]))
    assert np.array_equal(ys[0], ys[1])
