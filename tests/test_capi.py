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
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(sgtsne_\w+)\(", src, re.M)))


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


def test_locality_order_components_and_isolated(sg):
    # two components and isolated vertices are all placed exactly once
    rp = np.array([0, 1, 2, 2, 3, 4, 4], np.int32)
    col = np.array([1, 0, 4, 3], np.int32)
    perm = sg.locality_order(sg.CSR(6, rp, col, np.ones(4, np.float32)))
    assert np.array_equal(np.sort(perm), np.arange(6))
