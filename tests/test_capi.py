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
