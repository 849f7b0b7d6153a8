"""sgtsne_grid_conv_noalloc: the grid convolution with the caller's device
arrays and the library's static workspace, through the C-ABI.

Two calls in a row, the second on a smaller grid (so the shared workspace
holds the first call's data), each equal bit for bit to sgtsne_grid_conv
(the same kernels on the same input); sizes beyond the workspace are refused.
Device arrays come from the HIP runtime the library itself links."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _conv_noalloc(sg, hip, V, h, zdirect):
    d, ng = V.ndim - 1, V.shape[1]
    mo = d + 2 if zdirect else d + 1
    Phi = np.empty((mo,) + V.shape[1:], np.float32)
    dv, dp = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dv), V.nbytes) == 0 and hip.hipMalloc(C.byref(dp), Phi.nbytes) == 0
    try:
        assert hip.hipMemcpy(dv, V.ctypes.data, V.nbytes, H2D) == 0
        st = sg.lib().sgtsne_grid_conv_noalloc(d, ng, float(h), int(zdirect), dv, dp, None)
        assert st == 0, sg.lib().sgtsne_last_error()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(Phi.ctypes.data, dp, Phi.nbytes, D2H) == 0
    finally:
        hip.hipFree(dv)
        hip.hipFree(dp)
    return Phi


@pytest.mark.parametrize("first,second", [((2, 300), (2, 52)), ((3, 60), (3, 20)), ((2, 130), (1, 117))])
def test_grid_conv_noalloc_matches_allocating_entry(sg, first, second):
    hip = _hip()
    rng = np.random.default_rng(first[1] + second[1])
    for d, ng in (first, second):
        V = rng.uniform(0, 1, size=(d + 1,) + (ng,) * d).astype(np.float32)
        for zdirect in (True, False):
            want = sg.grid_conv(V, 0.3, zdirect=zdirect)
            got = _conv_noalloc(sg, hip, V, 0.3, zdirect)
            assert np.array_equal(got, want), (d, ng, zdirect)


def test_grid_conv_noalloc_refuses_oversize(sg):
    hip = _hip()
    dv = C.c_void_p()
    assert hip.hipMalloc(C.byref(dv), 1024) == 0
    try:
        st = sg.lib().sgtsne_grid_conv_noalloc(3, 80, 0.3, 1, dv, dv, None)
        assert st == -1  # SGTSNE_ERR_ARG
        assert b"allocation-free entry holds" in sg.lib().sgtsne_last_error()
    finally:
        hip.hipFree(dv)
