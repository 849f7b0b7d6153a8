"""ctypes binding of the FP64 oracle (oracle/sgtsne_oracle.c).

TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU baseline.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Y is AoS [n][d] float64 throughout; CSR uses int64 row pointers.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    sig = {
        "or_best_grid_size": (C.c_int, [C.c_int]),
        "or_knn_exact": (None, [_f32p, C.c_int, C.c_int, C.c_int, _i32p, _f64p]),
        "or_knn_rows": (None, [_f32p, C.c_int, C.c_int, C.c_int, _i32p, C.c_int, _i32p, _f64p]),
        "or_equalize_vertex": (C.c_int, [_f64p, _f64p, C.c_double, C.c_int, C.POINTER(C.c_double)]),
        "or_perplexity_equalize": (None, [_i32p, _f64p, C.c_int, C.c_int, C.c_double, _i64p, _i32p, _f64p]),
        "or_make_stochastic": (C.c_int, [C.c_int, _i64p, _f64p]),
        "or_lambda_rescale": (None, [C.c_int, _i64p, _f64p, C.c_double, C.c_int]),
        "or_symmetrize": (C.c_int64, [C.c_int, _i64p, _i32p, _f64p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "or_attractive": (None, [C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p]),
        "or_repulsive_exact": (C.c_double, [C.c_int, C.c_int, _f64p, _f64p]),
        "or_phi_exact": (None, [C.c_int, C.c_int, _f64p, _f64p]),
        "or_lagrange_weights": (None, [C.c_double, _f64p]),
        "or_s2g": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p]),
        "or_g2s": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p]),
        "or_conv_nopad_fft": (None, [C.c_int, C.c_int, C.c_int, C.c_double, _f64p, _f64p]),
        "or_conv_nopad_fft_k": (None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _f64p, _f64p]),
        "or_repulsive_interp_z": (C.c_double, [C.c_int, C.c_int, _f64p, C.c_double, C.c_int, _f64p,
                                               C.POINTER(C.c_int)]),
        "or_run_z": (None, [C.c_int, C.c_int, _i64p, _i32p, _f64p, C.c_double, C.c_int, C.c_double, C.c_int,
                            C.c_int, C.c_int, _f64p, _f64p, _f64p, C.c_void_p, C.c_void_p]),
        "or_conv_nopad_direct": (None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _f64p, _f64p]),
        "or_fft": (None, [C.c_int, _f64p, C.c_int]),
        "or_repulsive_interp": (C.c_double, [C.c_int, C.c_int, _f64p, C.c_double, _f64p, C.POINTER(C.c_int)]),
        "or_update": (None, [C.c_int, C.c_int, _f64p, _f64p, C.c_double, C.c_double, C.c_double,
                             _f64p, _f64p, _f64p]),
        "or_kl": (C.c_double, [C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double]),
        "or_run": (None, [C.c_int, C.c_int, _i64p, _i32p, _f64p, C.c_double, C.c_int, C.c_double,
                          C.c_int, C.c_int, _f64p, _f64p, _f64p, C.c_void_p, C.c_void_p]),
        "or_repulsive_interp_g": (C.c_double, [C.c_int, C.c_int, _f64p, C.c_double, C.c_int, C.c_int, _f64p,
                                               C.POINTER(C.c_int)]),
        "or_run_g": (None, [C.c_int, C.c_int, _i64p, _i32p, _f64p, C.c_double, C.c_int, C.c_double, C.c_int,
                            C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p, C.c_void_p, C.c_void_p]),
        "or_set_threads": (None, [C.c_int]),
        "or_get_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class CSR:
    """CSR matrix (int64 rowptr, int32 col, float64 val)."""

    def __init__(self, n, rowptr, col, val):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def dense(self):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            s, e = self.rowptr[i], self.rowptr[i + 1]
            np.add.at(A[i], self.col[s:e], self.val[s:e])
        return A


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def best_grid_size(n):
    return lib().or_best_grid_size(int(n))


def knn_exact(X, k):
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, dim = X.shape
    I = np.empty((n, k + 1), np.int32)
    D = np.empty((n, k + 1), np.float64)
    lib().or_knn_exact(X, n, dim, k, I, D)
    return I, D


def knn_rows(X, k, rows):
    """knn_exact's rows for the query points `rows` only."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    n, dim = X.shape
    I = np.empty((rows.size, k + 1), np.int32)
    D = np.empty((rows.size, k + 1), np.float64)
    lib().or_knn_rows(X, n, dim, k, rows, rows.size, I, D)
    return I, D


def equalize_vertex(dist, u):
    dist = _f64(dist)
    k = dist.size - 1
    val = np.empty(k)
    beta = C.c_double()
    it = lib().or_equalize_vertex(val, dist, float(u), k, C.byref(beta))
    return val, beta.value, it


def perplexity_equalize(I, D, u):
    I = np.ascontiguousarray(I, dtype=np.int32)
    D = _f64(D)
    n, k1 = I.shape
    k = k1 - 1
    rowptr = np.empty(n + 1, np.int64)
    col = np.empty(n * k, np.int32)
    val = np.empty(n * k, np.float64)
    lib().or_perplexity_equalize(I, D, n, k, float(u), rowptr, col, val)
    return CSR(n, rowptr, col, val)


def make_stochastic(P):
    return lib().or_make_stochastic(P.n, P.rowptr, P.val)


def lambda_rescale(P, lam, drop_leaf=False):
    lib().or_lambda_rescale(P.n, P.rowptr, P.val, float(lam), int(bool(drop_leaf)))


def symmetrize(P):
    L = lib()
    nnz = L.or_symmetrize(P.n, P.rowptr, P.col, P.val, None, None, None)
    rp = np.empty(P.n + 1, np.int64)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    L.or_symmetrize(P.n, P.rowptr, P.col, P.val, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return CSR(P.n, rp, col, val)


def attractive(P, Y):
    Y = _f64(Y)
    n, d = Y.shape
    F = np.empty_like(Y)
    lib().or_attractive(n, d, P.rowptr, P.col, P.val, Y, F)
    return F


def repulsive_exact(Y):
    Y = _f64(Y)
    n, d = Y.shape
    F = np.zeros_like(Y)
    Z = lib().or_repulsive_exact(n, d, Y, F)
    return F, Z


def phi_exact(Y):
    Y = _f64(Y)
    n, d = Y.shape
    Phi = np.empty((n, d + 1))
    lib().or_phi_exact(n, d, Y, Phi)
    return Phi


def lagrange_weights(delta):
    w = np.empty(4)
    lib().or_lagrange_weights(float(delta), w)
    return w


def s2g(y, q, ng):
    y = _f64(y)
    q = _f64(q)
    n, d = y.shape
    nvec = q.shape[1]
    V = np.zeros((nvec,) + (ng,) * d)
    lib().or_s2g(n, d, ng, nvec, y, q, V)
    return V


def g2s(y, V):
    y = _f64(y)
    V = _f64(V)
    n, d = y.shape
    nvec, ng = V.shape[0], V.shape[1]
    Phi = np.empty((n, nvec))
    lib().or_g2s(n, d, ng, nvec, y, V, Phi)
    return Phi


def conv_nopad_fft(V, h, kpow=2):
    V = _f64(V)
    nvec, ng = V.shape[0], V.shape[1]
    d = V.ndim - 1
    Phi = np.empty_like(V)
    lib().or_conv_nopad_fft_k(d, ng, nvec, float(h), int(kpow), V, Phi)
    return Phi


def conv_nopad_direct(V, h, kpow=2):
    V = _f64(V)
    nvec, ng = V.shape[0], V.shape[1]
    d = V.ndim - 1
    Phi = np.empty_like(V)
    lib().or_conv_nopad_direct(d, ng, nvec, float(h), int(kpow), V, Phi)
    return Phi


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    buf = x.view(np.float64).copy()
    lib().or_fft(x.size, buf, int(sign))
    return buf.view(np.complex128)


ZETA_DIRECT = 1      # oracle zmode numbering: 0 = algebraic (reference), 1 = direct
ZETA_ALGEBRAIC = 0


def repulsive_interp(Y, h, zmode=ZETA_ALGEBRAIC, fixed_ng=0):
    """Returns (Frep, Z, ng).  Y is shifted to min 0 in place (reference
    semantics); pass a copy to keep the original.  fixed_ng > 0: the CUDA
    path's fixed grid (nGrid = best_grid_size(fixed_ng)) instead of the h rule."""
    assert Y.dtype == np.float64 and Y.flags.c_contiguous
    n, d = Y.shape
    F = np.zeros_like(Y)
    ng = C.c_int()
    Z = lib().or_repulsive_interp_g(n, d, Y, float(h), int(zmode), int(fixed_ng), F, C.byref(ng))
    return F, Z, ng.value


def update(Fattr, Frep, alpha, momentum, eta, Y, uY, gains):
    n, d = Y.shape
    lib().or_update(n, d, _f64(Fattr), _f64(Frep), float(alpha), float(momentum), float(eta), Y, uY, gains)


def kl(P, Y, Z, alpha=1.0):
    Y = _f64(Y)
    n, d = Y.shape
    return lib().or_kl(n, d, P.rowptr, P.col, P.val, Y, float(Z), float(alpha))


def run(P, Y, uY, gains, it0, it1, alpha0=12.0, early_iter=250, h=None, zmode=ZETA_ALGEBRAIC, fixed_ng=0):
    """Runs iterations [it0, it1) in place; returns (Z per iteration, ng per iteration)."""
    n, d = Y.shape
    if h is None or h <= 0:
        h = {1: 0.5, 2: 0.7, 3: 1.2}[d]
    cnt = it1 - it0
    Z = np.empty(max(cnt, 1))
    ng = np.empty(max(cnt, 1), np.int32)
    lib().or_run_g(n, d, P.rowptr, P.col, P.val, float(alpha0), int(early_iter), float(h), int(zmode),
                   int(fixed_ng), int(it0), int(it1), Y, uY, gains, Z.ctypes.data, ng.ctypes.data)
    return Z[:cnt], ng[:cnt]
