"""Python binding of the gfx950 SG-t-SNE engine (C-ABI in include/sgtsne.h).

The names mirror the reference's interface:
  tsneparams               CU/types.hpp:55-72 (HY/sgtsne.hpp:25-37)
  sgtsne(P, params, y_in)  sgtsneCUDA CU/sgtsne.cu:26-128 / sgtsne HY/sgtsne.cpp:37-197
  perplexityEqualization   CU/perplexity_search.cu:114-141 / HY/sgtsne.cpp:257-297
plus the stepwise Context used for parity tests, batching and serving.

Every call goes to libsgtsne_hip.so; there is no CPU fallback: importing this
module without the built library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsgtsne_hip.so")

FORMAT_CSR = 0
FORMAT_COO = 2
FORMAT_HYBRID = 3
FORMAT_CSR_SEG = 4
GRID_H = 0
GRID_FIXED = 1
LAYOUT_SOA = 0
LAYOUT_AOS = 1
ZETA_DIRECT = 0
ZETA_ALGEBRAIC = 1
ERR_FULL = -8
JOB_RUNNING = 0
JOB_DONE = 1
SPREAD_AUTO = 0
SPREAD_ATOMIC = 1
SPREAD_MFMA = 2
KNN_AUTO = 0
KNN_EXACT_FP64 = 1
KNN_FILTERED = 2

EXPORTED = [
    "sgtsne_params_default", "sgtsne_last_error", "sgtsne_version", "sgtsne_knn_exact", "sgtsne_knn",
    "sgtsne_perplexity_equalize", "sgtsne_grid_conv", "sgtsne_run", "sgtsne_ctx_create", "sgtsne_ctx_upload_P",
    "sgtsne_ctx_get_P", "sgtsne_ctx_set_Y", "sgtsne_ctx_get_Y", "sgtsne_ctx_step",
    "sgtsne_ctx_terms", "sgtsne_ctx_kl", "sgtsne_ctx_repulsive_exact", "sgtsne_ctx_timing",
    "sgtsne_ctx_set_timing", "sgtsne_ctx_errors", "sgtsne_ctx_info", "sgtsne_ctx_destroy",
    "sgtsne_ctx_tuning", "sgtsne_ctx_set_attractive", "sgtsne_ctx_reorders",
    "sgtsne_batch_create", "sgtsne_batch_destroy", "sgtsne_batch_add", "sgtsne_batch_add_points",
    "sgtsne_batch_add_points_async", "sgtsne_batch_add_poll", "sgtsne_batch_step", "sgtsne_batch_sync", "sgtsne_batch_mark", "sgtsne_batch_wait",
    "sgtsne_batch_passed", "sgtsne_batch_job_left",
    "sgtsne_batch_job_status", "sgtsne_batch_get_Y",
    "sgtsne_batch_terms", "sgtsne_batch_remove", "sgtsne_batch_info", "sgtsne_batch_tuning",
    "sgtsne_batch_set_graph", "sgtsne_batch_set_attractive", "sgtsne_copy_probe", "sgtsne_locality_order",
    "sgtsne_free", "sgtsne_read_X_text", "sgtsne_read_P_mtx", "sgtsne_write_embedding_text",
    "sgtsne_write_embedding_text_f64", "sgtsne_write_embedding_bin", "sgtsne_write_time_info",
    "sgtsne_simulation_data", "sgtsne_random_start", "sgtsne_grid_conv_noalloc",
]


class SgtsneError(RuntimeError):
    pass


class tsneparams(C.Structure):
    """Mirror of sgtsne_params (every tsneparams field, CU/types.hpp:55-72)."""
    _fields_ = [
        ("d", C.c_int), ("lambda_", C.c_double), ("alpha", C.c_double), ("maxIter", C.c_int),
        ("earlyIter", C.c_int), ("n", C.c_int), ("h", C.c_double), ("dropLeaf", C.c_int),
        ("np", C.c_int), ("format", C.c_int), ("bs", C.c_int), ("ComputeError", C.c_int),
        ("ng", C.c_int), ("grid_mode", C.c_int), ("layout", C.c_int), ("seed", C.c_uint64),
        ("ell_width", C.c_int), ("zeta_mode", C.c_int), ("reorder", C.c_int), ("spread", C.c_int),
    ]

    def __init__(self, **kw):
        super().__init__()
        _check(lib().sgtsne_params_default(C.byref(self)))
        for k, v in kw.items():
            if k == "lambda":
                k = "lambda_"
            if not hasattr(self, k):
                raise AttributeError(f"tsneparams has no field {k}")
            setattr(self, k, v)


_lib = None


def lib():
    """Loads libsgtsne_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SgtsneError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    P = C.POINTER(tsneparams)
    sig = {
        "sgtsne_params_default": (C.c_int, [P]),
        "sgtsne_last_error": (C.c_char_p, []),
        "sgtsne_version": (C.c_char_p, []),
        "sgtsne_knn_exact": (C.c_int, [f32p, C.c_int, C.c_int, C.c_int, i32p, f64p]),
        "sgtsne_knn": (C.c_int, [f32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p, f64p, C.POINTER(C.c_int)]),
        "sgtsne_perplexity_equalize": (C.c_int, [i32p, f64p, C.c_int, C.c_int, C.c_double, i32p, i32p, f32p]),
        "sgtsne_grid_conv": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, f32p, f32p]),
        "sgtsne_run": (C.c_int, [C.c_int, C.c_int64, i32p, i32p, f32p, P, vp, f32p, vp]),
        "sgtsne_ctx_create": (C.c_int, [C.POINTER(vp), P]),
        "sgtsne_ctx_upload_P": (C.c_int, [vp, C.c_int, C.c_int64, i32p, i32p, f32p, C.c_int]),
        "sgtsne_ctx_get_P": (C.c_int, [vp, C.POINTER(C.c_int64), vp, vp, vp]),
        "sgtsne_ctx_set_Y": (C.c_int, [vp, f32p, C.c_int, C.c_int]),
        "sgtsne_ctx_get_Y": (C.c_int, [vp, f32p, C.c_int]),
        "sgtsne_ctx_step": (C.c_int, [vp, C.c_int]),
        "sgtsne_ctx_terms": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]),
        "sgtsne_ctx_kl": (C.c_int, [vp, C.c_double, C.POINTER(C.c_double)]),
        "sgtsne_ctx_repulsive_exact": (C.c_int, [vp, vp, C.POINTER(C.c_double), C.c_int]),
        "sgtsne_ctx_timing": (C.c_int, [vp, f64p]),
        "sgtsne_ctx_set_timing": (C.c_int, [vp, C.c_int]),
        "sgtsne_ctx_tuning": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), vp, C.c_int, C.POINTER(C.c_int)]),
        "sgtsne_ctx_errors": (C.c_int, [vp, vp, C.c_int, C.POINTER(C.c_int)]),
        "sgtsne_ctx_set_attractive": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "sgtsne_ctx_reorders": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "sgtsne_ctx_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64)]),
        "sgtsne_ctx_destroy": (C.c_int, [vp]),
        "sgtsne_batch_create": (C.c_int, [C.POINTER(vp), P, C.c_int, C.c_int64, C.c_int64, C.c_int64]),
        "sgtsne_batch_destroy": (C.c_int, [vp]),
        "sgtsne_batch_add": (C.c_int, [vp, C.c_int, C.c_int64, i32p, i32p, f32p, C.c_int, vp, C.c_int, C.c_int,
                                       C.POINTER(C.c_int)]),
        "sgtsne_batch_add_points": (C.c_int, [vp, f32p, C.c_int, C.c_int, C.c_int, C.c_double, vp, C.c_int,
                                              C.c_int, C.POINTER(C.c_int)]),
        "sgtsne_batch_add_points_async": (C.c_int, [vp, f32p, C.c_int, C.c_int, C.c_int, C.c_double, vp, C.c_int,
                                                    C.c_int, C.POINTER(C.c_int64)]),
        "sgtsne_batch_add_poll": (C.c_int, [vp, C.c_int64, C.c_int, C.POINTER(C.c_int)]),
        "sgtsne_batch_step": (C.c_int, [vp, C.c_int]),
        "sgtsne_batch_sync": (C.c_int, [vp]),
        "sgtsne_batch_mark": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "sgtsne_batch_wait": (C.c_int, [vp, C.c_int64]),
        "sgtsne_batch_passed": (C.c_int, [vp, C.c_int64, C.POINTER(C.c_int)]),
        "sgtsne_batch_job_left": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64)]),
        "sgtsne_batch_job_status": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "sgtsne_batch_get_Y": (C.c_int, [vp, C.c_int, f32p, C.c_int]),
        "sgtsne_batch_terms": (C.c_int, [vp, C.c_int, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]),
        "sgtsne_batch_remove": (C.c_int, [vp, C.c_int]),
        "sgtsne_batch_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "sgtsne_batch_tuning": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "sgtsne_batch_set_graph": (C.c_int, [vp, C.c_int]),
        "sgtsne_batch_set_attractive": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "sgtsne_copy_probe": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
        "sgtsne_locality_order": (C.c_int, [C.c_int, i32p, i32p, i32p]),
        "sgtsne_free": (None, [vp]),
        "sgtsne_read_X_text": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(vp)]),
        "sgtsne_read_P_mtx": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                        C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
        "sgtsne_write_embedding_text": (C.c_int, [C.c_char_p, f32p, C.c_int, C.c_int, C.c_int]),
        "sgtsne_write_embedding_text_f64": (C.c_int, [C.c_char_p, f64p, C.c_int, C.c_int, C.c_int]),
        "sgtsne_write_embedding_bin": (C.c_int, [C.c_char_p, f64p, C.c_int, C.c_int]),
        "sgtsne_write_time_info": (C.c_int, [C.c_char_p, f64p, C.c_int, C.c_int, C.c_int]),
        "sgtsne_simulation_data": (C.c_int, [C.c_int, C.c_int, C.c_int, f32p]),
        "sgtsne_random_start": (C.c_int, [C.c_int, C.c_int, C.c_int, f32p]),
        "sgtsne_grid_conv_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(status):
    if status != 0:
        msg = lib().sgtsne_last_error().decode(errors="replace")
        raise SgtsneError(f"sgtsne error {status}: {msg}")


def locality_order(P):
    """Reverse Cuthill-McKee order of P's graph (host): perm[r] = vertex at row r."""
    perm = np.empty(P.n, np.int32)
    _check(lib().sgtsne_locality_order(P.n, P.rowptr, P.col if P.nnz else np.zeros(1, np.int32), perm))
    return perm


def copy_probe(nbytes, reps=10):
    """ms of one device-to-device streaming copy of nbytes (plain kernel)."""
    ms = C.c_double()
    _check(lib().sgtsne_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def _take(ptr, count, ctype, dtype):
    """Copies a library-allocated array into numpy and frees it."""
    try:
        if count == 0:
            return np.zeros(0, dtype)
        buf = C.cast(ptr, C.POINTER(ctype * count)).contents
        return np.frombuffer(buf, dtype=dtype, count=count).copy()
    finally:
        lib().sgtsne_free(ptr)


def read_X_text(path, n=0):
    """Points file of the reference's demo (CU/utils.cu:216-256): float32 [n][dim]."""
    nn, dim, ptr = C.c_int(int(n)), C.c_int(), C.c_void_p()
    _check(lib().sgtsne_read_X_text(os.fsencode(path), C.byref(nn), C.byref(dim), C.byref(ptr)))
    return _take(ptr, nn.value * dim.value, C.c_float, np.float32).reshape(nn.value, dim.value)


def read_P_mtx(path):
    """Matrix Market graph (CU/utils.cu:335-409): (m, n, colptr, row, val) in
    compressed columns, 0-based."""
    m, n, nnz = C.c_int(), C.c_int(), C.c_int64()
    cp, ro, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
    _check(lib().sgtsne_read_P_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), C.byref(cp),
                                   C.byref(ro), C.byref(va)))
    colptr = _take(cp, n.value + 1, C.c_int32, np.int32)
    row = _take(ro, nnz.value, C.c_int32, np.int32)
    val = _take(va, nnz.value, C.c_double, np.float64)
    return m.value, n.value, colptr, row, val


def write_embedding_text(path, Y, layout=LAYOUT_AOS):
    """One line per point, '%.6g ' per coordinate (the reference's writers).
    Y is [n][d] for LAYOUT_AOS and [d][n] for LAYOUT_SOA."""
    Y = np.ascontiguousarray(Y)
    n, d = (Y.shape if layout == LAYOUT_AOS else Y.shape[::-1])
    if Y.dtype == np.float64:
        _check(lib().sgtsne_write_embedding_text_f64(os.fsencode(path), Y.reshape(-1), n, d, int(layout)))
    else:
        Y = Y.astype(np.float32, copy=False)
        _check(lib().sgtsne_write_embedding_text(os.fsencode(path), Y.reshape(-1), n, d, int(layout)))


def write_embedding_bin(path, Y):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    _check(lib().sgtsne_write_embedding_bin(os.fsencode(path), Y.reshape(-1), Y.shape[0], Y.shape[1]))


def write_time_info(path, time_info, n, d, iters):
    ti = np.ascontiguousarray(time_info, dtype=np.float64)
    assert ti.size >= 7 * iters + 3
    _check(lib().sgtsne_write_time_info(os.fsencode(path), ti, int(n), int(d), int(iters)))


def simulation_data(n, dim, clusters):
    """The reference's synthetic points for its -s option (CU/utils.cu:310-332)."""
    X = np.empty((n, dim), np.float32)
    _check(lib().sgtsne_simulation_data(int(n), int(dim), int(clusters), X.reshape(-1)))
    return X


def random_start(n, d, layout=LAYOUT_AOS):
    """The reference's N(0, 1e-4^2) start (CU/utils.cu:185-192)."""
    y = np.empty((n, d) if layout == LAYOUT_AOS else (d, n), np.float32)
    _check(lib().sgtsne_random_start(int(n), int(d), int(layout), y.reshape(-1)))
    return y


def version():
    return lib().sgtsne_version().decode()


class CSR:
    """Row-compressed P: int32 rowptr [n+1], int32 col [nnz], float32 val [nnz]."""

    def __init__(self, n, rowptr, col, val):
        self.n = int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float32)

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def knn_exact(X, k):
    """Exact kNN on the GPU: I [n][k+1] int32 (self first), D [n][k+1] float64."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, dim = X.shape
    I = np.empty((n, k + 1), np.int32)
    D = np.empty((n, k + 1), np.float64)
    _check(lib().sgtsne_knn_exact(X, n, dim, int(k), I, D))
    return I, D


def knn(X, k, method=KNN_AUTO):
    """knn_exact with a chosen method (KNN_AUTO, KNN_EXACT_FP64, KNN_FILTERED);
    returns (I, D, rows the filtered method recomputed exactly, or -1)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, dim = X.shape
    I = np.empty((n, k + 1), np.int32)
    D = np.empty((n, k + 1), np.float64)
    fb = C.c_int(-1)
    _check(lib().sgtsne_knn(X, n, dim, int(k), int(method), I, D, C.byref(fb)))
    return I, D, fb.value


def grid_conv(V, h, zdirect=False):
    """GPU non-periodic convolution of V [d+1][ng]^d by (1 + h^2 r^2)^-2; with
    zdirect an extra last channel holds V[0] convolved by (1 + h^2 r^2)^-1."""
    V = np.ascontiguousarray(V, dtype=np.float32)
    m, ng, d = V.shape[0], V.shape[1], V.ndim - 1
    if m != d + 1:
        raise ValueError("grid_conv needs d+1 channels")
    Phi = np.empty((m + (1 if zdirect else 0),) + V.shape[1:], np.float32)
    _check(lib().sgtsne_grid_conv(d, ng, float(h), int(bool(zdirect)), V.reshape(-1), Phi.reshape(-1)))
    return Phi


def perplexityEqualization(I, D, n, nn, u):
    """Row-stochastic kNN CSR after perplexity equalisation (reference name)."""
    I = np.ascontiguousarray(I, dtype=np.int32).reshape(n, nn + 1)
    D = np.ascontiguousarray(D, dtype=np.float64).reshape(n, nn + 1)
    rowptr = np.empty(n + 1, np.int32)
    col = np.empty(n * nn, np.int32)
    val = np.empty(n * nn, np.float32)
    _check(lib().sgtsne_perplexity_equalize(I, D, int(n), int(nn), float(u), rowptr, col, val))
    return CSR(n, rowptr, col, val)


def sgtsne(P, params, y_in=None, time_info=None):
    """Drop-in for sgtsneCUDA / sgtsne: returns Y in params.layout ([d][N] for
    SoA, [N][d] for AoS).  time_info: optional float64 array of 7*maxIter+1."""
    n = P.n
    d = params.d
    shape = (d, n) if params.layout == LAYOUT_SOA else (n, d)
    y_out = np.empty(shape, np.float32)
    yp = None
    if y_in is not None:
        y_in = np.ascontiguousarray(y_in, dtype=np.float32).reshape(shape)
        yp = y_in.ctypes.data
    tp = None
    if time_info is not None:
        assert time_info.dtype == np.float64 and time_info.size >= 7 * params.maxIter + 1
        tp = time_info.ctypes.data
    _check(lib().sgtsne_run(n, P.nnz, P.rowptr, P.col, P.val, C.byref(params), yp, y_out, tp))
    return y_out


class Context:
    """One embedding job on the device (stepwise API)."""

    def __init__(self, params):
        self.params = params
        self._h = C.c_void_p()
        _check(lib().sgtsne_ctx_create(C.byref(self._h), C.byref(params)))
        self.d = params.d

    def close(self):
        if self._h:
            lib().sgtsne_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_P(self, P, symmetrize=True):
        _check(lib().sgtsne_ctx_upload_P(self._h, P.n, P.nnz, P.rowptr, P.col, P.val, int(bool(symmetrize))))
        self.n = P.n

    def get_P(self):
        nnz = C.c_int64()
        _check(lib().sgtsne_ctx_get_P(self._h, C.byref(nnz), None, None, None))
        rp = np.empty(self.n + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float32)
        _check(lib().sgtsne_ctx_get_P(self._h, C.byref(nnz), rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return CSR(self.n, rp, col, val)

    def set_Y(self, y, layout=LAYOUT_AOS, reset_state=True):
        y = np.ascontiguousarray(y, dtype=np.float32)
        _check(lib().sgtsne_ctx_set_Y(self._h, y.reshape(-1), int(layout), int(bool(reset_state))))

    def get_Y(self, layout=LAYOUT_AOS):
        shape = (self.n, self.d) if layout == LAYOUT_AOS else (self.d, self.n)
        y = np.empty(shape, np.float32)
        _check(lib().sgtsne_ctx_get_Y(self._h, y.reshape(-1), int(layout)))
        return y

    def step(self, iters=1):
        _check(lib().sgtsne_ctx_step(self._h, int(iters)))

    def terms(self, layout=LAYOUT_AOS):
        """(Fattr, Frep, Z, ng) for the current Y (Frep normalised by Z)."""
        shape = (self.n, self.d) if layout == LAYOUT_AOS else (self.d, self.n)
        Fa = np.empty(shape, np.float32)
        Fr = np.empty(shape, np.float32)
        Z = C.c_double()
        ng = C.c_int()
        _check(lib().sgtsne_ctx_terms(self._h, Fa.ctypes.data, Fr.ctypes.data, C.byref(Z), C.byref(ng), int(layout)))
        return Fa, Fr, Z.value, ng.value

    def kl(self, alpha=1.0):
        out = C.c_double()
        _check(lib().sgtsne_ctx_kl(self._h, float(alpha), C.byref(out)))
        return out.value

    def repulsive_exact(self, layout=LAYOUT_AOS):
        shape = (self.n, self.d) if layout == LAYOUT_AOS else (self.d, self.n)
        Fr = np.empty(shape, np.float32)
        Z = C.c_double()
        _check(lib().sgtsne_ctx_repulsive_exact(self._h, Fr.ctypes.data, C.byref(Z), int(layout)))
        return Fr, Z.value

    def timing(self):
        out = np.zeros(7, np.float64)
        _check(lib().sgtsne_ctx_timing(self._h, out))
        return out

    def tuning(self):
        g, u, rpb, win, cnt = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ms = np.zeros(64)
        _check(lib().sgtsne_ctx_tuning(self._h, C.byref(g), C.byref(u), C.byref(rpb), C.byref(win), ms.ctypes.data,
                                       64, C.byref(cnt)))
        return {"group": g.value, "unroll": u.value, "rows_per_block": rpb.value, "window": win.value,
                "candidates_ms": [round(float(x), 4) for x in ms[:cnt.value]]}

    def set_timing(self, on):
        _check(lib().sgtsne_ctx_set_timing(self._h, int(bool(on))))

    def reorders(self):
        a, dcl = C.c_int(), C.c_int()
        _check(lib().sgtsne_ctx_reorders(self._h, C.byref(a), C.byref(dcl)))
        return {"applied": a.value, "declined": dcl.value}

    def set_attractive(self, group, unroll=0, rows_per_block=0, window=0):
        """Pins the attractive kernel shape (CSR_SEG format)."""
        _check(lib().sgtsne_ctx_set_attractive(self._h, int(group), int(unroll), int(rows_per_block), int(window)))

    def errors(self):
        cnt = C.c_int()
        _check(lib().sgtsne_ctx_errors(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(max(cnt.value, 1), np.float64)
        _check(lib().sgtsne_ctx_errors(self._h, out.ctypes.data, cnt.value, C.byref(cnt)))
        return out[:cnt.value]

    def info(self):
        it, n, d, nnz = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        _check(lib().sgtsne_ctx_info(self._h, C.byref(it), C.byref(n), C.byref(d), C.byref(nnz)))
        return {"iter": it.value, "n": n.value, "d": d.value, "nnz": nnz.value}


class Full(SgtsneError):
    """No room in the batch right now (slots, arena or grid budget)."""


class Batch:
    """Many embedding jobs advanced together on the device (sgtsne_batch_*).

    Jobs are admitted from a CSR P (add) or from their points (add_points:
    kNN, perplexity and P + P^T on the device); step() queues iterations for
    every admitted job and returns at once; job_status / get_Y synchronise.
    """

    def __init__(self, params, max_jobs, max_points, max_nnz, grid_budget_bytes=0):
        self.params = params
        self.d = params.d
        self._h = C.c_void_p()
        _check(lib().sgtsne_batch_create(C.byref(self._h), C.byref(params), int(max_jobs), int(max_points),
                                         int(max_nnz), int(grid_budget_bytes)))
        self.n = {}
        self._fronts = {}  # ticket -> (X, y_in, n): arrays the library reads until the admission

    def close(self):
        if self._h:
            lib().sgtsne_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _y_arg(self, y_in, n):
        if y_in is None:
            return None, None
        shape = (self.d, n) if self.params.layout == LAYOUT_SOA else (n, self.d)
        y = np.ascontiguousarray(y_in, dtype=np.float32).reshape(shape)
        return y, y.ctypes.data

    def _admit(self, status, job, n):
        if status == ERR_FULL:
            raise Full(lib().sgtsne_last_error().decode(errors="replace"))
        _check(status)
        self.n[job.value] = n
        return job.value

    def add(self, P, symmetrize=True, y_in=None, max_iter=0, ng_cap=0):
        job = C.c_int(-1)
        keep, yp = self._y_arg(y_in, P.n)
        st = lib().sgtsne_batch_add(self._h, P.n, P.nnz, P.rowptr, P.col, P.val, int(bool(symmetrize)), yp,
                                    int(max_iter), int(ng_cap), C.byref(job))
        return self._admit(st, job, P.n)

    def add_points(self, X, k=90, u=30.0, y_in=None, max_iter=0, ng_cap=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, dim = X.shape
        job = C.c_int(-1)
        keep, yp = self._y_arg(y_in, n)
        st = lib().sgtsne_batch_add_points(self._h, X, n, dim, int(k), float(u), yp, int(max_iter), int(ng_cap),
                                           C.byref(job))
        return self._admit(st, job, n)

    def add_points_async(self, X, k=90, u=30.0, y_in=None, max_iter=0, ng_cap=0):
        """Queues the first part of a job's front end (kNN) and returns a
        ticket at once; poll_points(ticket) takes it to the admission.  Up to
        8 front ends per batch in flight."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, dim = X.shape
        keep, yp = self._y_arg(y_in, n)
        t = C.c_int64(-1)
        _check(lib().sgtsne_batch_add_points_async(self._h, X, n, dim, int(k), float(u), yp, int(max_iter),
                                                   int(ng_cap), C.byref(t)))
        self._fronts[t.value] = (X, keep, n)
        return t.value

    def poll_points(self, ticket, wait=False):
        """The job once the front end of `ticket` has been admitted, else None
        (the device has not passed its queued part yet: the next part is
        queued by a later poll).  Raises Full when the job does not fit."""
        job = C.c_int(-1)
        st = lib().sgtsne_batch_add_poll(self._h, int(ticket), int(bool(wait)), C.byref(job))
        if st != 0:
            self._fronts.pop(ticket, None)  # a failed front end is released (an unknown ticket: nothing)
            return self._admit(st, job, 0)  # raises
        if job.value >= 0:
            _, _, n = self._fronts.pop(ticket)
            return self._admit(st, job, n)
        return None

    def step(self, iters=1):
        _check(lib().sgtsne_batch_step(self._h, int(iters)))

    def sync(self):
        _check(lib().sgtsne_batch_sync(self._h))

    def mark(self):
        """A point on the iteration stream after everything queued so far."""
        m = C.c_int64()
        _check(lib().sgtsne_batch_mark(self._h, C.byref(m)))
        return m.value

    def wait(self, mark):
        """Returns once the device has passed `mark`."""
        _check(lib().sgtsne_batch_wait(self._h, int(mark)))

    def passed(self, mark):
        """True once the device has passed `mark` (does not block)."""
        p = C.c_int()
        _check(lib().sgtsne_batch_passed(self._h, int(mark), C.byref(p)))
        return bool(p.value)

    def job_left(self, job):
        """Batch iterations the job needs before its last one is queued (no
        device access)."""
        v = C.c_int64()
        _check(lib().sgtsne_batch_job_left(self._h, int(job), C.byref(v)))
        return v.value

    def job_status(self, job):
        st, it = C.c_int(), C.c_int()
        _check(lib().sgtsne_batch_job_status(self._h, int(job), C.byref(st), C.byref(it)))
        return st.value, it.value

    def get_Y(self, job, layout=LAYOUT_AOS):
        n = self.n[job]
        y = np.empty((n, self.d) if layout == LAYOUT_AOS else (self.d, n), np.float32)
        _check(lib().sgtsne_batch_get_Y(self._h, int(job), y.reshape(-1), int(layout)))
        return y

    def terms(self, job, layout=LAYOUT_AOS):
        n = self.n[job]
        shape = (n, self.d) if layout == LAYOUT_AOS else (self.d, n)
        Fa = np.empty(shape, np.float32)
        Fr = np.empty(shape, np.float32)
        Z = C.c_double()
        ng = C.c_int()
        _check(lib().sgtsne_batch_terms(self._h, int(job), Fa.ctypes.data, Fr.ctypes.data, C.byref(Z),
                                        C.byref(ng), int(layout)))
        return Fa, Fr, Z.value, ng.value

    def remove(self, job):
        _check(lib().sgtsne_batch_remove(self._h, int(job)))
        self.n.pop(job, None)

    def info(self):
        j, p, z, g, r = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().sgtsne_batch_info(self._h, C.byref(j), C.byref(p), C.byref(z), C.byref(g), C.byref(r)))
        return {"jobs": j.value, "points": p.value, "nnz": z.value, "grid_bytes": g.value, "graph_replays": r.value}

    def tuning(self):
        g, u, rpb, win, gr = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().sgtsne_batch_tuning(self._h, C.byref(g), C.byref(u), C.byref(rpb), C.byref(win), C.byref(gr)))
        return {"group": g.value, "unroll": u.value, "rows_per_block": rpb.value, "window": win.value,
                "graph": bool(gr.value)}

    def set_graph(self, on):
        _check(lib().sgtsne_batch_set_graph(self._h, int(bool(on))))

    def set_attractive(self, unroll, rows_per_block=0, window=0):
        """Pins the attractive-kernel shape (the lanes per row stay the engine's)."""
        _check(lib().sgtsne_batch_set_attractive(self._h, int(unroll), int(rows_per_block), int(window)))
