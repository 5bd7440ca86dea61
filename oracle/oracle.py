"""numpy/ctypes front-end of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, as the checker or the timed CPU baseline; the
product path never does. Each wrapper names the reference function it
restates (see oracle.h for file:line citations).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c_int = ctypes.c_int


def build():
    """Compile liboracle.so with gcc (no GPU, no reference sources)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_csr_to_csc.argtypes = [_c_int, _c_int, _c_int, _i32p, _i32p, _f32p, _i32p, _i32p, _f32p]
        L.oracle_serialspmv_f32.argtypes = [_c_int, _i32p, _i32p, _f32p, _f32p, ctypes.c_float, _f32p]
        L.oracle_spmspv_csc_host_f32.argtypes = [_c_int, _i32p, _i32p, _f32p, _c_int, _i32p, _f32p, _f32p, _i32p, _f32p]
        L.oracle_spmspv_csc_host_f32.restype = _c_int
        L.oracle_spmspv_f64.argtypes = [_c_int, _i32p, _i32p, _f32p, _c_int, _i32p, _f32p, _f64p, _f64p, _i32p, _f64p, _f64p]
        L.oracle_spmspv_f64.restype = _c_int
        L.oracle_spmv_f64.argtypes = [_c_int, _i32p, _i32p, _f32p, _f32p, _f64p, _f64p]
        L.oracle_bfs.argtypes = [_c_int, _i32p, _i32p, _c_int, _i32p, _i32p]
        L.oracle_bfs.restype = _c_int
        L.oracle_mat_features.argtypes = [_c_int, _c_int, _c_int, _i32p, _i32p, _f64p]
        L.oracle_std_row_trunc.argtypes = [_c_int, _c_int, _i32p]
        L.oracle_std_row_trunc.restype = ctypes.c_double
        L.oracle_vec_features.argtypes = [_i32p, _c_int, _i32p, _i64p]
        L.oracle_gm.argtypes = [ctypes.c_int64] * 4 + [_f64p]
        L.oracle_scan_col_lengths.argtypes = [_i32p, _c_int, _i32p, _i64p]
        L.oracle_scan_col_lengths.restype = ctypes.c_int64
        L.oracle_mlp_loss_grad.argtypes = [_c_int] * 5 + [_f64p, _f64p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_mlp_loss_grad.restype = ctypes.c_double
        L.oracle_spmspv_csc_host_f32_omp.argtypes = [_c_int, _i32p, _i32p, _f32p, _c_int, _i32p, _f32p, _f32p,
                                                     _i32p, _f32p, _c_int]
        L.oracle_spmspv_csc_host_f32_omp.restype = _c_int
        L.oracle_pagerank_f64.argtypes = [_c_int, _i32p, _i32p, _f32p, ctypes.c_double, ctypes.c_double, _c_int,
                                          _f64p, ctypes.c_void_p]
        L.oracle_pagerank_f64.restype = _c_int
        L.oracle_mlp_forward.argtypes = [_c_int] * 5 + [_f64p, _f64p, _f64p]
        _LIB = L
    return _LIB


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def csr_to_csc(m, n, row_ptr, col, val):
    """ToCSC semantics (cusparseScsr2csc, spmspv.h:383-419)."""
    row_ptr, col, val = _i32(row_ptr), _i32(col), _f32(val)
    nnz = int(row_ptr[-1])
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(max(nnz, 1), np.int32)
    cv = np.zeros(max(nnz, 1), np.float32)
    lib().oracle_csr_to_csc(m, n, nnz, row_ptr, _i32(col if nnz else np.zeros(1)), _f32(val if nnz else np.zeros(1)), cp, ri, cv)
    return cp, ri[:nnz], cv[:nnz]


def serialspmv_f32(m, row_ptr, col, val, x_dense, alpha=1.0):
    """serialspmv, la/test/test_application.cu:107-116 (FP32)."""
    y = np.zeros(m, np.float32)
    col = _i32(col) if len(col) else np.zeros(1, np.int32)
    val = _f32(val) if len(val) else np.zeros(1, np.float32)
    lib().oracle_serialspmv_f32(m, _i32(row_ptr), col, val, _f32(x_dense), alpha, y)
    return y


def spmspv_csc_host_f32(m, col_ptr, row_idx, cval, x_key, x_val):
    """SpmSpvCSCHost, la/test/test_sort_spmspv.cu:329-377 (FP32, no int x)."""
    yd = np.zeros(m, np.float32)
    yk = np.zeros(max(m, 1), np.int32)
    yv = np.zeros(max(m, 1), np.float32)
    ri = _i32(row_idx) if len(row_idx) else np.zeros(1, np.int32)
    cv = _f32(cval) if len(cval) else np.zeros(1, np.float32)
    xk = _i32(x_key) if len(x_key) else np.zeros(1, np.int32)
    xv = _f32(x_val) if len(x_val) else np.zeros(1, np.float32)
    k = lib().oracle_spmspv_csc_host_f32(m, _i32(col_ptr), ri, cv, len(x_key), xk, xv, yd, yk, yv)
    return yd, yk[:k], yv[:k]


def spmspv_csc_host_f32_omp(m, col_ptr, row_idx, cval, x_key, x_val, threads):
    """The build's all-core port of SpmSpvCSCHost (row ranges per thread):
    the same outputs as spmspv_csc_host_f32, bit for bit."""
    yd = np.zeros(m, np.float32)
    yk = np.zeros(max(m, 1), np.int32)
    yv = np.zeros(max(m, 1), np.float32)
    ri = _i32(row_idx) if len(row_idx) else np.zeros(1, np.int32)
    cv = _f32(cval) if len(cval) else np.zeros(1, np.float32)
    xk = _i32(x_key) if len(x_key) else np.zeros(1, np.int32)
    xv = _f32(x_val) if len(x_val) else np.zeros(1, np.float32)
    k = lib().oracle_spmspv_csc_host_f32_omp(m, _i32(col_ptr), ri, cv, len(x_key), xk, xv, yd, yk, yv, threads)
    return yd, yk[:k], yv[:k]


def spmspv_f64(m, col_ptr, row_idx, cval, x_key, x_val):
    """FP64 SpMSpV with structural support, the results are checked against. Returns (keys, vals, bounds)."""
    y64 = np.zeros(max(m, 1), np.float64)
    bnd = np.zeros(max(m, 1), np.float64)
    yk = np.zeros(max(m, 1), np.int32)
    yv = np.zeros(max(m, 1), np.float64)
    yb = np.zeros(max(m, 1), np.float64)
    ri = _i32(row_idx) if len(row_idx) else np.zeros(1, np.int32)
    cv = _f32(cval) if len(cval) else np.zeros(1, np.float32)
    xk = _i32(x_key) if len(x_key) else np.zeros(1, np.int32)
    xv = _f32(x_val) if len(x_val) else np.zeros(1, np.float32)
    k = lib().oracle_spmspv_f64(m, _i32(col_ptr), ri, cv, len(x_key), xk, xv, y64, bnd, yk, yv, yb)
    return yk[:k].copy(), yv[:k].copy(), yb[:k].copy()


def spmv_f64(m, row_ptr, col, val, x_dense):
    """FP64 dense-x SpMV the results are checked against. Returns (y, bound)."""
    y = np.zeros(max(m, 1), np.float64)
    b = np.zeros(max(m, 1), np.float64)
    col = _i32(col) if len(col) else np.zeros(1, np.int32)
    val = _f32(val) if len(val) else np.zeros(1, np.float32)
    lib().oracle_spmv_f64(m, _i32(row_ptr), col, val, _f32(x_dense), y, b)
    return y[:m], b[:m]


def bfs(n, col_ptr, row_idx, src):
    """Queue BFS over the CSC pattern; returns int32 levels (-1 unreached)."""
    lev = np.zeros(max(n, 1), np.int32)
    q = np.zeros(max(n, 1), np.int32)
    ri = _i32(row_idx) if len(row_idx) else np.zeros(1, np.int32)
    nl = lib().oracle_bfs(n, _i32(col_ptr), ri, src, lev, q)
    return lev[:n].copy(), nl


def mat_features(m, n, row_ptr):
    """computeMatFeture_serial, spmspv.h:2274-2334 ->
    (max, min, avg, range, std, edge_equality, gini)."""
    out = np.zeros(7, np.float64)
    row_ptr = _i32(row_ptr)
    lib().oracle_mat_features(m, n, int(row_ptr[-1]), row_ptr, np.zeros(m, np.int32), out)
    return out


def std_row_trunc(m, row_ptr):
    """extractMatFeature stdRow with the int-truncated mean (spmspv.h:929)."""
    row_ptr = _i32(row_ptr)
    return lib().oracle_std_row_trunc(m, int(row_ptr[-1]), row_ptr)


def vec_features(col_ptr, x_key):
    """computeVecFeture_serial, spmspv.h:2337-2361 -> (bin_len, max, min)."""
    out = np.zeros(3, np.int64)
    xk = _i32(x_key) if len(x_key) else np.zeros(1, np.int32)
    lib().oracle_vec_features(_i32(col_ptr), len(x_key), xk, out)
    return out


def gm(m, nnz, xnnz, bin_len):
    """computeGM1/2/3, spmspv.h:2724-2753."""
    out = np.zeros(3, np.float64)
    lib().oracle_gm(m, nnz, xnnz, bin_len, out)
    return out


def scan_col_lengths(col_ptr, x_key):
    """ScanFuseDriver result: ptr[xnnz+1] (exclusive scan of column lengths)."""
    ptr = np.zeros(len(x_key) + 1, np.int64)
    xk = _i32(x_key) if len(x_key) else np.zeros(1, np.int32)
    lib().oracle_scan_col_lengths(_i32(col_ptr), len(x_key), xk, ptr)
    return ptr


def pagerank_f64(n, col_ptr, row_idx, cval, damping=0.85, eps=1e-4, max_iter=500):
    """Residual-push PageRank in FP64: (rank, x size per iteration)."""
    rank = np.zeros(n, np.float64)
    sizes = np.zeros(max_iter, np.int32)
    ri = _i32(row_idx) if len(row_idx) else np.zeros(1, np.int32)
    cv = _f32(cval) if len(cval) else np.zeros(1, np.float32)
    it = lib().oracle_pagerank_f64(n, _i32(col_ptr), ri, cv, damping, eps, max_iter, rank,
                                   sizes.ctypes.data)
    return rank, sizes[:it]


def mlp_param_count(F, H1, H2, C):
    return H1 * F + H1 + H2 * H1 + H2 + C * H2 + C


def mlp_loss_grad(F, H1, H2, C, params, X, labels=None, soft=None, want_grad=True):
    """FP64 selector MLP loss (mean softmax cross-entropy) and gradient."""
    params = _f64(params)
    X = _f64(X)
    B = X.shape[0]
    grad = np.zeros(params.size, np.float64) if want_grad else None
    lab = _i32(labels) if labels is not None else None
    sft = _f64(soft) if soft is not None else None
    loss = lib().oracle_mlp_loss_grad(
        F, H1, H2, C, B, params, X,
        lab.ctypes.data if lab is not None else None,
        sft.ctypes.data if sft is not None else None,
        grad.ctypes.data if grad is not None else None)
    return loss, grad


def mlp_forward(F, H1, H2, C, params, X):
    X = _f64(X)
    out = np.zeros((X.shape[0], C), np.float64)
    lib().oracle_mlp_forward(F, H1, H2, C, X.shape[0], _f64(params), X, out)
    return out
