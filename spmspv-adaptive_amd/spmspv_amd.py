"""ctypes binding of libspmspv_amd.so (the C-ABI in include/spmspv_amd.h).

Host-side mirror of the reference's operator class
``SpmspvHandle<int, unsigned int, float>``
(hice-spmspv/hice/la/include/spmspv/csc-spmspv/spmspv.h:88-344): the class
below keeps the reference's method names and argument meaning, and every
call goes through the C-ABI. Device memory is managed with the library's
own allocation helpers (no torch needed). Errors raise ``SpmspvError``
carrying the library's code and text; the library itself never exits.

A missing library raises at import of the first call: there is no CPU
fallback anywhere in this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspmspv_amd.so")
_lib = None

# kernel slots (include/spmspv_amd.h)
K_COL_NAIVE_ATOMIC = 0
K_COL_NAIVE_SORT = 1
K_COL_LB_ATOMIC = 2
K_COL_LB_SORT = 3
K_ROW_NAIVE = 4
K_ROW_LB = 5
K_SPMV_NAIVE = 6
K_SPMV_LB = 7
K_COL_BUCKET = 8
K_COL_TWO_LAYER_ATOMIC = 9
K_COL_TWO_LAYER_SORT = 10
NUM_SLOTS = 8
# vector types (set_vector_type)
SPMSPV_VEC_SPARSE, SPMSPV_VEC_DENSE, SPMSPV_VEC_BITARRAY = 0, 1, 2
KERNEL_NAMES = ["col-naive-atomic", "col-naive-sort", "col-lb-atomic", "col-lb-sort",
                "row-naive", "row-lb", "spmv-naive", "spmv-lb", "col-bucket", "col-two-layer-atomic",
                "col-two-layer-sort"]
MLP_F, MLP_H1, MLP_H2, MLP_C = 16, 64, 64, 8
MLP_NPARAM = MLP_H1 * MLP_F + MLP_H1 + MLP_H2 * MLP_H1 + MLP_H2 + MLP_C * MLP_H2 + MLP_C
MLP_NWEIGHTS = MLP_NPARAM + 2 * MLP_F

SUCCESS = 0
NOT_IMPLEMENTED = -24


class SpmspvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"spmspv error {code}: {msg}")
        self.code = code


_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_ll = ctypes.c_longlong
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64
_d = ctypes.c_double

_SIGS = {
    "spmspv_create": [ctypes.POINTER(_vp), _i, _i, _i],
    "spmspv_destroy": [_vp],
    "spmspv_set_stream": [_vp, _vp],
    "spmspv_warmup": [_vp],
    "spmspv_input_csr": [_vp, _i, _vp, _vp, _vp],
    "spmspv_input_csc": [_vp, _i, _vp, _vp, _vp],
    "spmspv_to_csc": [_vp],
    "spmspv_alloc_pre_buffer": [_vp],
    "spmspv_dealloc_pre_buffer": [_vp],
    "spmspv_hola_preprocess": [_vp],
    "spmspv_hola_postprocess": [_vp],
    "spmspv_csr5_preprocess": [_vp],
    "spmspv_csr5_postprocess": [_vp],
    "spmspv_set_sigma": [_vp, _i],
    "spmspv_set_spmv_type": [_vp, _i],
    "spmspv_set_vector_type": [_vp, _i],
    "spmspv_set_sparse_x": [_vp, _vp, _vp, _i],
    "spmspv_set_x": [_vp, _vp],
    "spmspv_set_bitvector": [_vp, _vp, _i],
    "spmspv_set_sparse_x_dense_x": [_vp, _vp, _vp, _vp, _i],
    "spmspv_sparse2dense": [_vp],
    "spmspv_sparse2bitarray": [_vp],
    "spmspv_dense2sparse": [_vp],
    "spmspv_dense2bitarray": [_vp],
    "spmspv_bitarray2sparse": [_vp],
    "spmspv_dense2sparse_outer": [_vp, _i, _vp, ctypes.POINTER(_i), _vp, _vp],
    "spmspv_sparse2dense_outer": [_vp, _i, _vp, _vp, _i, _vp],
    "spmspv_compute_binlen": [_vp, ctypes.POINTER(_ll)],
    "spmspv_vec_features": [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "spmspv_mat_features": [_vp, _vp],
    "spmspv_extract_mat_feature": [_vp, ctypes.POINTER(_f)],
    "spmspv_compute_gm": [_vp, _vp],
    "spmspv_features": [_vp, _vp],
    "spmspv_run": [_vp, _i, _f, _vp, _vp, _vp, ctypes.POINTER(_i)],
    "spmspv_csc_nosort_naive": [_vp, _f, ctypes.POINTER(_i), _vp],
    "spmspv_csc_nosort_naive_keyval": [_vp, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_sort_naive": [_vp, _f, ctypes.POINTER(_i), _vp, _vp, _i],
    "spmspv_csc_nosort_merge": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_nosort_merge_keyval": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_sort_merge": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_bucket": [_vp, _f, ctypes.POINTER(_i), _vp, _vp, _i],
    "spmspv_csc_my_nosort": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_my_nosort_keyval": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_csc_sort_my": [_vp, _i, _f, ctypes.POINTER(_i), _vp, _vp, _vp],
    "spmspv_naivespmspv": [_vp, _f, _vp],
    "spmspv_holaspmspv": [_vp, _f, _vp],
    "spmspv_csr5spmspv": [_vp, _f, _vp],
    "spmspv_spmspv": [_vp, _f, _vp],
    "spmspv_naivespmv": [_vp, _f, _vp],
    "spmspv_holaspmv": [_vp, _f, _vp],
    "spmspv_csr5spmv": [_vp, _f, _vp],
    "spmspv_spmv": [_vp, _f, _vp],
    "spmspv_autotune2": [_vp, _i, _i, _vp],
    "spmspv_time_main_kernel": [_vp, _i],
    "spmspv_set_speculation": [_vp, _i],
    "spmspv_set_small_x_path": [_vp, _i],
    "spmspv_set_switch_penalty": [_vp, _f],
    "spmspv_set_bucket_range": [_vp, _ll, _ll],
    "spmspv_bfs_trace": [_vp, _i, _vp, _i, _i, _vp, _i, _vp, _vp],
    "spmspv_bfs_selector_set_weights": [_vp, _vp, _i],
    "spmspv_speculation_stats": [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
    "spmspv_main_kernel_ms": [_vp, _vp],
    "spmspv_autotune": [_vp, _i, _vp],
    "spmspv_autotune_report": [_vp, _vp],
    "spmspv_selector_set_weights": [_vp, _vp, _i],
    "spmspv_selector_mode": [_vp, _i, _i],
    "spmspv_select": [_vp, ctypes.POINTER(_i)],
    "spmspv_run_auto": [_vp, _f, _vp, _vp, _vp, ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "spmspv_mlp_train": [_vp, _vp, _i, _i, _i, _f, ctypes.c_uint, _vp, ctypes.POINTER(_f)],
    "spmspv_mlp_loss_grad": [_vp, _vp, _vp, _i, ctypes.POINTER(_f), _vp],
    "spmspv_mlp_forward": [_vp, _vp, _i, _vp],
    "spmspv_bfs": [_vp, _i, _vp, _i, _i, _vp],
    "spmspv_bfs_multi": [_vp, _i, _vp, _vp, _vp],
    "spmspv_run_batched": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "spmspv_pagerank": [_vp, _f, _f, _i, _i, _vp, _vp, _vp],
    "spmspv_csc_spmspv_noalloc": [_i, _i, _ll, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "spmspv_device_count": [ctypes.POINTER(_i)],
    "spmspv_malloc": [ctypes.POINTER(_vp), _sz],
    "spmspv_free": [_vp],
    "spmspv_memcpy_h2d": [_vp, _vp, _sz],
    "spmspv_memcpy_d2h": [_vp, _vp, _sz],
    "spmspv_memset": [_vp, _i, _sz],
    "spmspv_memcpy_d2d": [_vp, _vp, _sz],
    "spmspv_synchronize": [],
    "spmspv_timer_start": [_vp],
    "spmspv_timer_stop": [_vp, ctypes.POINTER(_f)],
    "spmspv_flush_cache": [_vp, _sz],
    "spmspv_measure_ceilings": [_vp, _i, _vp],
    "spmspv_gen_uniform_csc": [_i, _i, _u64, _vp, _vp, _vp],
    "spmspv_gen_rmat_csr": [_i, _i, _d, _d, _d, _u64, ctypes.POINTER(_ll), _vp, _vp, _vp],
    "spmspv_gen_sparse_vector": [_i, _i, _u64, _f, _f, _vp, _vp],
    "spmspv_gen_banded_csr": [_i, _i, _u64, ctypes.POINTER(_ll), _vp, _vp, _vp],
    "spmspv_gen_blockdiag_csr": [_i, _i, _u64, ctypes.POINTER(_ll), _vp, _vp, _vp],
    "spmspv_gen_random_csr": [_i, _i, _d, _u64, ctypes.POINTER(_ll), _vp, _vp, _vp],
    "spmspv_mtx_read_csr": [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_ll),
                            _vp, _vp, _vp],
    "spmspv_csr_load_header": [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "spmspv_csr_load": [ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp],
    "spmspv_csr_store": [ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp],
    "spmspv_frontier_sizes": [ctypes.c_char_p, ctypes.POINTER(_i), _vp],
    "spmspv_frontier_read": [ctypes.c_char_p, _i, _i, ctypes.POINTER(_i), _vp, _vp],
}


def lib():
    """Loads the built library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpmspvError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i
        L.spmspv_last_error.argtypes = [_vp]
        L.spmspv_last_error.restype = ctypes.c_char_p
        L.spmspv_ynnz_device.argtypes = [_vp]
        L.spmspv_ynnz_device.restype = _vp
        _lib = L
    return _lib


def declared_symbols():
    return sorted(list(_SIGS.keys()) + ["spmspv_last_error", "spmspv_ynnz_device"])


def _check(rc, h=None):
    if rc != SUCCESS:
        msg = lib().spmspv_last_error(h)
        raise SpmspvError(rc, msg.decode() if msg else "")
    return rc


def _np(a):
    return a.ctypes.data_as(_vp)


def device_count():
    c = _i(0)
    rc = lib().spmspv_device_count(ctypes.byref(c))
    return c.value if rc == SUCCESS else 0


def synchronize():
    _check(lib().spmspv_synchronize())


class DeviceArray:
    """A device buffer owned by Python (freed on close / garbage collection)."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(1, self.count * self.dtype.itemsize)
        p = _vp()
        _check(lib().spmspv_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            _check(lib().spmspv_memcpy_h2d(d.ptr, _np(a), a.nbytes))
        return d

    def to_host(self, count=None):
        count = self.count if count is None else int(count)
        out = np.empty(count, self.dtype)
        if count:
            _check(lib().spmspv_memcpy_d2h(_np(out), self.ptr, count * self.dtype.itemsize))
        return out

    def to_host_range(self, offset, count):
        """Elements [offset, offset + count) to the host."""
        out = np.empty(int(count), self.dtype)
        if count:
            src = _vp(self.ptr.value + int(offset) * self.dtype.itemsize)
            _check(lib().spmspv_memcpy_d2h(_np(out), src, int(count) * self.dtype.itemsize))
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        if a.size:
            _check(lib().spmspv_memcpy_h2d(self.ptr, _np(a), a.nbytes))

    def fill_zero(self):
        _check(lib().spmspv_memset(self.ptr, 0, self.nbytes))

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().spmspv_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpmspvHandle:
    """Mirror of SpmspvHandle<int, unsigned int, float> (spmspv.h:88-344)."""

    def __init__(self, m, n, nnz):
        self.h = _vp()
        _check(lib().spmspv_create(ctypes.byref(self.h), m, n, nnz))
        self.m, self.n, self.nnz = m, n, nnz
        # The handle borrows the device arrays it is given (the C-ABI keeps
        # only their addresses): the wrapper holds the arrays, so that they
        # cannot be freed while the handle may still read them.
        self._borrowed = {}

    def _c(self, rc):
        return _check(rc, self.h)

    def Destroy(self):
        if self.h is not None and self.h.value:
            lib().spmspv_destroy(self.h)
        self.h = None
        self._borrowed = {}

    def __del__(self):
        try:
            self.Destroy()
        except Exception:
            pass

    # ---- set-up (spmspv.h:353-419, 1653-1763)
    def InputCSR(self, nnz, row_ptr, col, val):
        rc = self._c(lib().spmspv_input_csr(self.h, nnz, row_ptr.ptr, col.ptr, val.ptr))
        self._borrowed["csr"] = (row_ptr, col, val)
        return rc

    def InputCSC(self, nnz, row, col_ptr, val):
        rc = self._c(lib().spmspv_input_csc(self.h, nnz, row.ptr, col_ptr.ptr, val.ptr))
        self._borrowed["csc"] = (row, col_ptr, val)
        return rc

    def ToCSC(self):
        return self._c(lib().spmspv_to_csc(self.h))

    def allocPreBuffer(self):
        return self._c(lib().spmspv_alloc_pre_buffer(self.h))

    def deallocPreBuffer(self):
        return self._c(lib().spmspv_dealloc_pre_buffer(self.h))

    def holaPreprocess(self):
        return self._c(lib().spmspv_hola_preprocess(self.h))

    def holaPostprocess(self):
        return self._c(lib().spmspv_hola_postprocess(self.h))

    def set_spmv_type(self, code):
        """0: merge-path row tiles (Hola role), 1: CSR5 tiles for K5/K7."""
        return self._c(lib().spmspv_set_spmv_type(self.h, code))

    def set_sigma(self, sigma):
        """CSR5 items per lane (<= 0: from nonzeros per row)."""
        return self._c(lib().spmspv_set_sigma(self.h, sigma))

    def csr5_preprocess(self):
        return self._c(lib().spmspv_csr5_preprocess(self.h))

    def csr5_postprocess(self):
        return self._c(lib().spmspv_csr5_postprocess(self.h))

    # ---- x (spmspv.h:977-1025)
    def set_vector_type(self, code):
        return self._c(lib().spmspv_set_vector_type(self.h, code))

    def set_sparse_x(self, key, val, xnnz):
        rc = self._c(lib().spmspv_set_sparse_x(self.h, key.ptr if key else None,
                                               val.ptr if val else None, xnnz))
        self._borrowed["x"] = (key, val)
        return rc

    def set_sparse_x_ptr(self, key_addr, val_addr, xnnz):
        """set_sparse_x with raw device addresses (slices of larger arrays)."""
        return self._c(lib().spmspv_set_sparse_x(self.h, _vp(key_addr), _vp(val_addr), xnnz))

    def set_x(self, x):
        rc = self._c(lib().spmspv_set_x(self.h, x.ptr if x else None))
        self._borrowed["dense_x"] = x
        return rc

    def set_bitvector(self, bits, num_ints):
        rc = self._c(lib().spmspv_set_bitvector(self.h, bits.ptr if bits else None, num_ints))
        self._borrowed["bits"] = bits
        return rc

    def sparse2dense(self):
        return self._c(lib().spmspv_sparse2dense(self.h))

    def sparse2bitarray(self):
        return self._c(lib().spmspv_sparse2bitarray(self.h))

    def dense2sparse(self):
        return self._c(lib().spmspv_dense2sparse(self.h))

    def dense2bitarray(self):
        return self._c(lib().spmspv_dense2bitarray(self.h))

    def bitarray2sparse(self):
        return self._c(lib().spmspv_bitarray2sparse(self.h))

    def dense2sparse_outer(self, m, dense, key, val):
        c = _i(0)
        self._c(lib().spmspv_dense2sparse_outer(self.h, m, dense.ptr, ctypes.byref(c), key.ptr, val.ptr))
        return c.value

    def sparse2dense_outer(self, ynnz, key, val, m, dense):
        return self._c(lib().spmspv_sparse2dense_outer(self.h, ynnz, key.ptr, val.ptr, m, dense.ptr))

    # ---- features (spmspv.h:2259-2753)
    def computeBinlen(self):
        b = _ll(0)
        self._c(lib().spmspv_compute_binlen(self.h, ctypes.byref(b)))
        return b.value

    computeBinlenVer2 = computeBinlen

    def computeVecFeature(self):
        b, mx, mn = _ll(0), _i(0), _i(0)
        self._c(lib().spmspv_vec_features(self.h, ctypes.byref(b), ctypes.byref(mx), ctypes.byref(mn)))
        return b.value, mx.value, mn.value

    def computeMatFeature(self):
        out = np.zeros(7, np.float32)
        self._c(lib().spmspv_mat_features(self.h, _np(out)))
        return out

    def extractMatFeature(self):
        v = _f(0)
        self._c(lib().spmspv_extract_mat_feature(self.h, ctypes.byref(v)))
        return v.value

    def computeGM(self):
        out = np.zeros(3, np.float64)
        self._c(lib().spmspv_compute_gm(self.h, _np(out)))
        return out

    def features(self):
        out = np.zeros(16, np.float32)
        self._c(lib().spmspv_features(self.h, _np(out)))
        return out

    # ---- kernels
    def run(self, kernel_id, y_dense=None, y_key=None, y_val=None, want_nnz=True):
        c = _i(-1)
        self._c(lib().spmspv_run(self.h, kernel_id, 1.0,
                                 y_dense.ptr if y_dense else None,
                                 y_key.ptr if y_key else None,
                                 y_val.ptr if y_val else None,
                                 ctypes.byref(c) if want_nnz else None))
        return c.value

    def run_auto(self, y_dense=None, y_key=None, y_val=None):
        c, k = _i(-1), _i(-1)
        self._c(lib().spmspv_run_auto(self.h, 1.0, y_dense.ptr if y_dense else None,
                                      y_key.ptr if y_key else None, y_val.ptr if y_val else None,
                                      ctypes.byref(c), ctypes.byref(k)))
        return c.value, k.value

    def run_auto_ptr(self, ykey_addr, yval_addr):
        """run_auto writing the sorted result at raw device addresses."""
        c, k = _i(-1), _i(-1)
        self._c(lib().spmspv_run_auto(self.h, 1.0, None, _vp(ykey_addr), _vp(yval_addr),
                                      ctypes.byref(c), ctypes.byref(k)))
        return c.value, k.value

    def autotune(self, reps=3, csr5=True):
        rep = np.zeros(5, np.int32)
        self._c(lib().spmspv_autotune2(self.h, reps, 1 if csr5 else 0, _np(rep)))
        return self.tuned()

    def tuned(self):
        """Every tuned parameter (spmspv_autotune_report)."""
        r = np.zeros(10, np.int32)
        self._c(lib().spmspv_autotune_report(self.h, _np(r)))
        return {"team_col": int(r[0]), "team_row": int(r[1]), "row_vt": int(r[2]), "spmv_type": int(r[3]),
                "sigma": int(r[4]), "buckets": int(r[5]), "small_x_products": int(r[6]),
                "one_workgroup_sort": int(r[7]), "bucket_rule": [int(r[8]), int(r[9])]}

    def set_bucket_range(self, lo, hi):
        """bin_len range in which auto mode runs the bucket slot in place of
        slot 2, 3, 5 or 7 (hi < lo: off)."""
        return self._c(lib().spmspv_set_bucket_range(self.h, int(lo), int(hi)))

    def select(self):
        k = _i(-1)
        self._c(lib().spmspv_select(self.h, ctypes.byref(k)))
        return k.value

    def selector_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        return self._c(lib().spmspv_selector_set_weights(self.h, _np(w), w.size))

    def selector_mode(self, mode, fixed=0):
        return self._c(lib().spmspv_selector_mode(self.h, mode, fixed))

    def bfs(self, src, levels, policy=-1, use_graphs=0):
        st = np.zeros(4, np.float64)
        self._c(lib().spmspv_bfs(self.h, src, levels.ptr, policy, use_graphs, _np(st)))
        return st

    def bfs_multi(self, srcs, levels):
        """Up to 64 BFSs at once; levels (device, n * len(srcs) ints) gets
        levels[v * nsrc + i]. Returns the stats."""
        srcs = np.ascontiguousarray(srcs, np.int32)
        st = np.zeros(4, np.float64)
        self._c(lib().spmspv_bfs_multi(self.h, int(srcs.size), _np(srcs), levels.ptr, _np(st)))
        return st

    def bfs_trace(self, src, levels, policy=-1, use_graphs=0, max_levels=128):
        """bfs that also returns, per level, its device time (ms) and the
        selector features of its frontier: (stats, level_ms, level_feat)."""
        st = np.zeros(4, np.float64)
        ms = np.zeros(max_levels, np.float32)
        ft = np.zeros((max_levels, MLP_F), np.float32)
        self._c(lib().spmspv_bfs_trace(self.h, src, levels.ptr, policy, use_graphs, _np(st), max_levels,
                                       _np(ms), _np(ft)))
        nl = min(int(st[0]), max_levels)
        return st, ms[:nl], ft[:nl]

    def bfs_selector_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        return self._c(lib().spmspv_bfs_selector_set_weights(self.h, _np(w), w.size))

    def pagerank(self, rank, damping=0.85, eps=1e-4, max_iter=500, policy=-1):
        """Residual-push PageRank (each iteration one SpMSpV); returns
        (stats, x size per iteration)."""
        st = np.zeros(4, np.float64)
        sizes = np.zeros(max_iter, np.int32)
        self._c(lib().spmspv_pagerank(self.h, damping, eps, max_iter, policy, rank.ptr, _np(st), _np(sizes)))
        return st, sizes[: int(st[0])]

    def run_batched(self, vptr, vkey, vval, yoff, ykey, yval, ynnz):
        """Batched SpMSpV over the vectors vptr/vkey/vval (device arrays);
        results at yoff (device int64 offsets) in ykey/yval, counts in ynnz.
        Returns the slot used per vector."""
        nvec = vptr.count - 1
        used = np.zeros(max(nvec, 1), np.int32)
        self._c(lib().spmspv_run_batched(self.h, nvec, vptr.ptr, vkey.ptr if vkey else None,
                                         vval.ptr if vval else None, yoff.ptr, ykey.ptr, yval.ptr,
                                         ynnz.ptr, _np(used)))
        return used[:nvec]

    def set_speculation(self, on=True):
        """Speculative queueing of run_auto (on by default)."""
        return self._c(lib().spmspv_set_speculation(self.h, 1 if on else 0))

    def set_switch_penalty(self, penalty):
        """Switching-aware selection: keep the previous call's kernel family
        within `penalty` logits of the top choice (0: off)."""
        return self._c(lib().spmspv_set_switch_penalty(self.h, penalty))

    def set_small_x_path(self, on=True):
        """Single-launch path of the column slots for a small x (on by default)."""
        return self._c(lib().spmspv_set_small_x_path(self.h, 1 if on else 0))

    def speculation_stats(self):
        """(right guesses, wrong guesses) of run_auto's speculation so far."""
        hits, misses = ctypes.c_longlong(0), ctypes.c_longlong(0)
        self._c(lib().spmspv_speculation_stats(self.h, ctypes.byref(hits), ctypes.byref(misses)))
        return hits.value, misses.value

    def time_main_kernel(self, on=True):
        return self._c(lib().spmspv_time_main_kernel(self.h, 1 if on else 0))

    def main_kernel_ms(self):
        v = _f(0)
        self._c(lib().spmspv_main_kernel_ms(self.h, ctypes.byref(v)))
        return v.value

    def timer_start(self):
        return self._c(lib().spmspv_timer_start(self.h))

    def timer_stop(self):
        v = _f(0)
        self._c(lib().spmspv_timer_stop(self.h, ctypes.byref(v)))
        return v.value

    def measure_ceilings(self, reps=3):
        """{read_gbs, copy_gbs, gather_gps}: the GPU's streaming read, copy
        and random-gather ceilings (each launch cold)."""
        out = np.zeros(3, np.float64)
        self._c(lib().spmspv_measure_ceilings(self.h, reps, _np(out)))
        return {"read_gbs": float(out[0]), "copy_gbs": float(out[1]), "gather_gps": float(out[2])}

    def flush_cache(self, nbytes=512 << 20):
        return self._c(lib().spmspv_flush_cache(self.h, nbytes))


# ---- synthetic inputs (host)
def gen_uniform_csc(n, k, seed):
    cp = np.zeros(n + 1, np.int32)
    r = np.zeros(max(n * k, 1), np.int32)
    v = np.zeros(max(n * k, 1), np.float32)
    _check(lib().spmspv_gen_uniform_csc(n, k, seed, _np(cp), _np(r), _np(v)))
    return cp, r[: n * k], v[: n * k]


def gen_rmat_csr(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    nnz = _ll(0)
    _check(lib().spmspv_gen_rmat_csr(scale, edge_factor, a, b, c, seed, ctypes.byref(nnz), None, None, None))
    n = 1 << scale
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1), np.float32)
    _check(lib().spmspv_gen_rmat_csr(scale, edge_factor, a, b, c, seed, ctypes.byref(nnz), _np(rp), _np(col), _np(val)))
    return rp, col[: nnz.value], val[: nnz.value]


def gen_sparse_vector(n, k, seed, lo=-1.0, hi=1.0):
    key = np.zeros(max(k, 1), np.int32)
    val = np.zeros(max(k, 1), np.float32)
    _check(lib().spmspv_gen_sparse_vector(n, k, seed, lo, hi, _np(key), _np(val)))
    return key[:k], val[:k]


def _gen_two_phase(fn, *args):
    nnz = _ll(0)
    _check(fn(*args, ctypes.byref(nnz), None, None, None))
    return nnz.value


def gen_banded_csr(n, bw, seed):
    nnz = _gen_two_phase(lib().spmspv_gen_banded_csr, n, bw, seed)
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    _check(lib().spmspv_gen_banded_csr(n, bw, seed, ctypes.byref(_ll(0)), _np(rp), _np(col), _np(val)))
    return rp, col[:nnz], val[:nnz]


def gen_blockdiag_csr(n, bs, seed):
    nnz = _gen_two_phase(lib().spmspv_gen_blockdiag_csr, n, bs, seed)
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    _check(lib().spmspv_gen_blockdiag_csr(n, bs, seed, ctypes.byref(_ll(0)), _np(rp), _np(col), _np(val)))
    return rp, col[:nnz], val[:nnz]


def gen_random_csr(m, n, mean, seed):
    nnz = _gen_two_phase(lib().spmspv_gen_random_csr, m, n, mean, seed)
    rp = np.zeros(m + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    _check(lib().spmspv_gen_random_csr(m, n, mean, seed, ctypes.byref(_ll(0)), _np(rp), _np(col), _np(val)))
    return rp, col[:nnz], val[:nnz]


# ---- matrix and frontier files (host)
def _path(p):
    return os.fsencode(p)


def read_mtx(path):
    """MatrixMarket file -> (m, n, row_ptr, col, val) with the reference's
    ReadMTX semantics (mtx.hpp:218-402)."""
    m, n, nnz = _i(0), _i(0), _ll(0)
    _check(lib().spmspv_mtx_read_csr(_path(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                                     None, None, None))
    rp = np.zeros(m.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1), np.float32)
    _check(lib().spmspv_mtx_read_csr(_path(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                                     _np(rp), _np(col), _np(val)))
    return m.value, n.value, rp, col[: nnz.value], val[: nnz.value]


def store_csr(path, m, n, row_ptr, col, val):
    """storeCSR (mtx.hpp:806-819): the binary CSR cache file."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float32)
    _check(lib().spmspv_csr_store(_path(path), m, n, int(row_ptr[-1]), _np(row_ptr),
                                  _np(col) if col.size else None, _np(val) if val.size else None))


def load_csr(path):
    """loadCSR_header + loadCSR (mtx.hpp:750-804) -> (m, n, row_ptr, col, val)."""
    m, n, nnz = _i(0), _i(0), _i(0)
    _check(lib().spmspv_csr_load_header(_path(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz)))
    rp = np.zeros(m.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1), np.float32)
    _check(lib().spmspv_csr_load(_path(path), m.value, n.value, nnz.value, _np(rp), _np(col), _np(val)))
    return m.value, n.value, rp, col[: nnz.value], val[: nnz.value]


def frontier_sizes(path):
    """readNNZXFromFile + readSparseXInxFromFile (readsvmdata.hpp:223-258)."""
    c = _i(0)
    _check(lib().spmspv_frontier_sizes(_path(path), ctypes.byref(c), None))
    out = np.zeros(max(c.value, 1), np.int32)
    _check(lib().spmspv_frontier_sizes(_path(path), ctypes.byref(c), _np(out)))
    return out[: c.value]


def frontier_read(path, line, cap):
    """extractSparseXfromFile (readsvmdata.hpp:308-348): (keys, values=1.0)."""
    k = np.zeros(max(cap, 1), np.int32)
    v = np.zeros(max(cap, 1), np.float32)
    c = _i(0)
    _check(lib().spmspv_frontier_read(_path(path), line, cap, ctypes.byref(c), _np(k), _np(v)))
    return k[: c.value], v[: c.value]


def replay_frontiers(A, sizes_path, x_path, y_key, y_val, slots=None):
    """The reference harness's frontier replay (la/test/test_application.cu:
    240-415, new_test_app.cu:280-469): iteration i takes x = the keys on
    line i of x_path (values 1.0; at most the i-th size of sizes_path,
    clamped to n), sets it as the handle's sparse x and runs the selector's
    slot (slots None) or each given slot, timed with the handle's events.
    Returns one dict per iteration: xnnz, keys, and per run (slot, ynnz, ms)."""
    sizes = frontier_sizes(sizes_path)
    out = []
    for i, want in enumerate(sizes):
        want = min(int(want), A.n)
        keys, vals = frontier_read(x_path, i, A.n)
        keys, vals = keys[:want], vals[:want]
        dk = DeviceArray.from_host(keys if keys.size else np.zeros(1, np.int32), np.int32)
        dv = DeviceArray.from_host(vals if vals.size else np.zeros(1, np.float32), np.float32)
        runs = []
        for slot in (slots if slots is not None else [None]):
            A.set_sparse_x(dk, dv, int(keys.size))
            A.timer_start()
            if slot is None:
                ynnz, used = A.run_auto(y_key=y_key, y_val=y_val)
            else:
                ynnz, used = A.run(slot, y_key=y_key, y_val=y_val), slot
            ms = A.timer_stop()
            runs.append({"slot": used, "ynnz": ynnz, "ms": ms})
        out.append({"xnnz": int(keys.size), "keys": keys, "runs": runs})
    return out


class Matrix:
    """A CSR matrix resident on the device plus a handle set up the way the
    reference harness does it (InputCSR, holaPreprocess, ToCSC,
    allocPreBuffer; la/test/new_test_app.cu:180-255)."""

    def __init__(self, m, n, row_ptr, col, val, to_csc=True):
        self.m, self.n = m, n
        self.nnz = int(row_ptr[-1])
        self.h_row_ptr, self.h_col, self.h_val = row_ptr, col, val
        self.d_row_ptr = DeviceArray.from_host(row_ptr, np.int32)
        self.d_col = DeviceArray.from_host(col if self.nnz else np.zeros(1, np.int32), np.int32)
        self.d_val = DeviceArray.from_host(val if self.nnz else np.zeros(1, np.float32), np.float32)
        self.A = SpmspvHandle(m, n, self.nnz)
        self.A.InputCSR(self.nnz, self.d_row_ptr, self.d_col, self.d_val)
        if to_csc:
            self.A.ToCSC()
        self.A.allocPreBuffer()
