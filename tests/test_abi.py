"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/spmspv_amd.h declares, rejects bad arguments without a GPU, and its
host generators produce the documented shapes deterministically."""
import os
import re
import subprocess

import numpy as np
import pytest

import spmspv_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spmspv_amd.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spmspv_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.check_output(["nm", "-D", "--defined-only", sa.LIB_PATH]).decode()
    exported = set(re.findall(r" T (spmspv_\w+)", out))
    declared = header_symbols()
    assert len(declared) > 60
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # the Python binding describes exactly the declared surface
    assert set(sa.declared_symbols()) == set(declared)


def test_library_loads_and_binds():
    L = sa.lib()
    for s in header_symbols():
        assert hasattr(L, s)


def test_null_handle_arguments_are_errors():
    import ctypes
    L = sa.lib()
    assert L.spmspv_create(None, 4, 4, 4) == -20
    assert L.spmspv_set_sparse_x(None, None, None, 0) == -20
    assert L.spmspv_input_csr(None, 0, None, None, None) == -20
    assert L.spmspv_run(None, 0, 1.0, None, None, None, None) == -20
    assert L.spmspv_selector_mode(None, 0, 0) == -20
    assert b"null" in L.spmspv_last_error(None)


def test_gen_uniform_csc_shape():
    cp, r, v = sa.gen_uniform_csc(1000, 16, 0x5EED0001)
    assert cp[-1] == 16000
    np.testing.assert_array_equal(np.diff(cp), 16)
    for j in (0, 1, 500, 999):
        col = r[cp[j]:cp[j + 1]]
        assert np.all(np.diff(col) > 0) and col.min() >= 0 and col.max() < 1000
    assert v.min() >= -1 and v.max() <= 1
    cp2, r2, v2 = sa.gen_uniform_csc(1000, 16, 0x5EED0001)
    np.testing.assert_array_equal(r, r2)
    np.testing.assert_array_equal(v, v2)


def test_gen_rmat_symmetric_dedup_no_loops():
    rp, col, val = sa.gen_rmat_csr(10, 8, seed=3)
    n = 1 << 10
    import scipy.sparse as sp
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    assert (A != A.T).nnz == 0
    assert A.diagonal().sum() == 0
    for i in range(0, n, 97):
        row = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
    assert val.min() > 0 and val.max() <= 1
    rp2, col2, _ = sa.gen_rmat_csr(10, 8, seed=3)
    np.testing.assert_array_equal(col, col2)
    # skew: R-MAT degrees are far from uniform
    deg = np.diff(rp)
    assert deg.max() > 8 * deg.mean()


def test_gen_sparse_vector_and_others():
    k, v = sa.gen_sparse_vector(100, 30, 9)
    assert len(np.unique(k)) == 30 and np.all(np.diff(k) > 0)
    k, v = sa.gen_sparse_vector(100, 100, 9)
    np.testing.assert_array_equal(k, np.arange(100))
    rp, col, val = sa.gen_banded_csr(50, 3, 1)
    assert rp[-1] == sum(min(49, i + 3) - max(0, i - 3) + 1 for i in range(50))
    rp, col, val = sa.gen_blockdiag_csr(70, 32, 1)
    assert rp[-1] == 32 * 32 * 2 + 6 * 6
    rp, col, val = sa.gen_random_csr(300, 200, 8.0, 4)
    assert abs(rp[-1] / 300 - 8.0) < 1.5


def test_noalloc_rejects_oversized_input_before_any_device_call():
    """The no-allocation entry point checks its sizes against the 32 MiB
    scratch before it touches the device (so this runs without a GPU)."""
    import ctypes
    L = sa.lib()
    dummy = (ctypes.c_int * 4)()
    big = 1 << 24  # the accumulator alone would need 64 MiB
    p = ctypes.cast(dummy, ctypes.c_void_p)
    assert L.spmspv_csc_spmspv_noalloc(big, big, 0, p, None, None, None, None, 0, p, p, p, None) == -20
    assert b"maximum" in L.spmspv_last_error(None)
    assert L.spmspv_csc_spmspv_noalloc(0, 4, 0, p, None, None, None, None, 0, p, p, p, None) == -20
    assert L.spmspv_csc_spmspv_noalloc(4, 4, 0, None, None, None, None, None, 0, p, p, p, None) == -20
