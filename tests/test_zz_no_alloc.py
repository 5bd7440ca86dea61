"""The no-allocation entry point spmspv_csc_spmspv_noalloc against the
handle path it is a variant of (spmspv_run with SPMSPV_K_COL_LB_ATOMIC on a
handle given the same CSC), on two calls in a row, the second with a
smaller input; both are also checked against the FP64 oracle. Index sets
bit-exact; values within 1e-5 * sum|a_ij x_j| of the oracle, so the two
paths (each with FP32 atomics in arbitrary order) agree within twice that."""
import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5


def noalloc(m, n, nnz, cp, ri, cv, xk, xv, xnnz, yk, yv, cnt):
    return sa.lib().spmspv_csc_spmspv_noalloc(m, n, nnz, cp.ptr, ri.ptr, cv.ptr,
                                              xk.ptr if xk else None, xv.ptr if xv else None, xnnz,
                                              yk.ptr, yv.ptr, cnt.ptr, None)


def one_case(n, k, xnnz, seed):
    print(f"noalloc case n={n} k={k} xnnz={xnnz} seed={seed}", flush=True)
    cp, ri, cv = sa.gen_uniform_csc(n, k, seed)
    nnz = int(cp[-1])
    key, val = sa.gen_sparse_vector(n, xnnz, seed + 1)
    d = [sa.DeviceArray.from_host(a, t) for a, t in ((cp, np.int32), (ri, np.int32), (cv, np.float32))]
    dk = sa.DeviceArray.from_host(key if xnnz else np.zeros(1, np.int32), np.int32)
    dv = sa.DeviceArray.from_host(val if xnnz else np.zeros(1, np.float32), np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    cnt = sa.DeviceArray.from_host(np.array([-1], np.int32), np.int32)
    rc = noalloc(n, n, nnz, d[0], d[1], d[2], dk, dv, xnnz, yk, yv, cnt)
    assert rc == 0, sa.lib().spmspv_last_error(None)
    sa.synchronize()
    c = int(cnt.to_host()[0])
    gk, gv = yk.to_host(c), yv.to_host(c)
    # the existing entry point on the same CSC and x
    A = sa.SpmspvHandle(n, n, nnz)
    A.InputCSC(nnz, d[1], d[0], d[2])
    A.allocPreBuffer()
    A.set_sparse_x(dk, dv, xnnz)
    hk = sa.DeviceArray(n, np.int32)
    hv = sa.DeviceArray(n, np.float32)
    hc = A.run(sa.K_COL_LB_ATOMIC, y_key=hk, y_val=hv)
    rk, rv = hk.to_host(hc), hv.to_host(hc)
    A.Destroy()
    assert c == hc
    np.testing.assert_array_equal(gk, rk)
    ek, ev, eb = oracle.spmspv_f64(n, cp, ri, cv, key, val)
    np.testing.assert_array_equal(gk, ek)
    lim = TOL * eb + 1e-30
    assert np.all(np.abs(gv.astype(np.float64) - ev) <= lim)
    assert np.all(np.abs(rv.astype(np.float64) - ev) <= lim)
    assert np.all(np.abs(gv.astype(np.float64) - rv.astype(np.float64)) <= 2 * lim)
    return c


def test_noalloc_matches_handle_path_twice(gpu):
    # the second call is smaller: the scratch still holds the first call's
    # (larger) state and must not leak into the result
    c1 = one_case(1 << 17, 16, 20000, 11)
    c2 = one_case(1 << 11, 8, 300, 12)
    assert c1 > c2 > 0
    # an empty x after them
    assert one_case(1 << 10, 4, 0, 13) == 0


def test_noalloc_rejects_oversized_input(gpu):
    big = 1 << 24  # accumulator alone 64 MiB > the 32 MiB scratch
    dummy = sa.DeviceArray(4, np.int32)
    rc = sa.lib().spmspv_csc_spmspv_noalloc(big, big, 0, dummy.ptr, None, None, None, None, 0,
                                            dummy.ptr, dummy.ptr, dummy.ptr, None)
    assert rc == -20
    assert b"maximum" in sa.lib().spmspv_last_error(None)
