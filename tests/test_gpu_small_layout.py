"""The single-launch small-x path and the one-workgroup sort size their items
per thread from the product count (ceil(products / threads)): product counts
on both sides of every multiple of the thread count, at the single-launch
cap (11,264) and one x entry past it, and at the one-workgroup sort's cap
(3,072) and past it, against the FP64 oracle through the C-ABI.

A uniform matrix with exactly 16 rows per column makes the product count 16
times the x entries. Index sets bit-exact; FP32 values within
|y - y64| <= 1e-5 * sum_j |a_ij x_j|."""
import sys

import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5
N, K = 8192, 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def uniform16(gpu):
    import scipy.sparse as sp
    cp, ri, cv = sa.gen_uniform_csc(N, K, 0x5EED0031)
    R = sp.csc_matrix((cv, ri, cp), shape=(N, N)).tocsr()
    R.sort_indices()
    M = sa.Matrix(N, N, R.indptr.astype(np.int32), R.indices.astype(np.int32), R.data.astype(np.float32))
    return M, (cp, ri, cv)


def check(cp, ri, cv, key, val, yk, yv, tag):
    ek, ev, eb = oracle.spmspv_f64(N, cp, ri, cv, key, val)
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    err = np.abs(yv.astype(np.float64) - ev)
    bad = np.nonzero(err > TOL * eb + 1e-30)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound"


def x_of(xnnz, seed):
    rng = np.random.default_rng(seed)
    key = np.sort(rng.choice(N, xnnz, replace=False)).astype(np.int32)
    val = rng.uniform(-1, 1, xnnz).astype(np.float32)
    return key, val


@pytest.mark.parametrize("xnnz", [1, 63, 64, 65, 127, 128, 129, 320, 703, 704, 705])
def test_single_launch_product_counts(uniform16, xnnz):
    """run_auto: up to 704 entries (11,264 products) the single launch
    finishes the call; 705 entries go on to the selected slot."""
    M, (cp, ri, cv) = uniform16
    A = M.A
    key, val = x_of(xnnz, 7000 + xnnz)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(N, np.int32)
    yv = sa.DeviceArray(N, np.float32)
    for rep in range(2):
        A.set_sparse_x(dk, dv, xnnz)
        say(f"single launch: xnnz {xnnz} products {K * xnnz} call {rep}")
        c, used = A.run_auto(y_key=yk, y_val=yv)
        check(cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"auto xnnz {xnnz} call {rep} slot {used}")


@pytest.mark.parametrize("xnnz", [1, 15, 16, 17, 100, 191, 192, 193])
@pytest.mark.parametrize("slot", [sa.K_COL_NAIVE_SORT, sa.K_COL_LB_SORT])
def test_one_workgroup_sort_counts(uniform16, xnnz, slot):
    """The sort slots: up to 192 entries (3,072 products) one workgroup
    sorts and reduces; 193 entries take the multi-launch passes."""
    M, (cp, ri, cv) = uniform16
    A = M.A
    key, val = x_of(xnnz, 8000 + xnnz)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(N, np.int32)
    yv = sa.DeviceArray(N, np.float32)
    A.set_sparse_x(dk, dv, xnnz)
    say(f"sort slot {slot}: xnnz {xnnz} products {K * xnnz}")
    c = A.run(slot, y_key=yk, y_val=yv)
    check(cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"slot {slot} xnnz {xnnz}")
