"""CSR5 copies (K5/K7 with spmv_type 1) against the FP64 oracle through the
C-ABI, on column distributions that move the share of items whose x entry
the tile kernel reads from its LDS copy of the 40,960 densest columns: over
99 % (a dense block and a thin tail), about half (81,920 columns of equal
weight) and under a tenth (2^19 columns of equal weight, one row with items
everywhere that spans tiles); empty rows; every sigma; sparse and dense
output.

Index sets bit-exact; FP32 values within |y - y64| <= 1e-5 * sum_j |a_ij x_j|."""
import sys

import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(m, cp, ri, cv, key, val, yk, yv, tag):
    ek, ev, eb = oracle.spmspv_f64(m, cp, ri, cv, key, val)
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    err = np.abs(yv.astype(np.float64) - ev)
    bad = np.nonzero(err > TOL * eb + 1e-30)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound, first rows {ek[bad[:5]]}"


def csr_from(m, n, rows, cols, rng):
    import scipy.sparse as sp
    vals = rng.uniform(-1, 1, rows.size).astype(np.float32)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    A.data[A.data == 0] = 0.5
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32))


def run_csr5(M, slot, key, val, dense=False):
    dk = sa.DeviceArray.from_host(key if key.size else np.zeros(1, np.int32), np.int32)
    dv = sa.DeviceArray.from_host(val if val.size else np.zeros(1, np.float32), np.float32)
    M.A.set_sparse_x(dk, dv, int(key.size))
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    c = M.A.run(slot, y_key=yk, y_val=yv)
    out = (yk.to_host(c), yv.to_host(c))
    if dense:
        yd = sa.DeviceArray(M.m, np.float32)
        M.A.set_sparse_x(dk, dv, int(key.size))
        M.A.run(slot, y_dense=yd, want_nnz=False)
        out = out + (yd.to_host(M.m),)
    return out


def cases():
    rng = np.random.default_rng(77)
    out = []
    # 40,000 dense columns and a thin tail: the 40,960 densest columns hold
    # over 99 % of the items; a quarter of the rows empty
    m, n = 30000, 140000
    rows = rng.integers(0, m, 250000)
    rows = rows[rows % 4 != 1]
    cols = np.where(rng.random(rows.size) < 0.98, rng.integers(0, 40000, rows.size),
                    rng.integers(0, n, rows.size))
    out.append(("dense block + tail", m, n) + csr_from(m, n, rows, cols, rng))
    # 81,920 columns of equal weight: about half the items in LDS
    m, n = 20000, 81920
    rows = rng.integers(0, m, 300000)
    cols = rng.integers(0, n, rows.size)
    out.append(("81,920 even columns", m, n) + csr_from(m, n, rows, cols, rng))
    # 2^19 columns of equal weight: under a tenth in LDS; one row with
    # items in every column range, long enough to span tiles
    m, n = 50000, 1 << 19
    rows = np.concatenate([rng.integers(0, m, 600000), np.full(9000, 4321)])
    cols = np.concatenate([rng.integers(0, n, 600000), rng.choice(n, 9000, replace=False)])
    out.append(("2^19 even columns", m, n) + csr_from(m, n, rows, cols, rng))
    return out


@pytest.fixture(scope="module", params=range(3), ids=["dense_block_tail", "even_81920", "even_2p19"])
def shape_case(request, gpu):
    name, m, n, rp, col, val = cases()[request.param]
    M = sa.Matrix(m, n, rp, col, val)
    yield name, M, oracle.csr_to_csc(m, n, rp, col, val)
    M.A.set_spmv_type(0)
    M.A.set_sigma(-1)


@pytest.mark.parametrize("sigma", [8, 16, 24, 32])
def test_csr5_shapes(shape_case, sigma):
    name, M, (cp, ri, cv) = shape_case
    M.A.set_sigma(sigma)
    M.A.set_spmv_type(1)
    for k in (1, 3000, M.n // 2, M.n):
        key, val = sa.gen_sparse_vector(M.n, k, 5000 + k + sigma)
        for slot in (sa.K_ROW_LB, sa.K_SPMV_LB):
            say(f"{name}: sigma {sigma} xnnz {k} slot {slot}")
            res = run_csr5(M, slot, key, val, dense=(k == M.n // 2))
            yk, yv = res[0], res[1]
            check(M.m, cp, ri, cv, key, val, yk, yv, f"{name} s={sigma} k={k} slot {slot}")
            if k == M.n // 2:
                yd = res[2]
                dense = np.zeros(M.m, np.float32)
                dense[yk] = yv
                np.testing.assert_array_equal(yd, dense)


def test_csr5_repeat_calls_keep_state(shape_case):
    """Alternating densities: the sums, carries and support bytes of one
    call must not leak into the next."""
    name, M, (cp, ri, cv) = shape_case
    M.A.set_sigma(24)
    M.A.set_spmv_type(1)
    for it, k in enumerate((M.n, 5, M.n // 3, 1, M.n)):
        key, val = sa.gen_sparse_vector(M.n, k, 900 + it)
        say(f"{name}: repeat {it} xnnz {k}")
        yk, yv = run_csr5(M, sa.K_ROW_LB, key, val)
        check(M.m, cp, ri, cv, key, val, yk, yv, f"{name} repeat {it}")
