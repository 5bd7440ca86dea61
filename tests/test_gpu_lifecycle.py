"""Handle life cycle and vector-format conversions through the C-ABI:
a new matrix on a handle that transposed the previous one, BFS level
graphs across matrix and arena changes, the bucket count argument, and
every x conversion (dense2sparse, dense2bitarray, bitarray2sparse,
sparse2bitarray, sparse2dense, the outer forms) checked against the CPU
oracle; dense x as input to every slot.

Index sets bit-exact, values |y - y64| <= 1e-5 * sum|a_ij x_j|, BFS
levels bit-exact."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse as sp

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
    bad = np.nonzero(np.abs(yv.astype(np.float64) - ev) > TOL * eb + 1e-30)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound"


def permuted(rp, col, val, n, seed):
    """P A P^T of a symmetric CSR: same nnz, different structure."""
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    p = np.random.default_rng(seed).permutation(n)
    B = A[p][:, p].tocsr()
    B.sort_indices()
    return B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.astype(np.float32)


def bits_of(keys, n):
    w = np.zeros((n + 31) // 32, np.uint32)
    for k in np.asarray(keys, np.int64):
        w[k >> 5] |= np.uint32(1) << np.uint32(k & 31)
    return w.view(np.int32)


@pytest.fixture(scope="module")
def two_graphs(gpu):
    n = 1 << 12
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=77)
    rp2, col2, val2 = permuted(rp, col, val, n, seed=5)
    assert rp2[-1] == rp[-1]
    return n, (rp, col, val), (rp2, col2, val2)


def test_new_csr_invalidates_transposed_csc(two_graphs):
    """InputCSR of a second matrix on a handle whose CSC was transposed from
    the first: every slot (column slots read the CSC) follows the new one."""
    n, g1, g2 = two_graphs
    M = sa.Matrix(n, n, *g1)
    key, val = sa.gen_sparse_vector(n, 300, 0x77)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    # first matrix, column slots, so the CSC and the x scan are in use
    cp1, ri1, cv1 = oracle.csr_to_csc(n, n, *g1)
    for slot in (2, 3):
        M.A.set_sparse_x(dk, dv, key.size)
        c = M.A.run(slot, y_key=yk, y_val=yv)
        check(n, cp1, ri1, cv1, key, val, yk.to_host(c), yv.to_host(c), f"first matrix slot {slot}")
    d2 = [sa.DeviceArray.from_host(a, t) for a, t in zip(g2, (np.int32, np.int32, np.float32))]
    M.A.InputCSR(int(g2[0][-1]), *d2)
    cp2, ri2, cv2 = oracle.csr_to_csc(n, n, *g2)
    for slot in range(9):
        say(f"second matrix slot {slot}")
        M.A.set_sparse_x(dk, dv, key.size)
        c = M.A.run(slot, y_key=yk, y_val=yv)
        check(n, cp2, ri2, cv2, key, val, yk.to_host(c), yv.to_host(c), f"second matrix slot {slot}")
    # the x was kept and its scan recomputed for the new CSC, also without set_sparse_x
    c = M.A.run(sa.K_COL_LB_ATOMIC, y_key=yk, y_val=yv)
    check(n, cp2, ri2, cv2, key, val, yk.to_host(c), yv.to_host(c), "second matrix, same x")


def test_bfs_graphs_follow_matrix_and_arena(two_graphs):
    """Captured BFS levels of the first graph are not replayed on the second,
    nor after the arena was freed and reserved again."""
    n, g1, g2 = two_graphs
    M = sa.Matrix(n, n, *g1)
    lev = sa.DeviceArray(n, np.int32)
    deg = np.diff(g1[0])
    src = int(np.nonzero(deg > 0)[0][3])
    for policy in (-1, sa.K_ROW_LB, sa.K_COL_LB_SORT):
        M.A.bfs(src, lev, policy=policy, use_graphs=1)
        ref, _ = oracle.bfs(n, g1[0], g1[1], src)
        np.testing.assert_array_equal(lev.to_host(), ref)
    d2 = [sa.DeviceArray.from_host(a, t) for a, t in zip(g2, (np.int32, np.int32, np.float32))]
    M.A.InputCSR(int(g2[0][-1]), *d2)
    src2 = int(np.nonzero(np.diff(g2[0]) > 0)[0][3])
    for policy in (-1, sa.K_ROW_LB, sa.K_COL_LB_SORT):
        say(f"bfs on the second graph, policy {policy}")
        M.A.bfs(src2, lev, policy=policy, use_graphs=1)
        ref, _ = oracle.bfs(n, g2[0], g2[1], src2)
        np.testing.assert_array_equal(lev.to_host(), ref, err_msg=f"second graph policy {policy}")
    M.A.deallocPreBuffer()
    M.A.allocPreBuffer()
    for policy in (-1, sa.K_COL_LB_ATOMIC):
        say(f"bfs after a new arena, policy {policy}")
        M.A.bfs(src2, lev, policy=policy, use_graphs=1)
        ref, _ = oracle.bfs(n, g2[0], g2[1], src2)
        np.testing.assert_array_equal(lev.to_host(), ref, err_msg=f"new arena policy {policy}")


def test_bucket_counts(two_graphs):
    n, g1, _ = two_graphs
    M = sa.Matrix(n, n, *g1)
    cp, ri, cv = oracle.csr_to_csc(n, n, *g1)
    key, val = sa.gen_sparse_vector(n, 900, 0x78)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    L = sa.lib()
    for b in (2, 16, 256, 1024, 0):
        say(f"bucket slot with {b} buckets")
        M.A.set_sparse_x(dk, dv, key.size)
        c = ctypes.c_int(-1)
        assert L.spmspv_csc_bucket(M.A.h, 1.0, ctypes.byref(c), yk.ptr, yv.ptr, b) == 0
        check(n, cp, ri, cv, key, val, yk.to_host(c.value), yv.to_host(c.value), f"{b} buckets")
    for bad in (1, 3, 2048, -4):
        assert L.spmspv_csc_bucket(M.A.h, 1.0, None, yk.ptr, yv.ptr, bad) == -20


@pytest.fixture(scope="module")
def mat12(gpu):
    n = 1 << 12
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=41)
    M = sa.Matrix(n, n, rp, col, val)
    cp, ri, cv = oracle.csr_to_csc(n, n, rp, col, val)
    return n, M, (cp, ri, cv)


def dense_x(n, seed):
    """A dense x with exact zeros and nonzeros (support = the nonzeros)."""
    rng = np.random.default_rng(seed)
    x = np.zeros(n, np.float32)
    keys = np.sort(rng.choice(n, size=n // 7, replace=False)).astype(np.int32)
    x[keys] = rng.uniform(-1, 1, keys.size).astype(np.float32)
    x[keys[::11]] = 0.0  # stored zeros are not part of a dense x's support
    nzk = np.nonzero(x)[0].astype(np.int32)
    return x, nzk, x[nzk]


@pytest.mark.parametrize("slot", list(range(9)))
def test_dense_x_every_slot(mat12, slot):
    n, M, (cp, ri, cv) = mat12
    x, nzk, nzv = dense_x(n, 100 + slot)
    dx = sa.DeviceArray.from_host(x, np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    M.A.set_vector_type(sa.SPMSPV_VEC_DENSE)
    M.A.set_x(dx)
    c = M.A.run(slot, y_key=yk, y_val=yv)
    check(n, cp, ri, cv, nzk, nzv, yk.to_host(c), yv.to_host(c), f"dense x slot {slot}")


def test_dense_to_bits_to_sparse(mat12):
    """dense2bitarray fills the caller's bit array with the NONZEROS of x
    (the reference's dense2bitarray_ker marked the zeros); bitarray2sparse
    and dense2sparse give x's nonzeros and leave that array intact."""
    n, M, (cp, ri, cv) = mat12
    x, nzk, nzv = dense_x(n, 7)
    dx = sa.DeviceArray.from_host(x, np.float32)
    words = (n + 31) // 32
    ub = sa.DeviceArray.from_host(np.full(words, -1, np.int32), np.int32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    M.A.set_vector_type(sa.SPMSPV_VEC_DENSE)
    M.A.set_x(dx)
    M.A.set_bitvector(ub, words)
    M.A.dense2bitarray()
    want = bits_of(nzk, n)
    np.testing.assert_array_equal(ub.to_host(), want)
    M.A.bitarray2sparse()
    np.testing.assert_array_equal(ub.to_host(), want)  # still the caller's bits
    for slot in (sa.K_COL_LB_ATOMIC, sa.K_ROW_LB, sa.K_ROW_NAIVE):
        c = M.A.run(slot, y_key=yk, y_val=yv)
        check(n, cp, ri, cv, nzk, nzv, yk.to_host(c), yv.to_host(c), f"bitarray2sparse slot {slot}")
    M.A.set_vector_type(sa.SPMSPV_VEC_DENSE)
    M.A.set_x(dx)
    M.A.dense2sparse()
    np.testing.assert_array_equal(ub.to_host(), want)
    c = M.A.run(sa.K_COL_LB_SORT, y_key=yk, y_val=yv)
    check(n, cp, ri, cv, nzk, nzv, yk.to_host(c), yv.to_host(c), "dense2sparse")


def test_sparse_to_bits_and_dense(mat12):
    """sparse2bitarray / sparse2dense write the caller's buffers: bit k set
    iff k is a key (zero values included), dense = the scattered values."""
    n, M, (cp, ri, cv) = mat12
    key, val = sa.gen_sparse_vector(n, 333, 0x55)
    val[::9] = 0.0  # a stored zero is in a sparse x's support
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    words = (n + 31) // 32
    ub = sa.DeviceArray.from_host(np.zeros(words, np.int32), np.int32)
    ud = sa.DeviceArray.from_host(np.full(n, 9.0, np.float32), np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    M.A.set_vector_type(sa.SPMSPV_VEC_SPARSE)
    M.A.set_sparse_x(dk, dv, key.size)
    M.A.set_x(ud)
    M.A.set_bitvector(ub, words)
    M.A.sparse2dense()
    want_d = np.zeros(n, np.float32)
    want_d[key] = val
    np.testing.assert_array_equal(ud.to_host(), want_d)
    M.A.sparse2bitarray()
    np.testing.assert_array_equal(ub.to_host(), bits_of(key, n))
    for slot in (sa.K_ROW_NAIVE, sa.K_ROW_LB, sa.K_SPMV_LB, sa.K_COL_NAIVE_SORT):
        c = M.A.run(slot, y_key=yk, y_val=yv)
        check(n, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"sparse2bitarray slot {slot}")


def test_outer_conversions_round_trip(mat12):
    n, M, _ = mat12
    x, nzk, nzv = dense_x(n, 9)
    dx = sa.DeviceArray.from_host(x, np.float32)
    k = sa.DeviceArray(n, np.int32)
    v = sa.DeviceArray(n, np.float32)
    c = M.A.dense2sparse_outer(n, dx, k, v)
    assert c == nzk.size
    np.testing.assert_array_equal(k.to_host(c), nzk)
    np.testing.assert_array_equal(v.to_host(c), nzv)
    back = sa.DeviceArray.from_host(np.full(n, 3.0, np.float32), np.float32)
    M.A.sparse2dense_outer(c, k, v, n, back)
    np.testing.assert_array_equal(back.to_host(), x)
