"""Pins the CPU oracle (oracle/oracle.c) against closed forms, brute force,
textbook library routines (scipy) and the reference's own known-answer
pattern. The reference ships no golden vectors, so these checks are what
fixes the oracle; each is chosen so that a plausible slip (dropped term,
wrong index, transposed operand, off-by-one) fails one of them."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

import oracle


def rand_csr(m, n, density, seed, values="uniform"):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="csr", random_state=rng, dtype=np.float32)
    if values == "ones":
        A.data[:] = 1.0
    else:
        A.data = rng.uniform(-1, 1, A.nnz).astype(np.float32)
    A.sort_indices()
    return A


def rand_x(n, k, seed):
    rng = np.random.default_rng(seed)
    key = np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)
    val = rng.uniform(-1, 1, k).astype(np.float32)
    return key, val


# ------------------------------------------------------------ transpose
@pytest.mark.parametrize("m,n,d", [(7, 5, 0.4), (50, 80, 0.05), (300, 200, 0.01), (1, 1, 1.0)])
def test_csr_to_csc_matches_scipy(m, n, d):
    A = rand_csr(m, n, d, seed=m * 31 + n)
    cp, ri, cv = oracle.csr_to_csc(m, n, A.indptr, A.indices, A.data)
    C = A.tocsc()
    C.sort_indices()
    np.testing.assert_array_equal(cp, C.indptr)
    np.testing.assert_array_equal(ri, C.indices)
    np.testing.assert_array_equal(cv, C.data)


def test_csr_to_csc_empty_rows_and_cols():
    # rows 0 and 3 empty, columns 1 and 4 empty
    dense = np.array([[0, 0, 0, 0, 0], [2, 0, 3, 0, 0], [0, 0, 0, 5, 0], [0, 0, 0, 0, 0]], np.float32)
    A = sp.csr_matrix(dense)
    cp, ri, cv = oracle.csr_to_csc(4, 5, A.indptr, A.indices, A.data)
    np.testing.assert_array_equal(cp, [0, 1, 1, 2, 3, 3])
    np.testing.assert_array_equal(ri, [1, 1, 2])
    np.testing.assert_array_equal(cv, [2, 3, 5])


# ------------------------------------------------------------- SpMSpV
@pytest.mark.parametrize("seed", range(6))
def test_spmspv_f64_brute_force(seed):
    m, n = 40 + seed, 30 + 2 * seed
    A = rand_csr(m, n, 0.15, seed)
    C = A.tocsc()
    C.sort_indices()
    key, val = rand_x(n, 1 + seed * 3, seed + 100)
    yk, yv, yb = oracle.spmspv_f64(m, C.indptr, C.indices, C.data, key, val)
    Ad = A.toarray().astype(np.float64)
    x = np.zeros(n)
    x[key] = val
    y = Ad @ x
    support = np.nonzero((Ad[:, key] != 0).any(axis=1))[0]
    np.testing.assert_array_equal(yk, support)
    np.testing.assert_allclose(yv, y[support], rtol=0, atol=1e-12)
    np.testing.assert_allclose(yb, (np.abs(Ad) @ np.abs(x))[support], rtol=1e-12)


def test_spmspv_f64_unit_vector_is_column():
    A = rand_csr(60, 60, 0.1, 7)
    C = A.tocsc()
    C.sort_indices()
    for j in (0, 13, 59):
        yk, yv, _ = oracle.spmspv_f64(60, C.indptr, C.indices, C.data, np.array([j]), np.array([1.0], np.float32))
        np.testing.assert_array_equal(yk, C.indices[C.indptr[j]:C.indptr[j + 1]])
        np.testing.assert_array_equal(yv, C.data[C.indptr[j]:C.indptr[j + 1]].astype(np.float64))


def test_spmspv_reference_kat_all_ones():
    """generate_model_data.cu:639-643, 942-944: every value 1, x = 1 on
    [0, k): y[i] = number of stored entries of row i in columns < k."""
    A = rand_csr(500, 400, 0.03, 3, values="ones")
    C = A.tocsc()
    C.sort_indices()
    for k in (1, 17, 400):
        key = np.arange(k, dtype=np.int32)
        yk, yv, _ = oracle.spmspv_f64(500, C.indptr, C.indices, C.data, key, np.ones(k, np.float32))
        counts = np.array([(A.indices[A.indptr[i]:A.indptr[i + 1]] < k).sum() for i in range(500)])
        np.testing.assert_array_equal(yk, np.nonzero(counts)[0])
        np.testing.assert_array_equal(yv, counts[counts > 0].astype(np.float64))
        # the FP32 reference paths are exact on this integer pattern
        yd, k32, v32 = oracle.spmspv_csc_host_f32(500, C.indptr, C.indices, C.data, key, np.ones(k, np.float32))
        np.testing.assert_array_equal(k32, yk)
        np.testing.assert_array_equal(v32, counts[counts > 0].astype(np.float32))
        xd = np.zeros(400, np.float32)
        xd[:k] = 1
        np.testing.assert_array_equal(oracle.serialspmv_f32(500, A.indptr, A.indices, A.data, xd), counts.astype(np.float32))


def test_spmspv_structural_support_keeps_exact_zeros():
    # row 0 = 1*x0 - 1*x1 cancels exactly; x value 0 at column 2 still counts
    dense = np.array([[1, -1, 0], [0, 0, 4], [0, 0, 0]], np.float32)
    C = sp.csc_matrix(dense)
    C.sort_indices()
    yk, yv, _ = oracle.spmspv_f64(3, C.indptr, C.indices, C.data, np.array([0, 1, 2]), np.array([1, 1, 0], np.float32))
    np.testing.assert_array_equal(yk, [0, 1])
    np.testing.assert_array_equal(yv, [0.0, 0.0])
    # the reference's FLT_EPSILON filter (test_sort_spmspv.cu:354-372) drops both
    _, k32, _ = oracle.spmspv_csc_host_f32(3, C.indptr, C.indices, C.data, np.array([0, 1, 2]), np.array([1, 1, 0], np.float32))
    assert k32.size == 0


def test_spmspv_structural_support_keeps_inf_and_nan_rows():
    # a NaN or Inf x value still makes its column's rows part of the result:
    # row 0 gets NaN (1*NaN), row 1 Inf (4*Inf), row 2 NaN (Inf - Inf), row 3 finite
    dense = np.array([[1, 0, 0], [0, 4, 0], [0, 1, -1], [0, 0, 0]], np.float32)
    dense[3, 2] = 2.0
    C = sp.csc_matrix(dense)
    C.sort_indices()
    x = np.array([np.nan, np.inf, np.inf], np.float32)
    yk, yv, yb = oracle.spmspv_f64(4, C.indptr, C.indices, C.data, np.array([0, 1, 2]), x)
    np.testing.assert_array_equal(yk, [0, 1, 2, 3])
    assert np.isnan(yv[0]) and yv[1] == np.inf and np.isnan(yv[2]) and yv[3] == np.inf
    # only x = (NaN at column 0): row 0 alone, NaN
    yk, yv, _ = oracle.spmspv_f64(4, C.indptr, C.indices, C.data, np.array([0]), x[:1])
    np.testing.assert_array_equal(yk, [0])
    assert np.isnan(yv[0])


def test_spmspv_column_sum_identity():
    A = rand_csr(200, 150, 0.05, 11)
    C = A.tocsc()
    C.sort_indices()
    key, val = rand_x(150, 40, 12)
    _, yv, _ = oracle.spmspv_f64(200, C.indptr, C.indices, C.data, key, val)
    colsum = np.asarray(A.astype(np.float64).sum(axis=0)).ravel()
    assert abs(yv.sum() - (val.astype(np.float64) * colsum[key]).sum()) < 1e-9


def test_spmspv_empty_inputs():
    A = rand_csr(20, 20, 0.2, 1)
    C = A.tocsc()
    yk, yv, yb = oracle.spmspv_f64(20, C.indptr, C.indices, C.data, np.zeros(0, np.int32), np.zeros(0, np.float32))
    assert yk.size == 0 and yv.size == 0
    cp = np.zeros(21, np.int32)
    yk, _, _ = oracle.spmspv_f64(20, cp, np.zeros(0, np.int32), np.zeros(0, np.float32), np.array([3]), np.array([1.0], np.float32))
    assert yk.size == 0


def test_spmv_f64_matches_dense():
    A = rand_csr(90, 70, 0.1, 5)
    x = np.random.default_rng(1).uniform(-1, 1, 70).astype(np.float32)
    y, b = oracle.spmv_f64(90, A.indptr, A.indices, A.data, x)
    Ad = A.toarray().astype(np.float64)
    np.testing.assert_allclose(y, Ad @ x.astype(np.float64), atol=1e-12)
    np.testing.assert_allclose(b, np.abs(Ad) @ np.abs(x.astype(np.float64)), rtol=1e-12)
    y32 = oracle.serialspmv_f32(90, A.indptr, A.indices, A.data, x)
    assert np.all(np.abs(y32 - y) <= 1e-5 * b + 1e-30)


# --------------------------------------------------------------- scan
def test_scan_col_lengths_with_empty_columns():
    cp = np.array([0, 0, 0, 3, 3, 5, 5], np.int32)  # columns 0,1,3,5 empty
    key = np.array([0, 1, 2, 3, 4, 5], np.int32)
    ptr = oracle.scan_col_lengths(cp, key)
    np.testing.assert_array_equal(ptr, [0, 0, 0, 3, 3, 5, 5])
    ptr = oracle.scan_col_lengths(cp, np.array([4, 2], np.int32))
    np.testing.assert_array_equal(ptr, [0, 2, 5])


# ---------------------------------------------------------------- BFS
@pytest.mark.parametrize("seed", range(4))
def test_bfs_matches_scipy_shortest_path(seed):
    A = rand_csr(300, 300, 0.008, seed, values="ones")
    A = ((A + A.T) > 0).astype(np.float32).tocsc()
    A.sort_indices()
    src = seed * 17
    lev, nl = oracle.bfs(300, A.indptr, A.indices, src)
    d = shortest_path(A, unweighted=True, indices=src, directed=True)
    ref = np.where(np.isinf(d), -1, d).astype(np.int32)
    # BFS over column u visits the rows of column u: that is A^T; A is symmetric
    np.testing.assert_array_equal(lev, ref)
    assert nl == ref.max() + 1


def test_bfs_levels_invariants_directed():
    A = rand_csr(200, 200, 0.02, 9, values="ones").tocsc()
    A.sort_indices()
    lev, _ = oracle.bfs(200, A.indptr, A.indices, 5)
    assert lev[5] == 0
    for u in range(200):
        if lev[u] < 0:
            continue
        for v in A.indices[A.indptr[u]:A.indptr[u + 1]]:
            assert lev[v] >= 0 and lev[v] <= lev[u] + 1


# ----------------------------------------------------------- features
def test_mat_features_regular_closed_form():
    """d-regular rows: std 0, equality 1, and gini 1/m, because S2 omits the
    largest degree (spmspv.h:2318-2324): S1 = d*m(m-1)/2, S2 = d*(m-1)."""
    m, d = 64, 5
    rp = np.arange(0, (m + 1) * d, d, dtype=np.int32)
    f = oracle.mat_features(m, 100, rp)
    assert f[0] == d and f[1] == d and f[2] == d
    assert f[3] == 0.0 and f[4] == 0.0
    assert abs(f[5] - 1.0) < 1e-12
    assert abs(f[6] - 1.0 / m) < 1e-12
    assert oracle.std_row_trunc(m, rp) == 0.0


def test_mat_features_hand_example():
    # degrees 3,0,2,1 -> sorted 0,1,2,3; nnz 6, m 4, n 10
    rp = np.array([0, 3, 3, 5, 6], np.int32)
    f = oracle.mat_features(4, 10, rp)
    avg = 1.5
    std = np.sqrt(((3 - avg) ** 2 + avg ** 2 + 0.25 + 0.25) / 4)
    eq = -sum(p * np.log(p) for p in (3 / 6, 2 / 6, 1 / 6)) / np.log(4)
    s1 = 1 * 1 + 2 * 2 + 3 * 3
    s2 = 0 + 1 + 2
    gini = abs(2 * s1 / (4 * s2) - 5 / 4)
    np.testing.assert_allclose(f, [3, 0, 1.5, 0.3, std, eq, gini], rtol=1e-12)
    # int-truncated mean 1: (4 + 1 + 1 + 0)/4
    assert abs(oracle.std_row_trunc(4, rp) - np.sqrt(6 / 4)) < 1e-12


def test_vec_features_and_gm():
    cp = np.array([0, 4, 4, 9, 10], np.int32)
    out = oracle.vec_features(cp, np.array([0, 1, 3], np.int32))
    np.testing.assert_array_equal(out, [5, 4, 0])
    out = oracle.vec_features(cp, np.zeros(0, np.int32))
    assert out[0] == 0 and out[1] == 0 and out[2] == 2**31 - 1
    g = oracle.gm(100, 1000, 10, 50)
    # GM1 = 4 xnnz + 4 bin + 4 bin + 4 xnnz + 4 xnnz + 4 bin
    assert g[0] == 4 * 10 * 3 + 4 * 50 * 3
    # GM2 = 4(m+1) + 4 bin + 4 nnz + 4 bin + 4 m
    assert g[1] == 4 * 101 + 4 * 50 + 4 * 1000 + 4 * 50 + 4 * 100
    # GM3 = 4(m+1) + 3*4 nnz + 4 m
    assert g[2] == 4 * 101 + 12 * 1000 + 4 * 100


# ---------------------------------------------------------------- MLP
def test_mlp_gradient_finite_differences():
    F, H1, H2, C, B = 16, 64, 64, 8, 5
    rng = np.random.default_rng(0)
    npar = oracle.mlp_param_count(F, H1, H2, C)
    p = rng.normal(0, 0.3, npar)
    X = rng.normal(0, 1, (B, F))
    lab = rng.integers(0, C, B).astype(np.int32)
    loss, g = oracle.mlp_loss_grad(F, H1, H2, C, p, X, labels=lab)
    idx = rng.choice(npar, 60, replace=False)
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = oracle.mlp_loss_grad(F, H1, H2, C, q, X, labels=lab, want_grad=False)
        q[i] -= 2 * eps
        lm, _ = oracle.mlp_loss_grad(F, H1, H2, C, q, X, labels=lab, want_grad=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # soft targets equal to one-hot give the same loss
    soft = np.eye(C)[lab]
    l2, _ = oracle.mlp_loss_grad(F, H1, H2, C, p, X, soft=soft, want_grad=False)
    assert abs(l2 - loss) < 1e-12
    # uniform logits: loss = ln C
    z = np.zeros(npar)
    l0, _ = oracle.mlp_loss_grad(F, H1, H2, C, z, X, labels=lab, want_grad=False)
    assert abs(l0 - np.log(C)) < 1e-12


def test_mlp_forward_hand_network():
    """A 3->2->3->2 network small enough to evaluate by hand. Sample 0 has
    one dead hidden unit per layer and a negative logit, so a ReLU applied to
    the logits, a transposed weight matrix or swapped layer offsets each
    change the result."""
    F, H1, H2, C = 3, 2, 3, 2
    W1 = [1, 1, 1, 2, 0, -1]
    b1 = [0.5, -4]
    W2 = [1, 3, -2, 1, 0.5, 0.5]
    b2 = [0, 1, -1]
    W3 = [1, 2, 4, -1, 0, -2]
    b3 = [0.25, 0]
    p = np.array(W1 + b1 + W2 + b2 + W3 + b3, np.float64)
    assert p.size == oracle.mlp_param_count(F, H1, H2, C)
    X = np.array([[1, 2, -1], [0, 0, 0]], np.float64)
    # sample 0: z1 = (2.5, -2) -> h1 = (2.5, 0); z2 = (2.5, -4, 0.25) ->
    # h2 = (2.5, 0, 0.25); logits = (2.5 + 1 + 0.25, -2.5 - 0.5) = (3.75, -3)
    # sample 1: h1 = relu(b1) = (0.5, 0); z2 = (0.5, 0, -0.75) ->
    # h2 = (0.5, 0, 0); logits = (0.75, -0.5)
    out = oracle.mlp_forward(F, H1, H2, C, p, X)
    np.testing.assert_array_equal(out, [[3.75, -3.0], [0.75, -0.5]])


def test_mlp_forward_affine_and_dead_regions():
    """At the selector's own shape (16->64->64->8): with positive inputs and
    positive first/second-layer parameters no ReLU clips, so the network is
    the composition of its three affine maps; with a very negative second
    bias every second-layer unit is dead and the logits are exactly b3."""
    F, H1, H2, C, B = 16, 64, 64, 8, 6
    rng = np.random.default_rng(7)
    W1 = rng.uniform(0.01, 0.5, (H1, F))
    b1 = rng.uniform(0.0, 0.2, H1)
    W2 = rng.uniform(0.01, 0.5, (H2, H1))
    b2 = rng.uniform(0.0, 0.2, H2)
    W3 = rng.normal(0, 0.5, (C, H2))
    b3 = rng.normal(0, 0.5, C)
    X = rng.uniform(0.0, 1.0, (B, F))
    pack = lambda w2b: np.concatenate([W1.ravel(), b1, W2.ravel(), w2b, W3.ravel(), b3])
    out = oracle.mlp_forward(F, H1, H2, C, pack(b2), X)
    want = ((X @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    np.testing.assert_allclose(out, want, rtol=1e-13, atol=1e-12)
    dead = oracle.mlp_forward(F, H1, H2, C, pack(np.full(H2, -1e6)), X)
    np.testing.assert_array_equal(dead, np.tile(b3, (B, 1)))


def test_mlp_forward_log_softmax_is_per_sample_loss():
    """The per-sample loss of the finite-difference-pinned loss function at
    label c equals -log_softmax(forward logits)[c], for every label."""
    F, H1, H2, C, B = 16, 64, 64, 8, 4
    rng = np.random.default_rng(11)
    p = rng.normal(0, 0.3, oracle.mlp_param_count(F, H1, H2, C))
    X = rng.normal(0, 1, (B, F))
    z = oracle.mlp_forward(F, H1, H2, C, p, X)
    lse = z.max(axis=1) + np.log(np.exp(z - z.max(axis=1, keepdims=True)).sum(axis=1))
    for s in range(B):
        for c in range(C):
            loss, _ = oracle.mlp_loss_grad(F, H1, H2, C, p, X[s:s + 1], labels=np.array([c]), want_grad=False)
            assert abs(loss - (lse[s] - z[s, c])) < 1e-12, (s, c)


# ----------------------------------------------------------- PageRank
def _pagerank_solve(n, A_csc, d):
    """p* = (I - d A W^-1)^-1 (1-d)/n, W = diag of column sums (0 -> 0)."""
    import scipy.sparse.linalg as spla
    w = np.asarray(A_csc.sum(axis=0)).ravel()
    winv = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0)
    P = A_csc @ sp.diags(winv)
    return spla.spsolve((sp.identity(n) - d * P).tocsc(), np.full(n, (1 - d) / n))


@pytest.mark.parametrize("seed", [0, 1])
def test_pagerank_matches_linear_solve(seed):
    n = 300
    A = rand_csr(n, n, 0.02, seed).tocsc()
    A.data = np.abs(A.data) + np.float32(0.05)   # positive weights
    A.sort_indices()
    want = _pagerank_solve(n, A.astype(np.float64), 0.85)
    got, sizes = oracle.pagerank_f64(n, A.indptr, A.indices, A.data, 0.85, 1e-13, 5000)
    assert sizes[0] == n
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


def test_pagerank_cycle_is_uniform_and_eps_bound_holds():
    n = 64
    # directed cycle u -> u+1: every column has one entry, p* = 1/n
    cp = np.arange(n + 1, dtype=np.int32)
    ri = ((np.arange(n) + 1) % n).astype(np.int32)
    cv = np.ones(n, np.float32)
    got, _ = oracle.pagerank_f64(n, cp, ri, cv, 0.85, 1e-12, 5000)
    np.testing.assert_allclose(got, 1.0 / n, rtol=1e-10)
    # a coarse eps stops early, within sum|p - p*| <= eps * d
    A = rand_csr(200, 200, 0.03, 9, values="ones").tocsc()
    A.sort_indices()
    want = _pagerank_solve(200, A.astype(np.float64), 0.85)
    for eps in (1e-1, 1e-2, 1e-3):
        got, sizes = oracle.pagerank_f64(200, A.indptr, A.indices, A.data, 0.85, eps, 5000)
        assert np.abs(got - want).sum() <= eps * 0.85
        assert sizes.size > 1 and sizes[0] == 200


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_allcore_port_equals_serial_bit_for_bit(threads):
    m, n = 500, 400
    A = rand_csr(m, n, 0.05, 21 + threads).tocsc()
    A.sort_indices()
    key, val = rand_x(n, 150, 5)
    key = key[np.random.default_rng(1).permutation(key.size)]  # any key order
    yd, yk, yv = oracle.spmspv_csc_host_f32(m, A.indptr, A.indices, A.data, key, val)
    pd, pk, pv = oracle.spmspv_csc_host_f32_omp(m, A.indptr, A.indices, A.data, key, val, threads)
    np.testing.assert_array_equal(pd.view(np.uint32), yd.view(np.uint32))
    np.testing.assert_array_equal(pk, yk)
    np.testing.assert_array_equal(pv.view(np.uint32), yv.view(np.uint32))
