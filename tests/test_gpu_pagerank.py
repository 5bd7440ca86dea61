"""The PageRank-style driver and the frontier replay harness through the
C-ABI. PageRank ranks against the FP64 solution of (I - d A W^-1) p =
(1 - d)/n: once converged, sum|p - p*| <= eps * d, plus FP32 rounding
(documented slack 3e-5 over about 70 iterations); iteration counts within
3 of the FP64 oracle's (a residual within rounding of the threshold may
flip). Replayed frontiers: index sets bit-exact and values within
1e-5 * sum|a_ij x_j| of the FP64 oracle."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5
D = 0.85


@pytest.fixture(scope="module")
def g12(gpu):
    n = 1 << 12
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=61)
    M = sa.Matrix(n, n, rp, col, val)
    cp, ri, cv = oracle.csr_to_csc(n, n, rp, col, val)
    A = sp.csc_matrix((cv.astype(np.float64), ri, cp), shape=(n, n))
    w = np.asarray(A.sum(axis=0)).ravel()
    winv = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0)
    pstar = spla.spsolve((sp.identity(n) - D * (A @ sp.diags(winv))).tocsc(), np.full(n, (1 - D) / n))
    return n, M, (cp, ri, cv), pstar


@pytest.mark.parametrize("policy", [-1, 0, 1, 2, 3, 4, 5, 6, 7])
def test_pagerank_parity(g12, policy):
    n, M, (cp, ri, cv), pstar = g12
    eps = 1e-5
    print(f"pagerank policy {policy}", flush=True)
    rank = sa.DeviceArray(n, np.float32)
    st, sizes = M.A.pagerank(rank, D, eps, 2000, policy)
    got = rank.to_host().astype(np.float64)
    ref, rsizes = oracle.pagerank_f64(n, cp, ri, cv, D, eps, 2000)
    assert st[3] == 0, "not converged"
    assert sizes[0] == n and rsizes[0] == n
    assert abs(sizes.size - rsizes.size) <= 3, (sizes.size, rsizes.size)
    assert np.abs(ref - pstar).sum() <= eps * D
    assert np.abs(got - pstar).sum() <= eps * D + 3e-5, np.abs(got - pstar).sum()
    # the active set shrinks from every vertex to none: the dense-to-sparse
    # sweep the selector sees
    assert sizes[-1] < n // 10


def test_frontier_replay_matches_oracle(g12, tmp_path):
    """BFS frontiers written in the reference's files (bfs_xnnz_*.txt: one
    size per iteration; bfs_x_*.txt: one line of keys per iteration) and
    replayed, selector-dispatched and for two fixed slots."""
    n, M, (cp, ri, cv), _ = g12
    rp = M.h_row_ptr
    src = int(np.nonzero(np.diff(rp) > 0)[0][0])
    lev, nl = oracle.bfs(n, cp, ri, src)
    levels = [np.nonzero(lev == L)[0].astype(np.int32) for L in range(nl)]
    sizes_p = tmp_path / "bfs_xnnz_g.txt"
    x_p = tmp_path / "bfs_x_g.txt"
    sizes_p.write_text("\n".join(str(f.size) for f in levels) + "\n")
    x_p.write_text("\n".join(" ".join(map(str, f)) for f in levels) + "\n")
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    for slots in (None, [sa.K_COL_LB_SORT, sa.K_SPMV_LB]):
        res = sa.replay_frontiers(M.A, str(sizes_p), str(x_p), yk, yv, slots=slots)
        assert len(res) == nl
        for it, f in zip(res, levels):
            np.testing.assert_array_equal(it["keys"], f)
        # each run's result lives in yk/yv only until the next one: run the
        # last slot of every iteration again and check it against the oracle
        for it, f in zip(res, levels):
            dk = sa.DeviceArray.from_host(f if f.size else np.zeros(1, np.int32), np.int32)
            dv = sa.DeviceArray.from_host(np.ones(max(f.size, 1), np.float32), np.float32)
            M.A.set_sparse_x(dk, dv, int(f.size))
            slot = it["runs"][-1]["slot"]
            c = M.A.run(slot, y_key=yk, y_val=yv)
            assert c == it["runs"][-1]["ynnz"]
            ek, ev, eb = oracle.spmspv_f64(n, cp, ri, cv, f, np.ones(f.size, np.float32))
            np.testing.assert_array_equal(yk.to_host(c), ek)
            assert np.all(np.abs(yv.to_host(c).astype(np.float64) - ev) <= TOL * eb + 1e-30)
