"""BFS levels through the C-ABI, bit-exact against the CPU BFS oracle,
for every fixed slot, the per-level selector, and hipGraph-captured
levels."""
import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph12(gpu):
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=31)
    n = 1 << 12
    M = sa.Matrix(n, n, rp, col, val)
    cp, ri, _ = oracle.csr_to_csc(n, n, rp, col, val)
    deg = np.diff(rp)
    srcs = [int(v) for v in np.nonzero(deg > 0)[0][[0, 7, 100, 555]]]
    isolated = int(np.nonzero(deg == 0)[0][0]) if (deg == 0).any() else None
    return M, cp, ri, srcs, isolated


@pytest.mark.parametrize("policy", [-1, 0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("graphs", [0, 1])
def test_bfs_levels_exact(graph12, policy, graphs):
    M, cp, ri, srcs, isolated = graph12
    lev = sa.DeviceArray(M.n, np.int32)
    for src in srcs + ([isolated] if isolated is not None else []):
        print(f"bfs policy {policy} graphs {graphs} src {src}", flush=True)
        st = M.A.bfs(src, lev, policy=policy, use_graphs=graphs)
        ref, nl = oracle.bfs(M.n, cp, ri, src)
        got = lev.to_host()
        np.testing.assert_array_equal(got, ref)
        assert int(st[0]) == nl
        reached = ref >= 0
        assert int(st[3]) == int(reached.sum())
        assert st[1] == np.diff(cp)[reached].sum() / 2


def test_bfs_disconnected_and_path(gpu):
    # a path 0-1-2-...-99 plus a separate triangle; long, thin levels
    import scipy.sparse as sp
    n = 128
    rows = list(range(99)) + list(range(1, 100)) + [110, 111, 112, 111, 112, 110]
    cols = list(range(1, 100)) + list(range(99)) + [111, 112, 110, 110, 111, 112]
    A = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)), shape=(n, n))
    A.sort_indices()
    M = sa.Matrix(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32))
    cp, ri, _ = oracle.csr_to_csc(n, n, A.indptr, A.indices, A.data)
    lev = sa.DeviceArray(n, np.int32)
    for policy in (-1, 0, 1, 2, 3, 4, 5, 6, 7):
        for src in (0, 50, 111, 120):
            M.A.bfs(src, lev, policy=policy, use_graphs=1)
            ref, _ = oracle.bfs(n, cp, ri, src)
            np.testing.assert_array_equal(lev.to_host(), ref)


def test_bfs_rmat22_full_size(gpu):
    """The BFS workload at full size: R-MAT scale 22, edge factor 16,
    symmetrised (4,194,304 vertices, about 128 M stored edges), the first 4
    of the 64 sources the benchmark draws, levels bit-exact against the CPU
    BFS for the learned per-level policy (with and without captured levels)
    and two fixed slots."""
    import os
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd")
    n = 1 << 22
    rp, col, val = sa.gen_rmat_csr(22, 16, 0.57, 0.19, 0.19, seed=0x5EED0004)
    M = sa.Matrix(n, n, rp, col, val)
    M.A.selector_set_weights(np.load(os.path.join(root, "selector_weights.npy")).astype(np.float32))
    M.A.bfs_selector_set_weights(np.load(os.path.join(root, "bfs_selector_weights.npy")).astype(np.float32))
    deg = np.diff(rp)
    cand = np.nonzero(deg > 0)[0]
    srcs = np.random.default_rng(0x5EED0003).choice(cand, size=64, replace=False)[:4]
    lev = sa.DeviceArray(n, np.int32)
    for src in srcs:
        ref, nl = oracle.bfs(n, rp, col, int(src))  # symmetric: the CSR is the CSC
        for policy, graphs in ((-1, 1), (-1, 0), (sa.K_ROW_LB, 1), (sa.K_COL_LB_SORT, 1)):
            print(f"rmat22 bfs src {src} policy {policy} graphs {graphs}", flush=True)
            st = M.A.bfs(int(src), lev, policy=policy, use_graphs=graphs)
            np.testing.assert_array_equal(lev.to_host(), ref, err_msg=f"src {src} policy {policy}")
            assert int(st[0]) == nl


@pytest.mark.parametrize("nsrc", [1, 7, 64])
def test_bfs_multi_levels_exact(graph12, nsrc):
    """Multi-source BFS: the levels of every source equal the CPU BFS's
    (vertex-major output), sources including an isolated vertex and a
    repeated one."""
    M, cp, ri, srcs4, isolated = graph12
    deg = np.diff(cp)
    rng = np.random.default_rng(nsrc)
    srcs = rng.choice(np.nonzero(deg > 0)[0], size=nsrc, replace=True).astype(np.int32)
    if nsrc > 2:
        srcs[1] = srcs[0]  # repeated source
        if isolated is not None:
            srcs[2] = isolated
    lev = sa.DeviceArray(M.n * nsrc, np.int32)
    print(f"bfs_multi nsrc {nsrc}", flush=True)
    st = M.A.bfs_multi(srcs, lev)
    got = lev.to_host().reshape(M.n, nsrc)
    edges, reached, maxl = 0.0, 0, 0
    for i, s_ in enumerate(srcs):
        ref, nl = oracle.bfs(M.n, cp, ri, int(s_))
        np.testing.assert_array_equal(got[:, i], ref, err_msg=f"source {i} ({s_})")
        edges += deg[ref >= 0].sum() / 2
        reached += int((ref >= 0).sum())
        maxl = max(maxl, nl)
    assert int(st[3]) == reached
    assert st[1] == edges
    assert int(st[0]) == maxl


def test_bfs_multi_path_and_components(gpu):
    import scipy.sparse as sp
    n = 300
    rows = list(range(199)) + list(range(1, 200)) + [210, 211, 212, 211, 212, 210]
    cols = list(range(1, 200)) + list(range(199)) + [211, 212, 210, 210, 211, 212]
    A = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)), shape=(n, n))
    A.sort_indices()
    M = sa.Matrix(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32))
    cp, ri, _ = oracle.csr_to_csc(n, n, A.indptr, A.indices, A.data)
    srcs = np.array([0, 199, 100, 211, 250, 5], np.int32)
    lev = sa.DeviceArray(n * srcs.size, np.int32)
    M.A.bfs_multi(srcs, lev)
    got = lev.to_host().reshape(n, srcs.size)
    for i, s_ in enumerate(srcs):
        ref, _ = oracle.bfs(n, cp, ri, int(s_))
        np.testing.assert_array_equal(got[:, i], ref, err_msg=f"source {s_}")
