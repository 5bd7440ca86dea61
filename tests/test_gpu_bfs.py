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
