"""Matrix and frontier file I/O (host code of the C-ABI library; no GPU):
MatrixMarket reading with the reference's semantics (ReadMTX / SortMTX,
la/include/spmspv/mtx.hpp:218-402), the binary CSR cache (mtx.hpp:709-819)
and the recorded-frontier files (readsvmdata.hpp:223-348).

Pinned against scipy.io.mmread where the semantics agree (general and
symmetric real, pattern, integer files without duplicates) and against
hand-built expectations where they do not (duplicates kept, complex read
as 1.0, skew-symmetric read as general)."""
import io
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import spmspv_amd as sa


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def csr_of(m, n, rp, col, val):
    return sp.csr_matrix((val, col, rp), shape=(m, n))


def scipy_csr(path):
    A = scipy.io.mmread(path).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("symmetry", ["general", "symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_mtx_matches_scipy(tmp_path, field, symmetry):
    rng = np.random.default_rng(hash((field, symmetry)) & 0xFFFF)
    m = n = 40
    A = sp.random(m, n, density=0.08, random_state=rng, format="coo")
    if symmetry == "symmetric":
        A = sp.tril(A).tocoo()  # a symmetric file stores one triangle
    if field == "integer":
        A.data = rng.integers(-9, 10, A.nnz).astype(np.float64)
    body = io.BytesIO()
    scipy.io.mmwrite(body, A, field=field, symmetry=symmetry)
    path = write(tmp_path, "a.mtx", body.getvalue().decode())
    m2, n2, rp, col, val = sa.read_mtx(path)
    assert (m2, n2) == (m, n)
    R = scipy_csr(path)
    np.testing.assert_array_equal(rp, R.indptr)
    np.testing.assert_array_equal(col, R.indices)
    np.testing.assert_array_equal(val, R.data.astype(np.float32))
    # rows' entries in ascending column order
    for r in range(m):
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)


def test_mtx_symmetric_mirrors_off_diagonal_only(tmp_path):
    text = """%%MatrixMarket matrix coordinate real symmetric
% a comment line
%another
3 3 4
1 1 2.5
2 1 -1.0
3 2 4.0
3 3 1.0
"""
    m, n, rp, col, val = sa.read_mtx(write(tmp_path, "s.mtx", text))
    D = csr_of(m, n, rp, col, val).toarray()
    want = np.array([[2.5, -1.0, 0], [-1.0, 0, 4.0], [0, 4.0, 1.0]], np.float32)
    np.testing.assert_array_equal(D, want)
    assert rp[-1] == 2 * 4 - 2  # nnz*2 - diagonal entries (mtx.hpp:318)


def test_mtx_keeps_duplicates_in_file_order(tmp_path):
    text = """%%MatrixMarket matrix coordinate real general
2 3 4
2 3 1.5
1 2 7
2 3 -0.5
1 1 3
"""
    m, n, rp, col, val = sa.read_mtx(write(tmp_path, "d.mtx", text))
    np.testing.assert_array_equal(rp, [0, 2, 4])
    np.testing.assert_array_equal(col, [0, 1, 2, 2])
    np.testing.assert_array_equal(val, [3, 7, 1.5, -0.5])


def test_mtx_complex_and_skew_follow_the_reference(tmp_path):
    # complex entries become 1.0 (mtx.hpp:288-294); skew-symmetric is read
    # as general (only "symmetric" is expanded, mtx.hpp:243-246)
    text = """%%MatrixMarket matrix coordinate complex general
2 2 2
1 2 3.0 4.0
2 1 -1.0 0.5
"""
    m, n, rp, col, val = sa.read_mtx(write(tmp_path, "c.mtx", text))
    np.testing.assert_array_equal(val, [1.0, 1.0])
    text = """%%MatrixMarket matrix coordinate real skew-symmetric
3 3 1
3 1 2.0
"""
    m, n, rp, col, val = sa.read_mtx(write(tmp_path, "k.mtx", text))
    np.testing.assert_array_equal(rp, [0, 0, 0, 1])
    np.testing.assert_array_equal(col, [0])


def test_mtx_rectangular_and_empty_rows(tmp_path):
    text = """%%MatrixMarket matrix coordinate pattern general
4 6 3
4 6
4 1
1 3
"""
    m, n, rp, col, val = sa.read_mtx(write(tmp_path, "r.mtx", text))
    assert (m, n) == (4, 6)
    np.testing.assert_array_equal(rp, [0, 1, 1, 1, 3])
    np.testing.assert_array_equal(col, [2, 0, 5])
    np.testing.assert_array_equal(val, [1, 1, 1])


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",   # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n",   # truncated
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",        # dense file
])
def test_mtx_bad_files_are_errors(tmp_path, text):
    with pytest.raises(sa.SpmspvError) as e:
        sa.read_mtx(write(tmp_path, "bad.mtx", text))
    assert e.value.code == -13
    with pytest.raises(sa.SpmspvError):
        sa.read_mtx(str(tmp_path / "missing.mtx"))


def test_binary_csr_round_trip_and_layout(tmp_path):
    rp, col, val = sa.gen_rmat_csr(8, 8, seed=3)
    m = n = 256
    path = str(tmp_path / "g.mtx_.csr")
    sa.store_csr(path, m, n, rp, col, val)
    raw = open(path, "rb").read()
    nnz = int(rp[-1])
    # 24-byte header, then values, column indices, row pointers (mtx.hpp:806-819)
    assert len(raw) == 24 + 4 * nnz + 4 * nnz + 4 * (m + 1)
    assert raw[:9] == b"Hola\x01CSR\x01"
    assert np.frombuffer(raw[12:24], np.int32).tolist() == [m, n, nnz]
    np.testing.assert_array_equal(np.frombuffer(raw[24:24 + 4 * nnz], np.float32), val)
    np.testing.assert_array_equal(np.frombuffer(raw[24 + 8 * nnz:], np.int32), rp)
    m2, n2, rp2, col2, val2 = sa.load_csr(path)
    assert (m2, n2) == (m, n)
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(col2, col)
    np.testing.assert_array_equal(val2, val)


def test_binary_csr_rejects_bad_files(tmp_path):
    p = tmp_path / "x.csr"
    p.write_bytes(b"NotACSR!!" + bytes(15))
    with pytest.raises(sa.SpmspvError):
        sa.load_csr(str(p))
    rp = np.array([0, 1, 2], np.int32)
    sa.store_csr(str(p), 2, 2, rp, np.array([1, 0], np.int32), np.array([1, 2], np.float32))
    raw = p.read_bytes()
    p.write_bytes(raw[:-4])  # truncated row pointers
    with pytest.raises(sa.SpmspvError):
        sa.load_csr(str(p))


def test_mtx_to_binary_cache_flow(tmp_path):
    """The harness's flow (new_test_app.cu:527-566): read the .mtx, store the
    cache, and a later run loads the same CSR from the cache."""
    text = """%%MatrixMarket matrix coordinate real symmetric
5 5 6
1 1 1
2 1 2
3 2 3
4 3 4
5 4 5
5 5 6
"""
    src = write(tmp_path, "p.mtx", text)
    m, n, rp, col, val = sa.read_mtx(src)
    cache = src + "_.csr"
    sa.store_csr(cache, m, n, rp, col, val)
    m2, n2, rp2, col2, val2 = sa.load_csr(cache)
    assert (m2, n2) == (m, n)
    for a, b in ((rp, rp2), (col, col2), (val, val2)):
        np.testing.assert_array_equal(a, b)


def test_frontier_files(tmp_path):
    sizes = write(tmp_path, "bfs_xnnz_g.txt", "1\n3\n12 5\n")
    np.testing.assert_array_equal(sa.frontier_sizes(sizes), [1, 3, 12, 5])
    front = write(tmp_path, "bfs_x_g.txt", "7\n1 9 4\n\n  2   3 \n")
    k, v = sa.frontier_read(front, 1, 10)
    np.testing.assert_array_equal(k, [1, 9, 4])
    np.testing.assert_array_equal(v, [1, 1, 1])
    k, v = sa.frontier_read(front, 2, 10)
    assert k.size == 0
    k, _ = sa.frontier_read(front, 3, 10)
    np.testing.assert_array_equal(k, [2, 3])
    k, _ = sa.frontier_read(front, 9, 10)  # past the end
    assert k.size == 0
    with pytest.raises(sa.SpmspvError):
        sa.frontier_read(front, 1, 2)  # capacity too small
