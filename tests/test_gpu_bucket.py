"""The bucket slot (row bins summed in LDS windows) and the size rule that
lets auto mode run it, against the FP64 oracle through the C-ABI.

Index sets bit-exact; FP32 values within |y - y64| <= 1e-5 * sum_j |a_ij x_j|
(the bound of every value check here)."""
import ctypes
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


def bucket(M, key, val, buckets=0):
    dk = sa.DeviceArray.from_host(key if key.size else np.zeros(1, np.int32), np.int32)
    dv = sa.DeviceArray.from_host(val if val.size else np.zeros(1, np.float32), np.float32)
    M.A.set_sparse_x(dk, dv, int(key.size))
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    c = ctypes.c_int(-1)
    rc = sa.lib().spmspv_csc_bucket(M.A.h, 1.0, ctypes.byref(c), yk.ptr, yv.ptr, buckets)
    assert rc == 0, f"spmspv_csc_bucket returned {rc}"
    return yk.to_host(c.value), yv.to_host(c.value)


@pytest.fixture(scope="module")
def rmat16(gpu):
    """R-MAT scale 16 (65,536 rows, 1 M nonzeros): bins of equal nonzero
    counts are a few rows long at the hub rows and thousands of rows long in
    the tail."""
    rp, col, val = sa.gen_rmat_csr(16, 16, seed=31)
    n = 1 << 16
    M = sa.Matrix(n, n, rp, col, val)
    return M, oracle.csr_to_csc(n, n, rp, col, val)


@pytest.mark.parametrize("buckets", [2, 16, 256, 1024])
@pytest.mark.parametrize("density", [1e-3, 0.05, 1.0])
def test_bins_and_chunks(rmat16, buckets, density):
    M, (cp, ri, cv) = rmat16
    k = max(1, int(round(density * M.n)))
    key, val = sa.gen_sparse_vector(M.n, k, 900 + k)
    say(f"bucket slot, rmat16, {buckets} bins, xnnz {k}")
    yk, yv = bucket(M, key, val, buckets)
    check(M.m, cp, ri, cv, key, val, yk, yv, f"{buckets} bins density {density}")


def test_many_rows_per_bin_limit(gpu):
    """2^20 rows: at most 2^13 rows per bin gives 128 bins at least, whatever
    the requested count (2 here); one hub row of 150,000 nonzeros is a bin of
    its own, which a dense x fills with 150,000 products: three chunks whose
    row sums meet in FP32 atomics."""
    import scipy.sparse as sp
    rng = np.random.default_rng(12)
    m = n = 1 << 20
    rows = np.concatenate([rng.integers(0, m, 400000), np.full(150000, 777)])
    cols = np.concatenate([rng.integers(0, n, 400000), rng.choice(n, 150000, replace=False)])
    vals = rng.uniform(-1, 1, rows.size).astype(np.float32)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    M = sa.Matrix(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32))
    cp, ri, cv = oracle.csr_to_csc(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32),
                                   A.data.astype(np.float32))
    for k in (1000, 100000, n):
        key, val = sa.gen_sparse_vector(n, k, 40 + k)
        for b in (2, 256):
            say(f"bucket slot, 2^20 rows, {b} bins requested, xnnz {k}")
            yk, yv = bucket(M, key, val, b)
            check(m, cp, ri, cv, key, val, yk, yv, f"2^20 rows xnnz {k} bins {b}")


def test_zero_products_mark_rows(rmat16):
    """x values of +0.0 and -0.0 (products of either sign of zero) still put
    their rows into the result (structural support), both in bins of a few
    hub rows (summed in registers) and in long bins (summed in the LDS
    window, where a row's sum starts at -0.0)."""
    M, (cp, ri, cv) = rmat16
    for it, (k, mode) in enumerate([(3000, "mixed"), (65536, "mixed"), (20000, "all -0.0"), (65536, "all +0.0")]):
        key, val = sa.gen_sparse_vector(M.n, k, 8100 + it)
        val = val.copy()
        if mode == "mixed":
            val[::2] = 0.0
            val[1::4] = -0.0
        else:
            val[:] = -0.0 if mode == "all -0.0" else 0.0
        say(f"bucket slot, zero x values ({mode}), xnnz {k}")
        yk, yv = bucket(M, key, val, 256)
        check(M.m, cp, ri, cv, key, val, yk, yv, f"zero x values {mode} xnnz {k}")


def test_csc_only_handle_builds_bins(rmat16):
    """A handle given only the CSC (no row pointer) counts the rows' nonzeros
    from the CSC's row indices to lay out its bins; results match the
    oracle for a few bin counts."""
    M, (cp, ri, cv) = rmat16
    m, n, nnz = M.m, M.n, M.nnz
    d = (sa.DeviceArray.from_host(cp, np.int32), sa.DeviceArray.from_host(ri, np.int32),
         sa.DeviceArray.from_host(cv, np.float32))
    A = sa.SpmspvHandle(m, n, nnz)
    try:
        A.InputCSC(nnz, d[1], d[0], d[2])
        A.allocPreBuffer()
        for it, (k, b) in enumerate([(3000, 256), (40000, 1024), (65536, 2)]):
            key, val = sa.gen_sparse_vector(n, k, 8700 + it)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(val, np.float32)
            A.set_sparse_x(dk, dv, k)
            yk = sa.DeviceArray(m, np.int32)
            yv = sa.DeviceArray(m, np.float32)
            c = ctypes.c_int(-1)
            say(f"bucket slot on a CSC-only handle, {b} bins, xnnz {k}")
            assert sa.lib().spmspv_csc_bucket(A.h, 1.0, ctypes.byref(c), yk.ptr, yv.ptr, b) == 0
            check(m, cp, ri, cv, key, val, yk.to_host(c.value), yv.to_host(c.value), f"CSC only, {b} bins")
    finally:
        A.Destroy()


def test_inf_and_nan_products(rmat16):
    """An x value of +Inf (and one of NaN) makes its column's products Inf or
    NaN: the chunks holding them sum in FP32 instead of fixed point, and
    every row's value is the oracle's (Inf, NaN and finite values alike);
    bins without such products keep their fixed-point sums."""
    M, (cp, ri, cv) = rmat16
    n = M.n
    key, val = sa.gen_sparse_vector(n, 20000, 8800)
    val = val.copy()
    full = np.nonzero(cp[key + 1] > cp[key])[0]  # x entries whose column holds nonzeros
    val[full[5]] = np.inf
    val[full[9000]] = np.nan
    say("bucket slot, Inf and NaN in x")
    yk, yv = bucket(M, key, val, 256)
    ek, ev, eb = oracle.spmspv_f64(M.m, cp, ri, cv, key, val)
    np.testing.assert_array_equal(yk, ek)
    y = yv.astype(np.float64)
    assert np.array_equal(np.isnan(y), np.isnan(ev)), "NaN rows differ"
    fin = np.isfinite(ev)
    assert np.array_equal(np.isinf(y) & ~np.isnan(ev), np.isinf(ev)), "Inf rows differ"
    inf = np.isinf(ev)
    assert np.array_equal(np.sign(y[inf]), np.sign(ev[inf]))
    assert np.all(np.abs(y[fin] - ev[fin]) <= TOL * eb[fin] + 1e-30)
    assert fin.sum() > 0 and (~fin).sum() > 0


def test_repeat_calls_leave_counts_clear(rmat16):
    """The per-bin counts, the accumulator and the marks return to zero
    after every call: alternating sizes give the oracle's result each time."""
    M, (cp, ri, cv) = rmat16
    for it, k in enumerate([50, 40000, 3, 65536, 700]):
        key, val = sa.gen_sparse_vector(M.n, k, 5000 + it)
        yk, yv = bucket(M, key, val)
        check(M.m, cp, ri, cv, key, val, yk, yv, f"repeat {it}")


def test_matrix_change_moves_bins(rmat16):
    """A new CSR with the same nnz (the rows in reverse order, so that the
    hub rows fall into the last bins) re-derives every bin's product slots:
    alternating the two matrices gives the oracle's result on each."""
    M, (cp, ri, cv) = rmat16
    m = M.m
    rp, col, val = M.h_row_ptr, M.h_col, M.h_val
    lens = np.diff(rp)[::-1]
    rp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    # entries of old row m-1-i, in order, become row i
    starts = rp[:-1][::-1]
    order = np.repeat(starts - rp2[:-1], lens) + np.arange(M.nnz)
    col2, val2 = col[order], val[order]
    ref2 = oracle.csr_to_csc(m, M.n, rp2, col2, val2)
    d2 = (sa.DeviceArray.from_host(rp2, np.int32), sa.DeviceArray.from_host(col2, np.int32),
          sa.DeviceArray.from_host(val2, np.float32))
    d1 = (M.d_row_ptr, M.d_col, M.d_val)
    try:
        for it, (d, ref) in enumerate([(d2, ref2), (d1, (cp, ri, cv)), (d2, ref2)]):
            M.A.InputCSR(M.nnz, *d)
            M.A.ToCSC()
            key, xv = sa.gen_sparse_vector(M.n, 20000, 300 + it)
            say(f"bucket slot after a matrix change, step {it}")
            yk, yv = bucket(M, key, xv, 256)
            check(m, *ref, key, xv, yk, yv, f"matrix switch {it}")
    finally:
        M.A.InputCSR(M.nnz, *d1)
        M.A.ToCSC()


def test_rule_routes_auto_calls(rmat16):
    """With the rule's range over every size, an auto call whose network
    choice is 2, 3, 5 or 7 runs the bucket slot (and gives the oracle's
    result); with the rule off the bucket slot never runs."""
    M, (cp, ri, cv) = rmat16
    A = M.A
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    try:
        # the rule fallback's choice (mode 1) is what the rule sees; a pinned
        # slot (mode 2) bypasses the rule, and so does the single-launch
        # small-x path, which is off here
        A.set_small_x_path(False)
        A.selector_mode(1)
        A.set_bucket_range(0, 1 << 62)
        for k in (700, 6000, 30000, 65536):
            key, val = sa.gen_sparse_vector(M.n, k, 6100 + k)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(val, np.float32)
            A.set_sparse_x(dk, dv, k)
            A.selector_mode(1)
            plain = A.select()
            c, used = A.run_auto(y_key=yk, y_val=yv)
            say(f"rule on: xnnz {k} rule fallback picks {plain}, ran {used}")
            expect = sa.K_COL_BUCKET if plain in (2, 3, 5, 7) else plain
            assert used == expect
            check(M.m, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"rule on xnnz {k}")
        A.set_bucket_range(0, -1)
        for k in (6000, 65536):
            key, val = sa.gen_sparse_vector(M.n, k, 6100 + k)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(val, np.float32)
            A.set_sparse_x(dk, dv, k)
            c, used = A.run_auto(y_key=yk, y_val=yv)
            assert used != sa.K_COL_BUCKET
            check(M.m, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"rule off xnnz {k}")
    finally:
        A.set_bucket_range(0, -1)
        A.set_small_x_path(True)


def test_speculated_bucket_calls(rmat16):
    """Auto calls that the rule sends to the bucket slot are queued behind
    the selector once their size class has chosen it: repeated calls guess
    right, a call whose choice changed (rule switched off) guesses wrong and
    runs the chosen slot, a dense output also works, and every result is the
    oracle's."""
    M, (cp, ri, cv) = rmat16
    A = M.A
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    yd = sa.DeviceArray(M.m, np.float32)
    k = 30000
    key, val = sa.gen_sparse_vector(M.n, k, 9100)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    try:
        A.set_small_x_path(False)
        A.selector_mode(1)
        A.set_sparse_x(dk, dv, k)
        plain = A.select()
        if plain not in (2, 3, 5, 7):
            pytest.skip(f"the rule fallback picks slot {plain}, which the bucket rule leaves alone")
        A.set_bucket_range(0, 1 << 62)
        h0, m0 = A.speculation_stats()
        for it in range(4):
            A.set_sparse_x(dk, dv, k)
            c, used = A.run_auto(y_key=yk, y_val=yv)
            say(f"speculated bucket call {it}: ran {used}")
            assert used == sa.K_COL_BUCKET
            check(M.m, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"speculated call {it}")
        h1, m1 = A.speculation_stats()
        assert h1 - h0 >= 3, (h0, h1)
        # a dense result through the speculated bucket slot
        A.set_sparse_x(dk, dv, k)
        c, used = A.run_auto(y_dense=yd)
        assert used == sa.K_COL_BUCKET
        ek, ev, eb = oracle.spmspv_f64(M.m, cp, ri, cv, key, val)
        dense = yd.to_host(M.m)
        assert np.count_nonzero(np.delete(dense, ek)) == 0
        assert np.all(np.abs(dense[ek].astype(np.float64) - ev) <= TOL * eb + 1e-30)
        A.set_bucket_range(0, -1)
        A.set_sparse_x(dk, dv, k)
        c, used = A.run_auto(y_key=yk, y_val=yv)
        assert used == plain
        h2, m2 = A.speculation_stats()
        assert m2 - m1 >= 1, (m1, m2)
        check(M.m, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), "after a wrong guess")
    finally:
        A.set_bucket_range(0, -1)
        A.set_small_x_path(True)


def test_new_matrix_turns_rule_off(rmat16):
    """The bucket rule's range belongs to the matrix it was tuned on:
    InputCSR turns it off."""
    M, _ = rmat16
    A = M.A
    try:
        A.set_bucket_range(1000, 10 ** 9)
        assert A.tuned()["bucket_rule"] == [1000, 10 ** 9]
        A.InputCSR(M.nnz, M.d_row_ptr, M.d_col, M.d_val)
        A.ToCSC()
        assert A.tuned()["bucket_rule"] == [0, -1]
    finally:
        A.set_bucket_range(0, -1)


def test_autotune_with_too_many_rows_for_bins(gpu):
    """2^25 rows need more than the bucket slot's 6,144 bins of at most
    4,096 rows: the autotuner leaves the slot and its rule off and tunes the
    rest; an explicit bucket call says INVALID_ARGUMENT; auto calls still
    match the oracle."""
    rng = np.random.default_rng(21)
    m = n = 1 << 25
    nz = 200000
    rows = np.sort(rng.integers(0, m, nz))
    cols = rng.integers(0, n, nz).astype(np.int32)
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    order = np.lexsort((cols, rows))
    col = cols[order]
    val = rng.uniform(-1, 1, nz).astype(np.float32)
    M = sa.Matrix(m, n, rp, col, val)
    A = M.A
    key = np.arange(0, n, 7, dtype=np.int32)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(np.ones(key.size, np.float32), np.float32)
    A.set_sparse_x(dk, dv, key.size)
    say("autotune on 2^25 rows")
    tuned = A.autotune(reps=1)
    assert tuned["bucket_rule"] == [0, -1]
    # a range set by hand is refused: the slot cannot serve this matrix
    assert sa.lib().spmspv_set_bucket_range(A.h, 0, 1 << 62) == -20
    assert A.tuned()["bucket_rule"] == [0, -1]
    yk = sa.DeviceArray(m, np.int32)
    yv = sa.DeviceArray(m, np.float32)
    c = ctypes.c_int(-1)
    assert sa.lib().spmspv_csc_bucket(A.h, 1.0, ctypes.byref(c), yk.ptr, yv.ptr, 0) == -20
    cp, ri, cv = oracle.csr_to_csc(m, n, rp, col, val)
    k = 5000
    xk, xv = sa.gen_sparse_vector(n, k, 2121)
    A.set_sparse_x(sa.DeviceArray.from_host(xk, np.int32), sa.DeviceArray.from_host(xv, np.float32), k)
    cnt, used = A.run_auto(y_key=yk, y_val=yv)
    assert used != sa.K_COL_BUCKET
    check(m, cp, ri, cv, xk, xv, yk.to_host(cnt), yv.to_host(cnt), "2^25 rows, auto")


def test_autotuned_rule_then_parity(rmat16):
    """The autotuner's range is off or ordered; auto calls across densities
    match the oracle whichever slot the rule sends them to."""
    M, (cp, ri, cv) = rmat16
    A = M.A
    key = np.arange(M.n, dtype=np.int32)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(np.ones(M.n, np.float32), np.float32)
    A.set_sparse_x(dk, dv, M.n)
    tuned = A.autotune(reps=2)
    lo, hi = tuned["bucket_rule"]
    say(f"autotuned bucket rule {lo}..{hi}")
    assert (lo, hi) == (0, -1) or 0 < lo <= hi
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    A.selector_mode(1)
    for d in (1e-3, 1e-2, 0.1, 0.5, 1.0):
        k = max(1, int(round(d * M.n)))
        key, val = sa.gen_sparse_vector(M.n, k, 7100 + k)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(val, np.float32)
        A.set_sparse_x(dk, dv, k)
        c, used = A.run_auto(y_key=yk, y_val=yv)
        say(f"autotuned rule: density {d} ran {used}")
        check(M.m, cp, ri, cv, key, val, yk.to_host(c), yv.to_host(c), f"autotuned density {d}")
    A.set_bucket_range(0, -1)
