"""Parity of every kernel slot with the CPU oracle, through the C-ABI.

Index sets must match bit-exactly (structural support, sorted ascending);
FP32 values must satisfy |y - y64| <= 1e-5 * sum_j |a_ij x_j| against the
FP64 oracle (the tolerance north of every value check below)."""
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


def csc_of(m, n, rp, col, val):
    return oracle.csr_to_csc(m, n, rp, col, val)


def check_result(m, cp, ri, cv, key, val, yk, yv, tag):
    ek, ev, eb = oracle.spmspv_f64(m, cp, ri, cv, key, val)
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    err = np.abs(yv.astype(np.float64) - ev)
    lim = TOL * eb + 1e-30
    bad = np.nonzero(err > lim)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound, first {bad[:5]}, err {err[bad[:5]]} lim {lim[bad[:5]]}"
    return float((err / np.maximum(eb, 1e-300)).max()) if err.size else 0.0


def run_slot(M, k, key, val, want_dense=False):
    dk = sa.DeviceArray.from_host(key if key.size else np.zeros(1, np.int32), np.int32)
    dv = sa.DeviceArray.from_host(val if val.size else np.zeros(1, np.float32), np.float32)
    M.A.set_sparse_x(dk, dv, int(key.size))
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    yd = sa.DeviceArray(M.m, np.float32) if want_dense else None
    nnz = M.A.run(k, y_dense=yd, y_key=yk, y_val=yv)
    out = (yk.to_host(nnz), yv.to_host(nnz), yd.to_host() if yd else None)
    return out


SLOTS = list(range(11))  # the eight selector slots, the bucket kernel and the two-layer pair


@pytest.fixture(scope="module")
def c1(gpu):
    """Config 1: uniform n=65,536, 16 distinct rows per column, U[-1,1];
    x with 655 distinct keys, U[-1,1]; seed 0x5EED0001."""
    n, k = 65536, 16
    cp, ri, cv = sa.gen_uniform_csc(n, k, 0x5EED0001)
    # CSR of the same matrix (the handle takes CSR like the reference)
    import scipy.sparse as sp
    C = sp.csc_matrix((cv, ri, cp), shape=(n, n))
    R = C.tocsr()
    R.sort_indices()
    M = sa.Matrix(n, n, R.indptr.astype(np.int32), R.indices.astype(np.int32), R.data.astype(np.float32))
    key, val = sa.gen_sparse_vector(n, 655, 0x5EED0001)
    return M, (cp, ri, cv), key, val


@pytest.mark.parametrize("slot", SLOTS)
def test_c1_every_slot(c1, slot):
    M, (cp, ri, cv), key, val = c1
    say(f"c1 slot {slot} ({sa.KERNEL_NAMES[slot]})")
    yk, yv, _ = run_slot(M, slot, key, val)
    err = check_result(M.m, cp, ri, cv, key, val, yk, yv, f"c1 slot {slot}")
    say(f"  ynnz {yk.size} max rel err {err:.3e}")


def test_c1_tocsc_bit_exact(c1):
    M, (cp, ri, cv), _, _ = c1
    say("c1 ToCSC vs oracle transpose")
    ocp, ori, ocv = oracle.csr_to_csc(M.m, M.n, M.h_row_ptr, M.h_col, M.h_val)
    np.testing.assert_array_equal(ocp, cp)
    # the handle's CSC is internal; check it through a unit vector per column
    for j in (0, 1, 777, 65535):
        yk, yv, _ = run_slot(M, sa.K_COL_NAIVE_SORT, np.array([j], np.int32), np.array([1.0], np.float32))
        np.testing.assert_array_equal(yk, ori[ocp[j]:ocp[j + 1]])
        np.testing.assert_array_equal(yv, ocv[ocp[j]:ocp[j + 1]])


def test_c1_vec_features_exact(c1):
    M, (cp, ri, cv), key, val = c1
    say("c1 vec features")
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    M.A.set_sparse_x(dk, dv, key.size)
    b, mx, mn = M.A.computeVecFeature()
    ref = oracle.vec_features(cp, key)
    assert (b, mx, mn) == tuple(int(v) for v in ref)
    gm = M.A.computeGM()
    np.testing.assert_allclose(gm, oracle.gm(M.m, M.nnz, key.size, b))


def test_c1_dense_output(c1):
    M, (cp, ri, cv), key, val = c1
    for slot in SLOTS:
        say(f"c1 dense output slot {slot}")
        yk, yv, yd = run_slot(M, slot, key, val, want_dense=True)
        check_result(M.m, cp, ri, cv, key, val, yk, yv, f"dense slot {slot}")
        dense = np.zeros(M.m, np.float32)
        dense[yk] = yv
        np.testing.assert_array_equal(yd, dense)


def build(m, n, rp, col, val):
    return sa.Matrix(m, n, np.asarray(rp, np.int32), np.asarray(col, np.int32), np.asarray(val, np.float32))


@pytest.fixture(scope="module")
def rmat12(gpu):
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=20)
    n = 1 << 12
    M = build(n, n, rp, col, val)
    return M, csc_of(n, n, rp, col, val)


@pytest.mark.parametrize("density", [1e-4, 1e-3, 1e-2, 0.1, 0.5, 1.0])
def test_rmat_density_sweep(rmat12, density):
    M, (cp, ri, cv) = rmat12
    k = max(1, int(round(density * M.n)))
    key, val = sa.gen_sparse_vector(M.n, k, 77 + k)
    for slot in SLOTS:
        say(f"rmat12 density {density} xnnz {k} slot {slot}")
        yk, yv, _ = run_slot(M, slot, key, val)
        check_result(M.m, cp, ri, cv, key, val, yk, yv, f"rmat d={density} slot {slot}")


def test_edge_cases(gpu):
    # rectangular, empty rows/columns, exact cancellation, zero x values
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    m, n = 3000, 2000
    A = sp.random(m, n, density=0.002, format="csr", random_state=rng, dtype=np.float32)
    A.data = rng.uniform(-1, 1, A.nnz).astype(np.float32)
    A = A.tolil()
    A[10, :] = 0
    A[:, 7] = 0
    A[20, 3] = 1.0
    A[20, 4] = -1.0  # row 20 cancels exactly when x3 = x4 = 1
    A = A.tocsr()
    A.sort_indices()
    A.eliminate_zeros()
    M = build(m, n, A.indptr, A.indices, A.data)
    cp, ri, cv = csc_of(m, n, A.indptr, A.indices, A.data)
    cases = {
        "empty x": (np.zeros(0, np.int32), np.zeros(0, np.float32)),
        "empty column only": (np.array([7], np.int32), np.array([2.0], np.float32)),
        "cancellation": (np.array([3, 4], np.int32), np.array([1.0, 1.0], np.float32)),
        "zero x values": (np.array([3, 4, 5, 100], np.int32), np.zeros(4, np.float32)),
        "full x": (np.arange(n, dtype=np.int32), rng.uniform(-1, 1, n).astype(np.float32)),
        "unsorted keys": (np.array([1999, 5, 1000, 3], np.int32), np.array([1, 2, 3, 4], np.float32)),
    }
    for name, (key, val) in cases.items():
        for slot in SLOTS:
            say(f"edge case '{name}' slot {slot}")
            yk, yv, _ = run_slot(M, slot, key, val)
            check_result(m, cp, ri, cv, key, val, yk, yv, f"{name} slot {slot}")


def test_hub_column_spans_many_tiles(gpu):
    # one column with 50,000 entries + long rows: load-balanced tiles and
    # the carry fix-ups must join partial sums across many tiles
    import scipy.sparse as sp
    rng = np.random.default_rng(8)
    m = n = 60000
    rows = np.concatenate([np.arange(50000), rng.integers(0, m, 20000), np.full(30000, 123)])
    cols = np.concatenate([np.full(50000, 9), rng.integers(0, n, 20000), rng.choice(n, 30000, replace=False)])
    vals = rng.uniform(-1, 1, rows.size).astype(np.float32)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    M = build(m, n, A.indptr, A.indices, A.data.astype(np.float32))
    cp, ri, cv = csc_of(m, n, A.indptr, A.indices, A.data.astype(np.float32))
    for key in (np.array([9], np.int32), np.array([9, 10, 11], np.int32), np.arange(0, n, 3, dtype=np.int32)):
        val = rng.uniform(-1, 1, key.size).astype(np.float32)
        for slot in SLOTS:
            say(f"hub xnnz {key.size} slot {slot}")
            yk, yv, _ = run_slot(M, slot, key, val)
            check_result(m, cp, ri, cv, key, val, yk, yv, f"hub slot {slot}")
    # the CSR5 tiles: row 123 runs through 59 tiles of 512 items at sigma 8
    # (its carries are joined in several batches), 20 at sigma 24
    for sigma in (8, 24):
        M.A.set_sigma(sigma)
        M.A.set_spmv_type(1)
        for key in (np.arange(0, n, 3, dtype=np.int32), np.arange(n, dtype=np.int32)):
            val = rng.uniform(-1, 1, key.size).astype(np.float32)
            for slot in (sa.K_ROW_LB, sa.K_SPMV_LB):
                say(f"hub csr5 sigma {sigma} xnnz {key.size} slot {slot}")
                yk, yv, _ = run_slot(M, slot, key, val)
                check_result(m, cp, ri, cv, key, val, yk, yv, f"hub csr5 s={sigma} slot {slot}")
    M.A.set_spmv_type(0)


def test_repeat_calls_keep_state_clean(rmat12):
    """The accumulator and bitmaps are reset by every call: results of a
    call must not depend on the previous one."""
    M, (cp, ri, cv) = rmat12
    for it in range(6):
        key, val = sa.gen_sparse_vector(M.n, 50 + 300 * it, 1000 + it)
        for slot in SLOTS:
            yk, yv, _ = run_slot(M, slot, key, val)
            check_result(M.m, cp, ri, cv, key, val, yk, yv, f"repeat {it} slot {slot}")
    say("repeat calls ok")


def test_mat_features_match_oracle(rmat12):
    M, _ = rmat12
    say("mat features")
    f = M.A.computeMatFeature()
    ref = oracle.mat_features(M.m, M.n, M.h_row_ptr)
    np.testing.assert_allclose(f, ref.astype(np.float32), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(M.A.extractMatFeature(), oracle.std_row_trunc(M.m, M.h_row_ptr), rtol=1e-5)


def test_conversions(rmat12):
    M, _ = rmat12
    say("conversions")
    n = M.n
    key, val = sa.gen_sparse_vector(n, 300, 4)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    dx = sa.DeviceArray(n, np.float32)
    M.A.set_sparse_x(dk, dv, key.size)
    M.A.set_x(dx)
    M.A.sparse2dense()
    ref = np.zeros(n, np.float32)
    ref[key] = val
    np.testing.assert_array_equal(dx.to_host(), ref)
    # dense -> sparse keeps nonzeros in order
    ok = sa.DeviceArray(n, np.int32)
    ov = sa.DeviceArray(n, np.float32)
    cnt = M.A.dense2sparse_outer(n, dx, ok, ov)
    np.testing.assert_array_equal(ok.to_host(cnt), key)
    np.testing.assert_array_equal(ov.to_host(cnt), val)
    back = sa.DeviceArray(n, np.float32)
    M.A.sparse2dense_outer(cnt, ok, ov, n, back)
    np.testing.assert_array_equal(back.to_host(), ref)
    # unbind the borrowed dense buffer before it is freed
    M.A.set_x(None)


def test_selector_tree_mode_and_run_auto(rmat12):
    M, (cp, ri, cv) = rmat12
    M.A.selector_mode(1)
    for k in (3, 400, 4096):
        key, val = sa.gen_sparse_vector(M.n, k, 5 + k)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(val, np.float32)
        M.A.set_sparse_x(dk, dv, key.size)
        yk = sa.DeviceArray(M.m, np.int32)
        yv = sa.DeviceArray(M.m, np.float32)
        nnz, used = M.A.run_auto(y_key=yk, y_val=yv)
        say(f"run_auto xnnz {k} -> slot {used}")
        assert 0 <= used < 8
        check_result(M.m, cp, ri, cv, key, val, yk.to_host(nnz), yv.to_host(nnz), f"auto {k}")
        f = M.A.features()
        assert np.all(np.isfinite(f))


def test_fused_mlp_selection_matches_oracle(rmat12):
    """The MLP inference fused into the scan kernel (committed weights)
    picks the slot the FP64 oracle network picks on the device-computed
    features (near-ties, top-2 logit margin < 1e-3, are skipped)."""
    import os
    w = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd",
                             "selector_weights.npy")).astype(np.float32)
    M, _ = rmat12
    M.A.selector_set_weights(w)
    M.A.selector_mode(0)
    F = sa.MLP_F
    mean, istd = w[sa.MLP_NPARAM:sa.MLP_NPARAM + F].astype(np.float64), w[sa.MLP_NPARAM + F:].astype(np.float64)
    checked = 0
    try:
        for k in (1, 3, 40, 400, 1500, 4096):
            key, val = sa.gen_sparse_vector(M.n, k, 70 + k)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(val, np.float32)
            M.A.set_sparse_x(dk, dv, k)
            used = M.A.select()
            f = M.A.features().astype(np.float64)
            z = oracle.mlp_forward(F, 64, 64, 8, w[:sa.MLP_NPARAM].astype(np.float64),
                                   ((f - mean) * istd)[None, :])[0]
            top = np.sort(z)[::-1]
            say(f"xnnz {k}: device slot {used}, oracle slot {int(np.argmax(z))}, margin {top[0] - top[1]:.4f}")
            if top[0] - top[1] < 1e-3:
                continue
            assert used == int(np.argmax(z))
            checked += 1
    finally:
        M.A.selector_mode(1)
    assert checked >= 3


def test_cpp_handle_adaptive_app(gpu):
    """The reference's adaptive-app flow written against the C++
    SpmspvHandle shim (examples/adaptive_app.cpp), checked against the
    reference's all-ones known-answer pattern inside the program."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(sa.LIB_PATH), "adaptive_app")
    out = subprocess.run([exe, "12"], capture_output=True, text=True, timeout=120)
    say(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("RESULT OK") == 6


@pytest.mark.parametrize("sigma", [8, 16, 24, 32])
def test_csr5_row_kernels(rmat12, sigma):
    """K5/K7 on the CSR5 tiles (spmv_type 1) for every instantiated sigma:
    sparse output, dense output with and without support tracking, every x
    density; then back to the merge-path tiles."""
    M, (cp, ri, cv) = rmat12
    M.A.set_sigma(sigma)
    M.A.set_spmv_type(1)
    try:
        for k in (1, 40, 400, M.n):
            key, val = sa.gen_sparse_vector(M.n, k, 600 + k + sigma)
            for slot in (sa.K_ROW_LB, sa.K_SPMV_LB):
                say(f"csr5 sigma {sigma} xnnz {k} slot {slot}")
                yk, yv, yd = run_slot(M, slot, key, val, want_dense=True)
                check_result(M.m, cp, ri, cv, key, val, yk, yv, f"csr5 s={sigma} k={k} slot {slot}")
                dense = np.zeros(M.m, np.float32)
                dense[yk] = yv
                np.testing.assert_array_equal(yd, dense)
                yd2 = sa.DeviceArray(M.m, np.float32)
                M.A.run(slot, y_dense=yd2, want_nnz=False)
                np.testing.assert_array_equal(yd2.to_host(), dense)
    finally:
        M.A.set_spmv_type(0)
        M.A.set_sigma(-1)


def test_csr5_edge_shapes(gpu):
    """CSR5 with empty leading/trailing rows, rectangular shape, a row that
    spans many tiles, a single nonzero, and the reference-named entry
    points (csr5spmv / holaspmv)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(21)
    cases = []
    m, n = 5000, 3000
    A = sp.random(m, n, density=0.003, format="lil", random_state=rng, dtype=np.float32)
    A[:50, :] = 0
    A[m - 70:, :] = 0
    A[777, :] = rng.uniform(-1, 1, n).astype(np.float32)  # spans several tiles
    cases.append(("rect+empty+long", A.tocsr()))
    B = sp.lil_matrix((300, 300), dtype=np.float32)
    B[123, 45] = 2.5
    cases.append(("single nonzero", B.tocsr()))
    for name, A in cases:
        A.sort_indices()
        A.eliminate_zeros()
        m, n = A.shape
        M = build(m, n, A.indptr, A.indices, A.data.astype(np.float32))
        cp, ri, cv = csc_of(m, n, A.indptr, A.indices, A.data.astype(np.float32))
        for sigma in (8, 32):
            M.A.set_sigma(sigma)
            M.A.set_spmv_type(1)
            for k in (1, n // 3, n):
                key, val = sa.gen_sparse_vector(n, k, 31 + k)
                for slot in (sa.K_ROW_LB, sa.K_SPMV_LB):
                    say(f"csr5 {name} sigma {sigma} xnnz {k} slot {slot}")
                    yk, yv, yd = run_slot(M, slot, key, val, want_dense=True)
                    check_result(m, cp, ri, cv, key, val, yk, yv, f"{name} s={sigma} k={k} slot {slot}")
                    dense = np.zeros(m, np.float32)
                    dense[yk] = yv
                    np.testing.assert_array_equal(yd, dense)
        # reference-named entry points: csr5spmv runs CSR5 whatever the type
        M.A.set_spmv_type(0)
        key = np.arange(n, dtype=np.int32)
        val = rng.uniform(-1, 1, n).astype(np.float32)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(val, np.float32)
        M.A.set_sparse_x(dk, dv, n)
        y1 = sa.DeviceArray(m, np.float32)
        y2 = sa.DeviceArray(m, np.float32)
        assert sa.lib().spmspv_csr5spmv(M.A.h, 1.0, y1.ptr) == 0
        assert sa.lib().spmspv_holaspmv(M.A.h, 1.0, y2.ptr) == 0
        ref = (A.astype(np.float64) @ val.astype(np.float64))
        bound = (abs(A).astype(np.float64) @ np.abs(val).astype(np.float64))
        for y in (y1.to_host(), y2.to_host()):
            assert np.all(np.abs(y - ref) <= TOL * bound + 1e-30), name
        M.A.set_sigma(-1)


def test_autotune_then_parity(rmat12):
    """The autotuner changes team sizes and the row tile width per matrix;
    every slot must stay exact afterwards, for every candidate width."""
    M, (cp, ri, cv) = rmat12
    key, val = sa.gen_sparse_vector(M.n, 400, 99)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    M.A.set_sparse_x(dk, dv, key.size)
    rep = M.A.autotune(reps=2)
    say(f"autotune picked {rep}")
    assert rep["team_col"] in (2, 4, 8, 16, 32, 64) and rep["row_vt"] in (5, 7, 11)
    for slot in SLOTS:
        yk, yv, _ = run_slot(M, slot, key, val)
        check_result(M.m, cp, ri, cv, key, val, yk, yv, f"after autotune slot {slot}")
    lev = sa.DeviceArray(M.n, np.int32)
    M.A.bfs(0, lev, policy=5, use_graphs=1)
    ref, _ = oracle.bfs(M.n, cp, ri, 0)
    np.testing.assert_array_equal(lev.to_host(), ref)


def test_main_kernel_timing(rmat12):
    """With tile-kernel timing on, a K5/K7 call reports its tile kernel's
    duration (positive, within the call's own); a column slot's call
    reports none, not the previous call's."""
    M, (cp, ri, cv) = rmat12
    key, val = sa.gen_sparse_vector(M.n, 800, 123)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    M.A.time_main_kernel(True)
    try:
        for slot in (sa.K_ROW_LB, sa.K_COL_LB_ATOMIC, sa.K_SPMV_LB, sa.K_COL_NAIVE_SORT):
            M.A.set_sparse_x(dk, dv, key.size)
            M.A.timer_start()
            nnz = M.A.run(slot, y_key=yk, y_val=yv)
            call_ms = M.A.timer_stop()
            check_result(M.m, cp, ri, cv, key, val, yk.to_host(nnz), yv.to_host(nnz), f"timed slot {slot}")
            if slot in (sa.K_ROW_LB, sa.K_SPMV_LB):
                km = M.A.main_kernel_ms()
                say(f"slot {slot}: tile kernel {km:.4f} ms of {call_ms:.4f} ms")
                assert 0.0 < km <= call_ms
            else:
                with pytest.raises(sa.SpmspvError):
                    M.A.main_kernel_ms()
    finally:
        M.A.time_main_kernel(False)


def test_handle_holds_borrowed_arrays(gpu):
    """The handle borrows the device arrays it is given; the wrapper keeps
    them alive, so a handle taken from a Matrix that is then dropped still
    reads valid arrays (keeping only `.A` once freed the CSR under it)."""
    import gc
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=43)
    n = 1 << 12
    A = sa.Matrix(n, n, rp, col, val).A
    gc.collect()
    assert A._borrowed["csr"][0].ptr is not None and A._borrowed["csr"][0].ptr.value
    cp, ri, cv = csc_of(n, n, rp, col, val)
    key, xv = sa.gen_sparse_vector(n, 500, 17)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(xv, np.float32)
    A.set_sparse_x(dk, dv, key.size)
    del dk, dv
    gc.collect()
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    A.set_spmv_type(1)
    for slot in (sa.K_COL_LB_ATOMIC, sa.K_SPMV_LB):
        nnz = A.run(slot, y_key=yk, y_val=yv)
        check_result(n, cp, ri, cv, key, xv, yk.to_host(nnz), yv.to_host(nnz), f"dropped owner slot {slot}")
