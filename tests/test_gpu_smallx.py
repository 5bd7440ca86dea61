"""The single-launch small-x column path and the sampled selector inputs of
a large x, through the C-ABI.

Every column slot with the path on and off, at x sizes on both sides of its
limits (4,096 keys; 11,264 products), with count-only and dense outputs
(which do not take it), through spmspv_run and spmspv_run_auto; index sets
bit-exact and values within 1e-5 * sum|a_ij x_j| of the FP64 oracle. For a
large x, the features-only selector launch (sampled inputs) and the scan
launch choose the same slot, and a column slot run after it is exact."""
import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5
CAP = 256 * 44


def check(m, csc, key, val, yk, yv, tag):
    cp, ri, cv = csc
    ek, ev, eb = oracle.spmspv_f64(m, cp, ri, cv, key, val)
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    bad = np.nonzero(np.abs(yv.astype(np.float64) - ev) > TOL * eb + 1e-30)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound"
    return ek, ev, eb


@pytest.fixture(scope="module")
def uni(gpu):
    """Uniform n = 16,384 with exactly 16 rows per column: bin_len = 16 xnnz,
    so 704 keys give exactly the product limit and 705 one past it."""
    n, k = 1 << 14, 16
    cp, ri, cv = sa.gen_uniform_csc(n, k, 0x5A11)
    rp, col, val = oracle.csr_to_csc(n, n, cp, ri, cv)
    M = sa.Matrix(n, n, rp, col, val)
    return n, M, (cp, ri, cv)


@pytest.mark.parametrize("xnnz", [0, 1, 17, 703, 704, 705, 4096])
@pytest.mark.parametrize("slot", [0, 1, 2, 3])
def test_fixed_slots_both_paths(uni, xnnz, slot):
    n, M, csc = uni
    key, val = sa.gen_sparse_vector(n, xnnz, 0x77 + xnnz)
    dk = sa.DeviceArray.from_host(key if xnnz else np.zeros(1, np.int32), np.int32)
    dv = sa.DeviceArray.from_host(val if xnnz else np.zeros(1, np.float32), np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    yd = sa.DeviceArray(n, np.float32)
    for on in (True, False):
        M.A.set_small_x_path(on)
        tag = f"slot {slot} xnnz {xnnz} path {'on' if on else 'off'}"
        print(tag, flush=True)
        M.A.set_sparse_x(dk, dv, xnnz)
        c = M.A.run(slot, y_key=yk, y_val=yv)
        ek, ev, eb = check(n, csc, key, val, yk.to_host(c), yv.to_host(c), tag)
        # count only (no outputs given: internal buffers), then dense output
        M.A.set_sparse_x(dk, dv, xnnz)
        assert M.A.run(slot) == ek.size, tag + " count only"
        M.A.set_sparse_x(dk, dv, xnnz)
        assert M.A.run(slot, y_dense=yd) == ek.size, tag + " dense"
        d = yd.to_host()
        dense = np.zeros(n)
        dense[ek] = ev
        bd = np.zeros(n)
        bd[ek] = eb
        assert np.all(np.abs(d - dense) <= TOL * bd + 1e-30), tag + " dense values"
    M.A.set_small_x_path(True)


def test_run_auto_small_and_after(uni):
    """run_auto on x of growing size (the small ones end in one launch,
    reported as slot 1), each followed by fixed row and column slots on the
    same x, the handle's scratch state staying consistent."""
    n, M, csc = uni
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    M.A.selector_mode(1)
    for rep in range(2):
        for xnnz in (1, 5, 300, 704, 705, 2000, 4096, 4097, 9000):
            key, val = sa.gen_sparse_vector(n, xnnz, 0x99 + xnnz + rep)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(val, np.float32)
            M.A.set_sparse_x(dk, dv, xnnz)
            c, used = M.A.run_auto(y_key=yk, y_val=yv)
            tag = f"auto xnnz {xnnz} rep {rep} slot {used}"
            print(tag, flush=True)
            check(n, csc, key, val, yk.to_host(c), yv.to_host(c), tag)
            if xnnz * 16 <= CAP:
                assert used == sa.K_COL_NAIVE_SORT, tag
            for slot in (sa.K_ROW_LB, sa.K_COL_LB_ATOMIC, sa.K_SPMV_LB):
                c2 = M.A.run(slot, y_key=yk, y_val=yv)
                check(n, csc, key, val, yk.to_host(c2), yv.to_host(c2), f"{tag} then slot {slot}")


def test_sampled_selector_inputs(gpu):
    """x of more than 65,536 keys: the features-only launch (sampled inputs,
    estimated bin_len) picks what the scan launch picks, and a column slot
    run after it (whose bin_len must be exact) is exact."""
    import os
    scale = 17
    n = 1 << scale
    rp, col, val = sa.gen_rmat_csr(scale, 16, seed=91)
    M = sa.Matrix(n, n, rp, col, val)
    w = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd",
                             "selector_weights.npy")).astype(np.float32)
    M.A.selector_set_weights(w)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    csc = (rp, col, val)  # symmetric: the CSR is the CSC
    agree = 0
    # x sizes of two size classes: after the first call of a class has chosen
    # a row slot, the next ones take the features-only (sampled) launch
    for k, xnnz in enumerate([66000, 70000, 66001, 70001, 120000, 100000, 65537, 65538]):
        key, xv = sa.gen_sparse_vector(n, xnnz, 0xABC + k)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(xv, np.float32)
        M.A.set_sparse_x(dk, dv, xnnz)
        want = M.A.select()
        M.A.set_sparse_x(dk, dv, xnnz)
        c, used = M.A.run_auto(y_key=yk, y_val=yv)
        print(f"xnnz {xnnz}: scan launch {want}, dispatched {used}", flush=True)
        check(n, csc, key, xv, yk.to_host(c), yv.to_host(c), f"auto xnnz {xnnz}")
        agree += int(used == want)
        for slot in (sa.K_COL_LB_SORT, sa.K_COL_LB_ATOMIC):
            c2 = M.A.run(slot, y_key=yk, y_val=yv)
            check(n, csc, key, xv, yk.to_host(c2), yv.to_host(c2), f"xnnz {xnnz} then slot {slot}")
    assert agree == 8


def test_autotuned_thresholds_then_parity(uni):
    """The autotuner's bucket count and single-workgroup thresholds stay in
    their ranges, and every column slot is still exact afterwards."""
    n, M, csc = uni
    key, val = sa.gen_sparse_vector(n, 3000, 0x1234)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    M.A.set_sparse_x(dk, dv, key.size)
    tuned = M.A.autotune(reps=2)
    print(tuned, flush=True)
    assert tuned["buckets"] in (256, 512, 1024)
    assert 0 <= tuned["small_x_products"] <= CAP
    assert 0 < tuned["one_workgroup_sort"] <= 3072
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    for xnnz in (50, 300, 3000):
        for slot in (0, 1, 2, 3, 8):
            M.A.set_sparse_x(dk, dv, xnnz)
            c = M.A.run(slot, y_key=yk, y_val=yv)
            check(n, csc, key[:xnnz], val[:xnnz], yk.to_host(c), yv.to_host(c), f"tuned slot {slot} xnnz {xnnz}")


def test_two_layer_tiles_spanning_empty_columns(gpu):
    """The two-layer kernels when one product tile spans more x entries
    than its LDS stage holds (20,000 mostly empty columns): each product
    then searches the column scan in memory."""
    import scipy.sparse as sp
    m, n = 5000, 20000
    rng = np.random.default_rng(3)
    cols = np.repeat(np.arange(0, n, 997), 3)
    rows = rng.integers(0, m, cols.size)
    A = sp.csr_matrix((rng.uniform(-1, 1, cols.size).astype(np.float32), (rows, cols)), shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    M = sa.Matrix(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32))
    csc = oracle.csr_to_csc(m, n, A.indptr, A.indices, A.data)
    key = np.arange(n, dtype=np.int32)
    val = rng.uniform(-1, 1, n).astype(np.float32)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(m, np.int32)
    yv = sa.DeviceArray(m, np.float32)
    for slot in (sa.K_COL_TWO_LAYER_ATOMIC, sa.K_COL_TWO_LAYER_SORT):
        M.A.set_sparse_x(dk, dv, n)
        c = M.A.run(slot, y_key=yk, y_val=yv)
        check(m, csc, key, val, yk.to_host(c), yv.to_host(c), f"two-layer slot {slot}")
