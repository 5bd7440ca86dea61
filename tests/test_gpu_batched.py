"""Batched multi-vector SpMSpV (spmspv_run_batched): every vector's result
matches the FP64 oracle (index sets exact, values within 1e-5 * sum|a x|)
and the slot the batch picks for it is the one the single-vector selector
picks; vectors of very different density, empty vectors included."""
import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5


def check(m, cp, ri, cv, key, val, yk, yv, tag):
    ek, ev, eb = oracle.spmspv_f64(m, cp, ri, cv, key, val)
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    err = np.abs(yv.astype(np.float64) - ev)
    assert np.all(err <= TOL * eb + 1e-30), f"{tag}: values out of bound"


def make_batch(n, dens, seed):
    keys, vals = [], []
    for i, d in enumerate(dens):
        k = 0 if d == 0 else max(1, int(round(d * n)))
        key, val = sa.gen_sparse_vector(n, k, seed + i) if k else (np.zeros(0, np.int32), np.zeros(0, np.float32))
        keys.append(key)
        vals.append(val)
    vptr = np.zeros(len(dens) + 1, np.int32)
    vptr[1:] = np.cumsum([k.size for k in keys])
    return keys, vals, vptr


@pytest.mark.parametrize("graph", ["rmat12", "uniform"])
def test_batched_matches_oracle_and_selector(gpu, graph):
    if graph == "rmat12":
        rp, col, val = sa.gen_rmat_csr(12, 16, seed=31)
        n = 1 << 12
    else:
        n = 20000
        rp, col, val = sa.gen_random_csr(n, n, 12.0, 5)
    M = sa.Matrix(n, n, rp, col, val)
    cp, ri, cv = oracle.csr_to_csc(n, n, rp, col, val)
    dens = [0.0, 1.0 / n, 1e-3, 0.3, 1e-2, 0.0, 0.1, 1.0, 0.5, 2e-3, 0.05]
    keys, vals, vptr = make_batch(n, dens, 900)
    nvec = len(dens)
    yoff = np.arange(nvec + 1, dtype=np.int64) * n
    d_vptr = sa.DeviceArray.from_host(vptr, np.int32)
    d_vkey = sa.DeviceArray.from_host(np.concatenate(keys) if vptr[-1] else np.zeros(1, np.int32), np.int32)
    d_vval = sa.DeviceArray.from_host(np.concatenate(vals) if vptr[-1] else np.zeros(1, np.float32), np.float32)
    d_yoff = sa.DeviceArray.from_host(yoff, np.int64)
    d_yk = sa.DeviceArray(nvec * n, np.int32)
    d_yv = sa.DeviceArray(nvec * n, np.float32)
    d_yn = sa.DeviceArray(nvec, np.int32)
    for rep in range(2):  # the second batch reuses the handle's state
        used = M.A.run_batched(d_vptr, d_vkey, d_vval, d_yoff, d_yk, d_yv, d_yn)
        yn = d_yn.to_host()
        ykall = d_yk.to_host()
        yvall = d_yv.to_host()
        for v in range(nvec):
            o = int(yoff[v])
            check(n, cp, ri, cv, keys[v], vals[v], ykall[o:o + yn[v]], yvall[o:o + yn[v]],
                  f"{graph} rep {rep} vector {v} (density {dens[v]}, slot {used[v]})")
    # the batch's choice per vector equals the single-vector selector's
    for v in range(nvec):
        if keys[v].size == 0:
            continue
        dk = sa.DeviceArray.from_host(keys[v], np.int32)
        dv = sa.DeviceArray.from_host(vals[v], np.float32)
        M.A.set_sparse_x(dk, dv, keys[v].size)
        assert M.A.select() == used[v], f"vector {v}"


def test_batched_rejects_small_capacity(gpu):
    rp, col, val = sa.gen_rmat_csr(10, 8, seed=3)
    n = 1 << 10
    M = sa.Matrix(n, n, rp, col, val)
    keys, vals, vptr = make_batch(n, [0.5, 0.5], 77)
    d_vptr = sa.DeviceArray.from_host(vptr, np.int32)
    d_vkey = sa.DeviceArray.from_host(np.concatenate(keys), np.int32)
    d_vval = sa.DeviceArray.from_host(np.concatenate(vals), np.float32)
    d_yoff = sa.DeviceArray.from_host(np.array([0, 5, 10], np.int64), np.int64)
    d_yk = sa.DeviceArray(10, np.int32)
    d_yv = sa.DeviceArray(10, np.float32)
    d_yn = sa.DeviceArray(2, np.int32)
    with pytest.raises(sa.SpmspvError) as e:
        M.A.run_batched(d_vptr, d_vkey, d_vval, d_yoff, d_yk, d_yv, d_yn)
    assert e.value.code == -20 and "capacity" in str(e.value)


def test_batched_rmat24_full_size(gpu):
    """The batched configuration at full size: R-MAT scale 24, edge factor
    16, symmetrised (about 5.2e8 stored entries; the CSR is the CSC), the
    benchmark's 256 vectors (density log-uniform in [1e-4, 0.2], the same
    seeds). One spmspv_run_batched call against 256 sequential
    selector-dispatched calls: counts, slots and index sets identical for
    every vector, values within 1e-5 * sum|a x| of each other's oracle; 8
    vectors spanning the density range against the FP64 oracle."""
    scale, nvec = 24, 256
    n = 1 << scale
    print("generating R-MAT s24", flush=True)
    rp, col, val = sa.gen_rmat_csr(scale, 16, 0.57, 0.19, 0.19, seed=0x5EED0005)
    M = sa.Matrix(n, n, rp, col, val)
    import os
    w = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd", "selector_weights.npy")
    M.A.selector_set_weights(np.load(w).astype(np.float32))
    deg = np.diff(rp).astype(np.int64)
    rng = np.random.default_rng(0x5EED0006)
    dens = np.exp(rng.uniform(np.log(1e-4), np.log(0.2), nvec))
    keys, vals, bls = [], [], []
    for v in range(nvec):
        k = max(1, int(round(dens[v] * n)))
        key, xv = sa.gen_sparse_vector(n, k, 0x5EED1000 + v)
        keys.append(key)
        vals.append(xv)
        bls.append(int(deg[key].sum()))
    vptr = np.zeros(nvec + 1, np.int32)
    vptr[1:] = np.cumsum([k.size for k in keys])
    cap = np.minimum(np.array(bls, np.int64), n)
    yoff = np.zeros(nvec + 1, np.int64)
    yoff[1:] = np.cumsum(cap)
    d_vptr = sa.DeviceArray.from_host(vptr, np.int32)
    d_vkey = sa.DeviceArray.from_host(np.concatenate(keys), np.int32)
    d_vval = sa.DeviceArray.from_host(np.concatenate(vals), np.float32)
    d_yoff = sa.DeviceArray.from_host(yoff, np.int64)
    d_yk = sa.DeviceArray(int(yoff[-1]), np.int32)
    d_yv = sa.DeviceArray(int(yoff[-1]), np.float32)
    d_yn = sa.DeviceArray(nvec, np.int32)
    print("batched call", flush=True)
    used = M.A.run_batched(d_vptr, d_vkey, d_vval, d_yoff, d_yk, d_yv, d_yn)
    yn = d_yn.to_host()
    s_yk = sa.DeviceArray(n, np.int32)
    s_yv = sa.DeviceArray(n, np.float32)
    print("sequential calls", flush=True)
    order = np.argsort(dens)
    pick = set(int(order[j]) for j in np.linspace(0, nvec - 1, 8).astype(int))
    for v in range(nvec):
        M.A.set_sparse_x_ptr(d_vkey.ptr.value + 4 * int(vptr[v]), d_vval.ptr.value + 4 * int(vptr[v]),
                             int(vptr[v + 1] - vptr[v]))
        c, u = M.A.run_auto(y_key=s_yk, y_val=s_yv)
        assert (c, u) == (int(yn[v]), int(used[v])), f"vector {v}: sequential ({c}, {u}) batched ({yn[v]}, {used[v]})"
        bk = d_yk.to_host_range(int(yoff[v]), c)
        sk = s_yk.to_host(c)
        np.testing.assert_array_equal(bk, sk, err_msg=f"vector {v}")
        if v in pick:
            print(f"vector {v}: density {dens[v]:.2e} slot {u} ynnz {c}: oracle check", flush=True)
            check(n, rp, col, val, keys[v], vals[v], bk, d_yv.to_host_range(int(yoff[v]), c), f"s24 vector {v}")
            check(n, rp, col, val, keys[v], vals[v], sk, s_yv.to_host(c), f"s24 sequential vector {v}")
