"""Parity at the density sweep's full size: R-MAT scale 20, edge factor 16
(a, b, c = 0.57, 0.19, 0.19; 1,048,576 rows, about 31.4 M nonzeros), the
six densities of the sweep with the sweep's own seeds. Every slot in the
merge-path row format, the two CSR5 slots at sigma 24, and the
selector-dispatched call (MLP weights, speculation on, each density called
twice so that the second call takes the speculative path), each against the
FP64 oracle: index sets exact, values within 1e-5 * sum|a_ij x_j|."""
import os

import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
TOL = 1e-5
DENSITIES = [1e-4, 1e-3, 1e-2, 0.1, 0.5, 1.0]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd",
                       "selector_weights.npy")


@pytest.fixture(scope="module")
def rmat20(gpu):
    n = 1 << 20
    rp, col, val = sa.gen_rmat_csr(20, 16, 0.57, 0.19, 0.19, seed=0x5EED0002)
    M = sa.Matrix(n, n, rp, col, val)
    M.A.selector_set_weights(np.load(WEIGHTS).astype(np.float32))
    yield M, oracle.csr_to_csc(n, n, rp, col, val)
    M.A.set_spmv_type(0)
    M.A.set_sigma(-1)


def check(m, expect, yk, yv, tag):
    ek, ev, eb = expect
    assert yk.size == ek.size, f"{tag}: ynnz {yk.size} != {ek.size}"
    np.testing.assert_array_equal(yk, ek, err_msg=tag)
    err = np.abs(yv.astype(np.float64) - ev)
    bad = np.nonzero(err > TOL * eb + 1e-30)[0]
    assert bad.size == 0, f"{tag}: {bad.size} values out of bound, first row {ek[bad[0]]}"


@pytest.mark.parametrize("di", range(len(DENSITIES)))
def test_rmat20_sweep_parity(rmat20, di):
    M, (cp, ri, cv) = rmat20
    A = M.A
    n = M.n
    k = max(1, int(round(DENSITIES[di] * n)))
    key, val = sa.gen_sparse_vector(n, k, 0x5EED0100 + di)
    expect = oracle.spmspv_f64(M.m, cp, ri, cv, key, val)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    runs = [(0, s) for s in range(9)] + [(1, sa.K_ROW_LB), (1, sa.K_SPMV_LB)]
    for fmt, slot in runs:
        A.set_spmv_type(fmt)
        A.set_sigma(24 if fmt else -1)
        A.set_sparse_x(dk, dv, k)
        print(f"density {DENSITIES[di]} xnnz {k} format {fmt} slot {slot}", flush=True)
        ynnz = A.run(slot, y_key=yk, y_val=yv)
        check(M.m, expect, yk.to_host(ynnz), yv.to_host(ynnz), f"d={DENSITIES[di]} format {fmt} slot {slot}")
    A.set_speculation(True)
    A.selector_mode(0)
    for rep in range(2):
        A.set_sparse_x(dk, dv, k)
        ynnz, used = A.run_auto(y_key=yk, y_val=yv)
        print(f"density {DENSITIES[di]} run_auto call {rep}: slot {used} stats {A.speculation_stats()}",
              flush=True)
        check(M.m, expect, yk.to_host(ynnz), yv.to_host(ynnz), f"d={DENSITIES[di]} auto {rep} slot {used}")
