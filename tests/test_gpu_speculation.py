"""Speculative dispatch of run_auto: the slot guessed from the previous
choice for the same x-size class is queued before the selector's choice
reaches the host, every launch gated on the device-side choice. Results
must not depend on whether the guess was right: each call is checked
against the FP64 oracle (index sets exact, values within
1e-5 * sum|a_ij x_j|) and the counts of right and wrong guesses against the
sequence of slots forced through the fixed-slot selector mode."""
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


@pytest.fixture(scope="module")
def rmat12_csr5(gpu):
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=41)
    n = 1 << 12
    M = sa.Matrix(n, n, rp, col, val)
    M.A.set_spmv_type(1)
    return M, oracle.csr_to_csc(n, n, rp, col, val)


def test_forced_slot_sequence(rmat12_csr5):
    """Fixed-slot mode forces the choice; the guess is the previous choice
    of the class, so the sequence gives a known mix of right and wrong
    guesses, each of which must leave the result intact (sparse output and
    the dense output of the same call)."""
    M, (cp, ri, cv) = rmat12_csr5
    A = M.A
    n = M.n
    key, val = sa.gen_sparse_vector(n, 300, 4242)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    # prepare the row formats once, outside the sequence (speculation only
    # guesses slots whose launches need no set-up)
    A.set_sparse_x(dk, dv, key.size)
    A.run(7, y_key=yk, y_val=yv)
    A.set_speculation(True)
    # this x is small enough for the single-launch column path, which would
    # finish the column calls without a guess; the guesses are the subject here
    A.set_small_x_path(False)
    h0, m0 = A.speculation_stats()
    seq = [2, 2, 7, 7, 0, 5, 5, 2, 3, 3, 1, 1, 2]
    try:
        for i, slot in enumerate(seq):
            A.selector_mode(2, slot)
            A.set_sparse_x(dk, dv, key.size)
            yd = sa.DeviceArray(n, np.float32) if i % 3 == 0 else None
            ynnz, used = A.run_auto(y_dense=yd, y_key=yk, y_val=yv)
            print(f"call {i}: forced {slot} used {used} ynnz {ynnz} stats {A.speculation_stats()}",
                  flush=True)
            assert used == slot
            check(M.m, cp, ri, cv, key, val, yk.to_host(ynnz), yv.to_host(ynnz), f"call {i} slot {slot}")
            if yd is not None:
                dense = yd.to_host()
                ek, ev, eb = oracle.spmspv_f64(M.m, cp, ri, cv, key, val)
                ref = np.zeros(n)
                ref[ek] = ev
                bound = np.zeros(n)
                bound[ek] = eb
                assert np.all(np.abs(dense - ref) <= TOL * bound + 1e-30), f"call {i}: dense output"
    finally:
        A.selector_mode(1)
        A.set_small_x_path(True)
    h1, m1 = A.speculation_stats()
    # (previous choice -> forced slot): none->2 no guess, 2->2 right, 2->7
    # wrong, 7->7 right, 7->0 wrong, 0->5 wrong, 5->5 right, 5->2 wrong,
    # 2->3 wrong, 3->3 right, 3->1 wrong, 1->1 right, 1->2 wrong
    assert (h1 - h0, m1 - m0) == (5, 7)


def test_speculation_on_off_same_results(rmat12_csr5):
    """MLP selector over an alternating density sequence, with speculation
    on and off: the same slots and the same outputs."""
    import os
    w = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd",
                             "selector_weights.npy")).astype(np.float32)
    M, (cp, ri, cv) = rmat12_csr5
    A = M.A
    n = M.n
    A.selector_set_weights(w)
    A.selector_mode(0)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    xs = []
    for i, d in enumerate([1e-3, 0.5, 1e-3, 0.05, 0.5, 1.0, 1e-3, 1.0, 0.05]):
        k = max(1, int(round(d * n)))
        key, val = sa.gen_sparse_vector(n, k, 900 + i)
        xs.append((key, val, sa.DeviceArray.from_host(key, np.int32), sa.DeviceArray.from_host(val, np.float32)))
    try:
        runs = {}
        for on in (False, True, True):
            A.set_speculation(on)
            got = []
            for i, (key, val, dk, dv) in enumerate(xs):
                A.set_sparse_x(dk, dv, key.size)
                ynnz, used = A.run_auto(y_key=yk, y_val=yv)
                k_h, v_h = yk.to_host(ynnz), yv.to_host(ynnz)
                check(M.m, cp, ri, cv, key, val, k_h, v_h, f"speculation {on} x {i} slot {used}")
                got.append(used)
            runs.setdefault(on, []).append(got)
            print(f"speculation {on}: slots {got} stats {A.speculation_stats()}", flush=True)
        assert runs[False][0] == runs[True][0] == runs[True][1]
    finally:
        A.selector_mode(1)
        A.set_speculation(True)


@pytest.mark.parametrize("slot", [0, 2, 3, 7])
def test_edge_sizes_under_speculation(rmat12_csr5, slot):
    """Empty x, a single entry and full x, each called twice in a row with
    the slot forced (the second call takes the speculative path: the
    features-only selector for slots 0 and 7, the column scan for 2 and 3),
    against the FP64 oracle."""
    M, (cp, ri, cv) = rmat12_csr5
    A = M.A
    n = M.n
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    A.set_speculation(True)
    try:
        A.selector_mode(2, slot)
        for k in (0, 1, n):
            key, val = sa.gen_sparse_vector(n, k, 5000 + k) if k else (np.zeros(0, np.int32),
                                                                         np.zeros(0, np.float32))
            dk = sa.DeviceArray.from_host(key if k else np.zeros(1, np.int32), np.int32)
            dv = sa.DeviceArray.from_host(val if k else np.zeros(1, np.float32), np.float32)
            for rep in range(2):
                A.set_sparse_x(dk, dv, k)
                ynnz, used = A.run_auto(y_key=yk, y_val=yv)
                assert used == slot
                check(M.m, cp, ri, cv, key, val, yk.to_host(ynnz), yv.to_host(ynnz),
                      f"slot {slot} xnnz {k} call {rep}")
    finally:
        A.selector_mode(1)


def test_enc5_buffers_across_inputs(rmat12_csr5):
    """The CSR5 slots read x through one of two encoding buffers; each tile
    kernel clears the other one, into which the next x is scattered without
    a fill. A sequence of fixed-slot and selector-dispatched calls (right
    and wrong guesses), sparse and dense x, a repeated x and changed ones,
    each checked against the FP64 oracle."""
    M, (cp, ri, cv) = rmat12_csr5
    A = M.A
    n = M.n
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    dx = sa.DeviceArray(n, np.float32)
    rng = np.random.default_rng(77)
    # (how, slot, density): run = fixed slot; auto = selector forced to the
    # slot (the guess is the class's previous choice); dense = dense x;
    # again = the same x as the step before, not set again
    plan = [("run", 7, 0.3), ("run", 5, 0.02), ("again", 7, None), ("auto", 7, 0.5),
            ("auto", 7, 0.4), ("auto", 2, 0.45), ("auto", 5, 0.35), ("dense", 7, 0.2),
            ("dense", 5, 0.6), ("run", 7, 1.0), ("auto", 5, 0.3), ("again", 5, None),
            ("auto", 7, 0.33), ("run", 1, 0.01), ("run", 5, 0.05)]
    A.set_speculation(True)
    key = val = None
    try:
        for i, (how, slot, d) in enumerate(plan):
            if how != "again":
                k = max(1, int(round(d * n)))
                key, val = sa.gen_sparse_vector(n, k, 7000 + i)
            if how == "dense":
                A.set_vector_type(1)
                xd = np.zeros(n, np.float32)
                xd[key] = val
                dx.upload(xd)
                A.set_x(dx)
                ynnz = A.run(slot, y_key=yk, y_val=yv)
                A.set_x(None)
                A.set_vector_type(0)
                ek = np.nonzero(xd)[0].astype(np.int32)
                kk, vv = ek, xd[ek]
            else:
                if how != "again":
                    dk = sa.DeviceArray.from_host(key, np.int32)
                    dv = sa.DeviceArray.from_host(val, np.float32)
                    A.set_sparse_x(dk, dv, key.size)
                if how == "auto":
                    A.selector_mode(2, slot)
                    ynnz, used = A.run_auto(y_key=yk, y_val=yv)
                    assert used == slot
                else:
                    ynnz = A.run(slot, y_key=yk, y_val=yv)
                kk, vv = key, val
            print(f"step {i}: {how} slot {slot} xnnz {kk.size} ynnz {ynnz} stats {A.speculation_stats()}",
                  flush=True)
            check(M.m, cp, ri, cv, kk, vv, yk.to_host(ynnz), yv.to_host(ynnz), f"step {i} {how} slot {slot}")
    finally:
        A.selector_mode(1)
        A.set_vector_type(0)


@pytest.mark.parametrize("slot", range(9))
def test_result_count_paths(rmat12_csr5, slot):
    """The result count reaches the host from the compaction that ends the
    call (slots 0, 2 and 4-7) or from a one-thread launch after the call
    (the sorting slots 1, 3 and the bucket slot 8): with the sorted output,
    with the count alone, with a dense output, through run and, for the
    selector's eight classes (the bucket slot is not one), through run_auto
    (its second call of the size class takes the speculative path), each
    against the FP64 oracle."""
    M, (cp, ri, cv) = rmat12_csr5
    A = M.A
    n = M.n
    key, val = sa.gen_sparse_vector(n, 400, 6100 + slot)
    ek, ev, eb = oracle.spmspv_f64(M.m, cp, ri, cv, key, val)
    ref = np.zeros(M.m)
    ref[ek] = ev
    bound = np.zeros(M.m)
    bound[ek] = eb
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(val, np.float32)
    yk = sa.DeviceArray(M.m, np.int32)
    yv = sa.DeviceArray(M.m, np.float32)
    yd = sa.DeviceArray(M.m, np.float32)

    def dense_ok(tag):
        d = yd.to_host()
        assert np.all(np.abs(d - ref) <= TOL * bound + 1e-30), f"slot {slot} {tag}: dense output"

    A.set_sparse_x(dk, dv, key.size)
    print(f"slot {slot}: run", flush=True)
    ynnz = A.run(slot, y_key=yk, y_val=yv)
    check(M.m, cp, ri, cv, key, val, yk.to_host(ynnz), yv.to_host(ynnz), f"slot {slot} run")
    assert A.run(slot) == ek.size, f"slot {slot}: count alone"
    assert A.run(slot, y_dense=yd) == ek.size, f"slot {slot}: count with dense output"
    dense_ok("run")
    if slot == sa.K_COL_BUCKET:
        return
    A.set_speculation(True)
    try:
        A.selector_mode(2, slot)
        for rep in range(2):
            print(f"slot {slot}: run_auto call {rep} stats {A.speculation_stats()}", flush=True)
            ynnz, used = A.run_auto(y_key=yk, y_val=yv)
            assert used == slot
            check(M.m, cp, ri, cv, key, val, yk.to_host(ynnz), yv.to_host(ynnz), f"slot {slot} auto {rep}")
            ynnz, used = A.run_auto(y_dense=yd)
            assert (ynnz, used) == (ek.size, slot), f"slot {slot} auto {rep}: count with dense output"
            dense_ok(f"auto {rep}")
    finally:
        A.selector_mode(1)
