#!/usr/bin/env python3
"""Benchmark of the adaptive SpMSpV engine on one MI355X.

Workload (the density sweep the metric is quoted on): R-MAT scale 20
(1,048,576 vertices), edge factor 16, a/b/c/d = .57/.19/.19/.05,
symmetrised, duplicates and self-loops removed, FP32 values U(0,1]; x at
densities 0.01%, 0.1%, 1%, 10%, 50%, 100% with random support and U[-1,1]
values. One step = one selector-dispatched SpMSpV (features + MLP or rule
fallback + the chosen kernel, sorted sparse output) per density.

Each timed call starts with a cold cache: a 512 MB write evicts the MALL
before it, outside the timed interval. Calls are timed with events on the
handle's stream. `value` = algorithmic bytes of the step / step time, with
the kernel-independent byte count B = 16*xnnz + 8*bin_len + 8*y_nnz.

Also reported, each outside the timed steps: every fixed slot on the same
inputs (best fixed kernel, oracle-best per density, selector regret), the
FP32 error of every density against the FP64 oracle, speculation hits and a
mixed-order sequence, the roofline of the dominant kernel against the
measured read / copy / gather ceilings, the R-MAT scale 22 target densities,
the level-synchronous BFS (R-MAT scale 22, 64 sources) with a CPU BFS, the
batched multi-vector configuration (R-MAT scale 24, 256 vectors), the
PageRank driver, the config-1 uniform matrix, and the reference's CPU path
(the oracle's FP32 restatement of SpmSpvCSCHost) on one core beside the
build's all-core port.
"""
import argparse
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spmspv-adaptive_amd"))

HBM_PEAK_GBS = 8000.0
DENSITIES = [1e-4, 1e-3, 1e-2, 0.1, 0.5, 1.0]
TARGET_DENSITIES = [0.01, 0.1, 1.0]
WEIGHTS = os.path.join(ROOT, "spmspv-adaptive_amd", "selector_weights.npy")
BFS_WEIGHTS = os.path.join(ROOT, "spmspv-adaptive_amd", "bfs_selector_weights.npy")
AUTO_SLOT = 99  # --only-slot value that times the selector-dispatched call


def alg_bytes(xnnz, bin_len, ynnz):
    return 16.0 * xnnz + 8.0 * bin_len + 8.0 * ynnz


def profile_slot(sa, A, n, args, yk, yv):
    """Profiling mode: every slot of --only-slot at every density of
    --density, cold, `reps` times each, slots outer, densities inner (rocprofv3
    runs read the kernel durations and counters of exactly these launches;
    each call follows one 512 MB fill, which delimits the calls in a trace).
    With --profile-autotune the per-matrix tuning runs first, on the dense x,
    as in the sweep; the calibration copy (--calib-mb) follows it, so every
    dispatch after the copy belongs to the profiled calls."""
    slots = [int(v) for v in args.only_slot.split(",")]
    dens = [float(v) for v in args.density.split(",")]
    xs = {}
    for d in dens:
        k = max(1, int(round(d * n)))
        key, xv = sa.gen_sparse_vector(n, k, 0x5EED0100 + DENSITIES.index(d) if d in DENSITIES else 0x5EED0200)
        xs[d] = (k, sa.DeviceArray.from_host(key, np.int32), sa.DeviceArray.from_host(xv, np.float32))
    tuned = None
    if args.profile_autotune:
        key = np.arange(n, dtype=np.int32)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(np.ones(n, np.float32), np.float32)
        A.set_sparse_x(dk, dv, n)
        tuned = A.autotune(reps=3)
    if args.spmv_type >= 0:  # a K5/K7 format pinned for the profile
        A.set_spmv_type(args.spmv_type)
        A.set_sigma(args.sigma)
    if args.buckets:  # the bucket slot's bin count pinned (one call sets it)
        import ctypes
        kx, dkx, dvx = xs[dens[0]]
        A.set_sparse_x(dkx, dvx, kx)
        c = ctypes.c_int(-1)
        sa.lib().spmspv_csc_bucket(A.h, 1.0, ctypes.byref(c), yk.ptr, yv.ptr, args.buckets)
    if args.bucket_range:  # the bucket rule pinned, so that every profiled run takes the same slots
        lo, hi = (int(v) for v in args.bucket_range.split(","))
        A.set_bucket_range(lo, hi)
    if tuned is not None:
        tuned = A.tuned()
    if args.no_small_x:
        A.set_small_x_path(False)
    bls = {}
    for d, (k, dk, dv) in xs.items():
        A.set_sparse_x(dk, dv, k)
        bls[d] = A.computeBinlen()
    sa.synchronize()
    if args.calib_mb:
        nb = args.calib_mb << 20
        src = sa.DeviceArray(nb // 4, np.float32)
        dst = sa.DeviceArray(nb // 4, np.float32)
        src.fill_zero()
        sa.synchronize()
        sa.lib().spmspv_memcpy_d2d(dst.ptr, src.ptr, nb)
        sa.synchronize()
    for slot in slots:
        for d in dens:
            k, dk, dv = xs[d]
            ms = []
            ynnz = used = 0
            for _ in range(args.reps):
                A.set_sparse_x(dk, dv, k)
                A.flush_cache(args.flush_mb << 20)
                A.timer_start()
                if slot == AUTO_SLOT:
                    ynnz, used = A.run_auto(y_key=yk, y_val=yv)
                else:
                    ynnz, used = A.run(slot, y_key=yk, y_val=yv), slot
                ms.append(A.timer_stop())
            name = "auto" if slot == AUTO_SLOT else sa.KERNEL_NAMES[slot]
            print(json.dumps({"slot": name, "slot_id": slot, "ran": sa.KERNEL_NAMES[used], "density": d,
                              "xnnz": k, "bin_len": int(bls[d]), "ynnz": int(ynnz), "ms": ms,
                              "alg_bytes": alg_bytes(k, bls[d], ynnz), "calib_bytes": args.calib_mb << 20,
                              "autotuned": tuned}), flush=True)


def cpu_info():
    """(model name, cores this process may use)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def result_error(oracle, m, cp, ri, cv, key, val, yk, yv):
    """Index sets equal to the FP64 oracle's, and max |y - y64| / sum|a x|."""
    ek, ev, eb = oracle.spmspv_f64(m, cp, ri, cv, key, val)
    same = bool(yk.size == ek.size and np.array_equal(yk, ek))
    if not same:
        return False, None
    if ek.size == 0:
        return True, 0.0
    return True, float((np.abs(yv.astype(np.float64) - ev) / np.maximum(eb, 1e-300)).max())


def run_bfs(sa, args, A, rp, col, n):
    """BFS: fixed slots vs per-level selection, on the SpMSpV graph or on a
    separate R-MAT of --bfs-scale (the level-synchronous BFS workload:
    scale 22, 64 sources). Levels of the first --bfs-check sources are
    compared with the CPU BFS for every policy."""
    bscale = args.bfs_scale or args.scale
    if bscale == args.scale:
        B, brp, bcol, bn = A, rp, col, n
    else:
        g = bfs_graph(sa, args)
        B, brp, bcol, bn = g["M"].A, g["rp"], g["col"], g["n"]
    if os.path.exists(BFS_WEIGHTS):  # the BFS levels' own selector
        B.bfs_selector_set_weights(np.load(BFS_WEIGHTS).astype(np.float32))
    deg = np.diff(brp)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(0x5EED0003)
    srcs = rng.choice(cand, size=min(args.bfs_sources, cand.size), replace=False)
    lev = sa.DeviceArray(bn, np.int32)
    refs = {}
    cpu_s, cpu_edges = 0.0, 0.0
    if args.bfs_check > 0:
        oracle = load_oracle()
        for s_ in srcs[:args.bfs_check]:
            t0 = time.perf_counter()
            lv = oracle.bfs(bn, brp, bcol, int(s_))[0]  # symmetric: CSR = CSC
            cpu_s += time.perf_counter() - t0
            refs[int(s_)] = lv
            cpu_edges += float(deg[lv >= 0].sum()) / 2.0
    res, levels, checked = {}, {}, 0
    roof_bytes, roof_s = 0.0, 0.0
    for pol in [-1] + list(range(sa.NUM_SLOTS)):
        B.bfs(int(srcs[0]), lev, policy=pol, use_graphs=1)  # warm-up + graph capture
        secs, edges = 0.0, 0.0
        for s_ in srcs:
            st = B.bfs(int(s_), lev, policy=pol, use_graphs=1)
            secs += st[2]
            edges += st[1]
            if pol < 0:
                # boolean level bytes summed over the levels: 12 per frontier
                # vertex + 4 per frontier edge + 4 per new vertex + n/8 per level
                roof_bytes += 16.0 * st[3] + 8.0 * st[1] + st[0] * ((bn + 7) // 8)
                roof_s += st[2]
            if int(s_) in refs:
                if not np.array_equal(lev.to_host(), refs[int(s_)]):
                    raise SystemExit(f"BFS levels differ from the CPU BFS: policy {pol}, source {s_}")
                checked += 1
            levels[int(s_)] = int(st[0])
        res["learned" if pol < 0 else sa.KERNEL_NAMES[pol]] = edges / secs / 1e9
    # the same sources as one multi-source BFS (64 visited bits per vertex)
    msb = None
    if len(srcs) <= 64:
        mlev = sa.DeviceArray(bn * len(srcs), np.int32)
        B.bfs_multi(srcs, mlev)  # warm-up
        msecs, medges, mreached = [], 0.0, 0
        for _ in range(3):
            st = B.bfs_multi(srcs, mlev)
            msecs.append(st[2])
            medges, mreached = st[1], int(st[3])
        mh = mlev.to_host().reshape(bn, len(srcs))
        mchecked = 0
        for i, s_ in enumerate(srcs):
            if int(s_) in refs:
                if not np.array_equal(mh[:, i], refs[int(s_)]):
                    raise SystemExit(f"multi-source BFS levels differ from the CPU BFS: source {s_}")
                mchecked += 1
        del mh, mlev
        ms_s = float(np.median(msecs))
        msb = {"gteps": round(medges / ms_s / 1e9, 3), "ms": round(1e3 * ms_s, 3), "sources": int(len(srcs)),
               "reached_pairs": mreached, "levels_checked_vs_cpu": mchecked,
               "edges": "undirected edges summed over the sources' components, as for the single-source runs"}
    fixed_best = max(v for kk, v in res.items() if kk != "learned")
    bfs = {"graph": f"rmat{bscale}", "sources": int(len(srcs)), "gteps": res,
           "learned_gteps": res["learned"], "best_fixed_gteps": fixed_best,
           "levels_checked_vs_cpu": checked,
           "mean_levels": round(float(np.mean(list(levels.values()))), 2),
           "roofline_learned": {"achieved_gbs": round(roof_bytes / roof_s / 1e9, 2),
                                "frac": round(roof_bytes / roof_s / 1e9 / HBM_PEAK_GBS, 5),
                                "bytes": "16*reached + 8*edges + levels*n/8 per BFS",
                                "timed": "host wall clock of each BFS, level round trips included"}}
    if msb is not None:
        bfs["multi_source"] = msb
    if cpu_s > 0:
        bfs["cpu_baseline"] = {"gteps": round(cpu_edges / cpu_s / 1e9, 4), "cores": 1, "kind": "port",
                               "sample": f"{len(refs)} sources of the same graph, textbook queue BFS "
                                         f"(oracle), {cpu_s:.2f} s"}
    return bfs


_BFS_GRAPH = {}


def bfs_graph(sa, args):
    """The R-MAT graph of the BFS workload (scale --bfs-scale, edge factor 16,
    seed 0x5EED0004), built once: host CSR (= CSC by symmetry) and a handle
    with the selector weights."""
    if "M" not in _BFS_GRAPH:
        bscale = args.bfs_scale
        bn = 1 << bscale
        brp, bcol, bval = sa.gen_rmat_csr(bscale, 16, 0.57, 0.19, 0.19, seed=0x5EED0004)
        BM = sa.Matrix(bn, bn, brp, bcol, bval)
        if os.path.exists(WEIGHTS):
            BM.A.selector_set_weights(np.load(WEIGHTS).astype(np.float32))
        else:
            BM.A.selector_mode(1)
        _BFS_GRAPH.update(M=BM, rp=brp, col=bcol, val=bval, n=bn, scale=bscale)
    return _BFS_GRAPH


def run_c1(sa, args, oracle):
    """The reference's CPU-runnable case: uniform n = 65,536 with 16 distinct
    rows per column, x with 655 keys (seed 0x5EED0001). The sort-based
    kernel (col-naive-sort) and the selector-dispatched call, cold, median
    of --reps, against SpmSpvCSCHost on one core; FP32 error vs FP64."""
    n, k = 65536, 16
    cp, ri, cv = sa.gen_uniform_csc(n, k, 0x5EED0001)
    rp, col, val = oracle.csr_to_csc(n, n, cp, ri, cv)  # CSR of A = CSC of its transpose
    M = sa.Matrix(n, n, rp, col, val)
    if os.path.exists(WEIGHTS):
        M.A.selector_set_weights(np.load(WEIGHTS).astype(np.float32))
    key, xv = sa.gen_sparse_vector(n, 655, 0x5EED0001)
    dk = sa.DeviceArray.from_host(key, np.int32)
    dv = sa.DeviceArray.from_host(xv, np.float32)
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    res = {"n": n, "nnz": int(cp[-1]), "xnnz": 655}
    for name, slot in (("col-naive-sort", sa.K_COL_NAIVE_SORT), ("auto", None)):
        ms = []
        for r in range(2 + args.reps):
            M.A.set_sparse_x(dk, dv, key.size)
            M.A.flush_cache(args.flush_mb << 20)
            M.A.timer_start()
            if slot is None:
                c, used = M.A.run_auto(y_key=yk, y_val=yv)
            else:
                c, used = M.A.run(slot, y_key=yk, y_val=yv), slot
            t = M.A.timer_stop()
            if r >= 2:
                ms.append(t)
        same, err = result_error(oracle, n, cp, ri, cv, key, xv, yk.to_host(c), yv.to_host(c))
        res[name] = {"ms": round(float(np.median(ms)), 5), "slot": sa.KERNEL_NAMES[used], "ynnz": int(c),
                     "index_sets_exact": same, "max_rel_err": err}
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        oracle.spmspv_csc_host_f32(n, cp, ri, cv, key, xv)
        ts.append(time.perf_counter() - t0)
    res["cpu_ms"] = round(1e3 * float(np.median(ts)), 4)
    return res


def run_target(sa, args, oracle):
    """SpMSpV on the BFS graph (R-MAT scale 22) at densities 1%, 10% and 100%:
    the selector-dispatched call, cold, median of --reps; the tile kernel's
    own time where a row slot ran; FP32 error against the FP64 oracle."""
    g = bfs_graph(sa, args)
    A, n, rp, col, val = g["M"].A, g["n"], g["rp"], g["col"], g["val"]
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    xs = []
    for i, d in enumerate(TARGET_DENSITIES):
        k = max(1, int(round(d * n)))
        key, xv = sa.gen_sparse_vector(n, k, 0x5EED0300 + i)
        xs.append((k, sa.DeviceArray.from_host(key, np.int32), sa.DeviceArray.from_host(xv, np.float32), key, xv))
    k, dk, dv = xs[-1][:3]
    A.set_sparse_x(dk, dv, k)
    tuned = A.autotune(reps=3)
    out = []
    A.time_main_kernel(True)
    for i, d in enumerate(TARGET_DENSITIES):
        k, dk, dv, key, xv = xs[i]
        A.set_sparse_x(dk, dv, k)
        bl = A.computeBinlen()
        ms, kms, ynnz, used = [], [], 0, -1
        for r in range(2 + args.reps):
            A.set_sparse_x(dk, dv, k)
            A.flush_cache(args.flush_mb << 20)
            A.timer_start()
            ynnz, used = A.run_auto(y_key=yk, y_val=yv)
            t = A.timer_stop()
            if r >= 2:
                ms.append(t)
                if used in (sa.K_ROW_LB, sa.K_SPMV_LB):
                    kms.append(A.main_kernel_ms())
        same, err = result_error(oracle, n, rp, col, val, key, xv, yk.to_host(ynnz), yv.to_host(ynnz))
        b = alg_bytes(k, bl, ynnz)
        call = float(np.median(ms))
        row = {"density": d, "xnnz": k, "bin_len": int(bl), "ynnz": int(ynnz), "slot": sa.KERNEL_NAMES[used],
               "call_ms": round(call, 5), "call_gbs": round(b / (call * 1e-3) / 1e9, 2),
               "call_frac": round(b / (call * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
               "index_sets_exact": same, "max_rel_err": err}
        if kms:
            km = float(np.median(kms))
            row.update(kernel_ms=round(km, 5), kernel_gbs=round(b / (km * 1e-3) / 1e9, 2),
                       kernel_frac=round(b / (km * 1e-3) / 1e9 / HBM_PEAK_GBS, 5))
        out.append(row)
    A.time_main_kernel(False)
    ceil = A.measure_ceilings(3)  # the gather rate over this graph's n columns
    return {"graph": f"rmat{g['scale']}", "nnz": int(g["M"].nnz), "autotuned": tuned, "per_density": out,
            "ceilings": {k: round(v, 2) for k, v in ceil.items()},
            "target": ">= 60% of HBM peak at >= 1% density"}


def run_batched_bench(sa, args):
    """Batched multi-vector SpMSpV: --batched-vectors vectors (density
    log-uniform in [1e-4, 0.2], random support, U[-1,1]) against one R-MAT
    of --batched-scale, as one spmspv_run_batched call vs the same vectors
    as sequential selector-dispatched calls (both warm, wall clock around
    the whole set, outputs resident on the device). The batched results
    are compared with the sequential ones (index sets exact, values within
    1e-5 relative)."""
    scale, nvec = args.batched_scale, args.batched_vectors
    n = 1 << scale
    t0 = time.time()
    rp, col, val = sa.gen_rmat_csr(scale, 16, 0.57, 0.19, 0.19, seed=0x5EED0005)
    M = sa.Matrix(n, n, rp, col, val)
    A = M.A
    if os.path.exists(WEIGHTS):
        A.selector_set_weights(np.load(WEIGHTS).astype(np.float32))
    else:
        A.selector_mode(1)
    # per-matrix tuning on a dense x, as for the sweep's matrix (tile widths,
    # CSR5 width, the bucket slot's bins and its rule's range)
    tuned = None
    if not args.no_autotune:
        dk = sa.DeviceArray.from_host(np.arange(n, dtype=np.int32), np.int32)
        dv = sa.DeviceArray.from_host(np.ones(n, np.float32), np.float32)
        A.set_sparse_x(dk, dv, n)
        tuned = A.autotune(reps=2)
        del dk, dv
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
    s_yk = sa.DeviceArray(int(yoff[-1]), np.int32)
    s_yv = sa.DeviceArray(int(yoff[-1]), np.float32)
    setup_s = time.time() - t0
    # warm-up of both paths
    A.run_batched(d_vptr, d_vkey, d_vval, d_yoff, d_yk, d_yv, d_yn)
    bt, st = [], []
    seq_nnz = np.zeros(nvec, np.int64)
    seq_used = np.zeros(nvec, np.int32)
    for rep in range(args.batched_reps):
        sa.synchronize()
        t = time.perf_counter()
        used = A.run_batched(d_vptr, d_vkey, d_vval, d_yoff, d_yk, d_yv, d_yn)
        sa.synchronize()
        bt.append(time.perf_counter() - t)
        t = time.perf_counter()
        for v in range(nvec):
            o = int(yoff[v])
            A.set_sparse_x_ptr(d_vkey.ptr.value + 4 * int(vptr[v]), d_vval.ptr.value + 4 * int(vptr[v]),
                               int(vptr[v + 1] - vptr[v]))
            seq_nnz[v], seq_used[v] = A.run_auto_ptr(s_yk.ptr.value + 4 * o, s_yv.ptr.value + 4 * o)
        sa.synchronize()
        st.append(time.perf_counter() - t)
    yn = d_yn.to_host()
    # compare with the sequential calls: counts, slots and index sets exact;
    # values agree up to summation order (the atomic slots add in any order)
    bk, bv, sk, sv = d_yk.to_host(), d_yv.to_host(), s_yk.to_host(), s_yv.to_host()
    diff_nnz = int(np.sum(yn != seq_nnz))
    diff_slot = int(np.sum(used != seq_used))
    diff_idx, max_dv = 0, 0.0
    for v in range(nvec):
        o, c = int(yoff[v]), int(min(yn[v], seq_nnz[v]))
        if not np.array_equal(bk[o:o + c], sk[o:o + c]):
            diff_idx += 1
        elif c:
            max_dv = max(max_dv, float(np.max(np.abs(bv[o:o + c].astype(np.float64) - sv[o:o + c]))))
    same = diff_nnz == 0 and diff_slot == 0 and diff_idx == 0
    b_ms, s_ms = 1e3 * float(np.median(bt)), 1e3 * float(np.median(st))
    prods = int(np.sum(bls))
    # 8 vectors spanning the density range against the FP64 oracle (the
    # symmetric CSR is the CSC), and the reference's CPU path on a sample
    oracle = load_oracle()
    order = np.argsort(dens)
    pick = [int(order[j]) for j in np.linspace(0, nvec - 1, min(8, nvec)).astype(int)]
    oracle_ok, oracle_err = True, 0.0
    for v in pick:
        o, c = int(yoff[v]), int(yn[v])
        ok, err = result_error(oracle, n, rp, col, val, keys[v], vals[v], bk[o:o + c], bv[o:o + c])
        oracle_ok &= ok
        oracle_err = max(oracle_err, err if err is not None else float("inf"))
    cpu_t, cpu_b, cpu_v = 0.0, 0.0, 0
    for v in order:  # sparsest first: a bounded sample of about 5 s
        t = time.perf_counter()
        _, ck, _ = oracle.spmspv_csc_host_f32(n, rp, col, val, keys[v], vals[v])
        cpu_t += time.perf_counter() - t
        cpu_b += alg_bytes(keys[v].size, bls[v], ck.size)
        cpu_v += 1
        if cpu_t > 5.0:
            break
    return {"graph": f"rmat{scale}", "nnz": int(M.nnz), "vectors": nvec,
            "density": "log-uniform [1e-4, 0.2]", "total_bin_len": prods, "autotuned": tuned,
            "batched_ms": round(b_ms, 3), "sequential_ms": round(s_ms, 3),
            "speedup": round(s_ms / b_ms, 3),
            "batched_gbs": round(sum(alg_bytes(keys[v].size, bls[v], int(yn[v])) for v in range(nvec))
                                 / (b_ms * 1e-3) / 1e9, 2),
            "slots_used": {sa.KERNEL_NAMES[k]: int((used == k).sum()) for k in np.unique(used)},
            "matches_sequential": same, "vectors_differing": {"ynnz": diff_nnz, "slot": diff_slot,
                                                              "index_set": diff_idx},
            "max_abs_value_diff": max_dv, "setup_s": round(setup_s, 1),
            "oracle_checked_vectors": len(pick), "oracle_index_sets_exact": bool(oracle_ok),
            "oracle_max_rel_err": oracle_err,
            "cpu_baseline": {"gbs": round(cpu_b / cpu_t / 1e9, 3), "cores": 1, "kind": "port",
                             "sample": f"SpmSpvCSCHost on the {cpu_v} sparsest of the {nvec} vectors, "
                                       f"{cpu_t:.1f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--reps-fixed", type=int, default=3)
    ap.add_argument("--bfs-sources", type=int, default=64)
    ap.add_argument("--bfs-scale", type=int, default=22,
                    help="R-MAT scale of the BFS graph (0: the SpMSpV graph); 22 with 64 sources is the BFS workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (R-MAT scale 24) section")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the sweep, its fixed slots and the CPU baseline (no scale-22/24 sections)")
    ap.add_argument("--batched", action="store_true",
                    help="run only the batched measurement (one JSON line)")
    ap.add_argument("--batched-scale", type=int, default=24)
    ap.add_argument("--batched-vectors", type=int, default=256)
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--bfs-only", action="store_true",
                    help="run only the BFS measurement (prints one JSON line with the BFS result)")
    ap.add_argument("--bfs-check", type=int, default=2,
                    help="sources whose levels are compared with the CPU BFS for every policy")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fixed", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--spmv-type", type=int, default=-1,
                    help="profiling mode: K5/K7 format (0 merge-path, 1 CSR5), as reported under autotuned")
    ap.add_argument("--sigma", type=int, default=-1, help="profiling mode: CSR5 sigma")
    ap.add_argument("--buckets", type=int, default=0, help="profiling mode: the bucket slot's bin count")
    ap.add_argument("--bucket-range", default="",
                    help="profiling mode: the bucket rule's bin_len range lo,hi (as reported under autotuned)")
    ap.add_argument("--flush-mb", type=int, default=512)
    ap.add_argument("--only-slot", default="",
                    help="profiling mode: time these slots (comma list) at the --density values only "
                         "(for rocprofv3 runs); 99 = the selector-dispatched call")
    ap.add_argument("--density", default="1.0", help="profiling mode: comma list of x densities")
    ap.add_argument("--no-small-x", action="store_true",
                    help="profiling mode: turn the single-launch small-x column path off")
    ap.add_argument("--profile-autotune", action="store_true",
                    help="profiling mode: run the per-matrix tuning first, as the sweep does")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--calib-mb", type=int, default=0,
                    help="profiling mode: first copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                          "(multi-GPU sharding is a design note only)", "n_gpus": args.gpus}))
        sys.exit(2)

    import spmspv_amd as sa
    ROW_TILED = (sa.K_ROW_LB, sa.K_SPMV_LB)  # the slots whose tile kernel is timed alone

    if args.batched:
        print(json.dumps({"batched": run_batched_bench(sa, args)}), flush=True)
        return

    if args.bfs_only:
        if not args.bfs_scale:
            args.bfs_scale = args.scale
        args.scale = -1  # the sweep's graph is not built
        print(json.dumps({"bfs": run_bfs(sa, args, None, None, None, 0)}), flush=True)
        return

    t_setup = time.time()
    n = 1 << args.scale
    rp, col, val = sa.gen_rmat_csr(args.scale, 16, 0.57, 0.19, 0.19, seed=0x5EED0002)
    M = sa.Matrix(n, n, rp, col, val)
    A = M.A
    sel_kind = "rules"
    if os.path.exists(WEIGHTS):
        A.selector_set_weights(np.load(WEIGHTS).astype(np.float32))
        sel_kind = "mlp"
    else:
        A.selector_mode(1)
    flush = args.flush_mb << 20

    # inputs: one x per density, resident on the device
    xs = []
    for i, d in enumerate(DENSITIES):
        k = max(1, int(round(d * n)))
        key, xv = sa.gen_sparse_vector(n, k, 0x5EED0100 + i)
        xs.append((k, sa.DeviceArray.from_host(key, np.int32), sa.DeviceArray.from_host(xv, np.float32), key, xv))
    yk = sa.DeviceArray(n, np.int32)
    yv = sa.DeviceArray(n, np.float32)
    if args.only_slot:
        profile_slot(sa, A, n, args, yk, yv)
        return
    meta = []
    for (k, dk, dv, _, _) in xs:
        A.set_sparse_x(dk, dv, k)
        bl = A.computeBinlen()
        meta.append(bl)
    # per-matrix tuning (team sizes, row tile width, K5/K7 format) on the
    # densest x, as a user would run it once per matrix
    k, dk, dv = xs[-1][:3]
    A.set_sparse_x(dk, dv, k)
    tuned = A.autotune(reps=3) if not args.no_autotune else None
    setup_s = time.time() - t_setup
    ceilings = A.measure_ceilings(3)

    def timed_call(i, slot=None):
        k, dk, dv = xs[i][:3]
        A.set_sparse_x(dk, dv, k)
        if flush:
            A.flush_cache(flush)
        A.timer_start()
        if slot is None:
            ynnz, used = A.run_auto(y_key=yk, y_val=yv)
        else:
            ynnz, used = A.run(slot, y_key=yk, y_val=yv), slot
        ms = A.timer_stop()
        return ms, ynnz, used

    # ---- selector-dispatched sweep (the measured step)
    for _ in range(args.warmup):
        for i in range(len(DENSITIES)):
            timed_call(i)
    step_ms = []
    per_density = [[] for _ in DENSITIES]
    kern_ms = [[] for _ in DENSITIES]  # the row slots' tile kernel, HIP events on its stream
    chosen = [None] * len(DENSITIES)
    ynnzs = [0] * len(DENSITIES)
    A.time_main_kernel(True)
    sa.synchronize()
    spec0 = A.speculation_stats()
    for _ in range(args.steps):
        tot = 0.0
        for i in range(len(DENSITIES)):
            ms, ynnz, used = timed_call(i)
            per_density[i].append(ms)
            if used in ROW_TILED:
                kern_ms[i].append(A.main_kernel_ms())
            chosen[i] = used
            ynnzs[i] = ynnz
            tot += ms
        step_ms.append(tot)
    sa.synchronize()
    A.time_main_kernel(False)
    spec1 = A.speculation_stats()
    step_bytes = sum(alg_bytes(xs[i][0], meta[i], ynnzs[i]) for i in range(len(DENSITIES)))
    ms_step = float(np.median(step_ms))
    value = step_bytes / (ms_step * 1e-3) / 1e9

    # ---- warm back-to-back wall clock (no flush): what a loop of calls sees
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for i in range(len(DENSITIES)):
            k, dk, dv = xs[i][:3]
            A.set_sparse_x(dk, dv, k)
            A.run_auto(y_key=yk, y_val=yv)
    sa.synchronize()
    warm_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    warm_value = step_bytes / (warm_ms * 1e-3) / 1e9

    # ---- FP32 error of every density against the FP64 oracle (symmetric
    # matrix: the CSR is the CSC), outside the timed steps
    oracle = load_oracle()
    errors = []
    for i in range(len(DENSITIES)):
        k, dk, dv, key, xv = xs[i]
        A.set_sparse_x(dk, dv, k)
        ynnz, used = A.run_auto(y_key=yk, y_val=yv)
        same, err = result_error(oracle, n, rp, col, val, key, xv, yk.to_host(ynnz), yv.to_host(ynnz))
        errors.append({"density": DENSITIES[i], "index_sets_exact": same, "max_rel_err": err})

    # ---- mixed order: fresh x of log-uniform density in random order, so
    # consecutive calls of one size class may want different slots; timed
    # with speculation on and off
    rng = np.random.default_rng(0x5EED0400)
    mixed = []
    for j in range(24):
        kk = max(1, int(round(float(np.exp(rng.uniform(np.log(1e-4), 0.0))) * n)))
        key, xv = sa.gen_sparse_vector(n, kk, 0x5EED0500 + j)
        mixed.append((kk, sa.DeviceArray.from_host(key, np.int32), sa.DeviceArray.from_host(xv, np.float32)))
    mixed_res = {}
    for spec_on in (True, False):
        A.set_speculation(spec_on)
        h0 = A.speculation_stats()
        tot = 0.0
        for kk, dk, dv in mixed:
            A.set_sparse_x(dk, dv, kk)
            if flush:
                A.flush_cache(flush)
            A.timer_start()
            A.run_auto(y_key=yk, y_val=yv)
            tot += A.timer_stop()
        h1 = A.speculation_stats()
        mixed_res["on" if spec_on else "off"] = {"ms": round(tot, 4), "hits": h1[0] - h0[0],
                                                 "misses": h1[1] - h0[1]}
    A.set_speculation(True)

    # ---- every fixed slot, same inputs, same cold start
    fixed = None
    if not args.no_fixed:
        fixed = np.zeros((sa.NUM_SLOTS, len(DENSITIES)))
        for s in range(sa.NUM_SLOTS):
            for i in range(len(DENSITIES)):
                timed_call(i, s)  # warm-up of this slot
                fixed[s, i] = float(np.median([timed_call(i, s)[0] for _ in range(args.reps_fixed)]))
        best_per_density = fixed.min(axis=0)
        best_fixed_slot = int(np.argmin(fixed.sum(axis=1)))
        sel_ms = np.array([np.median(v) for v in per_density])
        regret = sel_ms / best_per_density - 1.0

    # ---- roofline of the dominant kernel: the slot the selector spends the
    # most time in over the step, over all of its launches in the timed
    # steps: algorithmic bytes per launch / average launch duration. For the
    # row slots the duration is the tile kernel's own (HIP events on the
    # stream it is launched on, recorded during the timed steps); for the
    # column slots, whose calls are chains of short kernels, the whole call.
    med = [float(np.median(v)) for v in per_density]
    slot_ms = {}
    for i in range(len(DENSITIES)):
        slot_ms[chosen[i]] = slot_ms.get(chosen[i], 0.0) + med[i]
    dom_slot = max(slot_ms, key=slot_ms.get)
    dom_is = [i for i in range(len(DENSITIES)) if chosen[i] == dom_slot]
    launches = [(alg_bytes(xs[i][0], meta[i], ynnzs[i]), km) for i in dom_is for km in kern_ms[i]]
    timed_kernel = dom_slot in ROW_TILED and len(launches) > 0
    if not timed_kernel:
        launches = [(alg_bytes(xs[i][0], meta[i], ynnzs[i]), cm) for i in dom_is for cm in per_density[i]]
    dom_bytes = float(np.mean([b for b, _ in launches]))
    dom_ms = float(np.mean([t for _, t in launches]))
    call_ms = float(np.mean([cm for i in dom_is for cm in per_density[i]]))
    dom_gbs = dom_bytes / (dom_ms * 1e-3) / 1e9
    fmt = ""
    if dom_slot in ROW_TILED and tuned and tuned["spmv_type"] == 1:
        fmt = f"-csr5-s{tuned['sigma']}"
    # measured HBM bytes per launch: the mean of the committed FETCH_SIZE +
    # WRITE_SIZE profiles of this kernel at each density it ran at (null
    # unless every one of them is there)
    traffic, traffic_src = None, []
    for i in dom_is:
        tpath = os.path.join(ROOT, "profiles",
                             f"{sa.KERNEL_NAMES[dom_slot]}{fmt}_rmat{args.scale}_d{DENSITIES[i]}.json")
        if not os.path.exists(tpath) and fmt.startswith("-csr5-s"):
            # the autotuner's CSR5 width can differ between runs where the
            # widths are within 2-3 %: the profile of another width
            alts = sorted(glob.glob(os.path.join(ROOT, "profiles",
                                                 f"{sa.KERNEL_NAMES[dom_slot]}-csr5-s*_rmat{args.scale}_d{DENSITIES[i]}.json")))
            tpath = alts[0] if alts else tpath
        if not os.path.exists(tpath):
            traffic_src = None
            break
        with open(tpath) as f:
            tj = json.load(f)
        traffic_src.append((tj["read_bytes"] + tj["write_bytes"], os.path.relpath(tpath, ROOT)))
    if traffic_src:
        traffic = float(np.mean([t for t, _ in traffic_src]))
        traffic_src = [p for _, p in traffic_src]

    bfs = run_bfs(sa, args, A, rp, col, n) if args.bfs_sources > 0 and not args.no_extra else None
    target = run_target(sa, args, oracle) if args.bfs_scale and not args.no_extra else None

    # ---- PageRank driver on the sweep's graph: learned per-iteration slot
    # vs fixed slots
    pagerank = None
    if not args.no_extra:
        rank = sa.DeviceArray(n, np.float32)
        pagerank = {"graph": f"rmat{args.scale}", "damping": 0.85, "eps": 1e-4}
        for pol in (-1, sa.K_SPMV_LB, sa.K_COL_LB_SORT):
            st, sizes = A.pagerank(rank, 0.85, 1e-4, 1000, pol)
            pagerank["learned" if pol < 0 else sa.KERNEL_NAMES[pol]] = {
                "iterations": int(st[0]), "seconds": round(float(st[1]), 4),
                "x_entries": int(st[2]), "first_sizes": [int(v) for v in sizes[:3]],
                "last_sizes": [int(v) for v in sizes[-3:]]}
        del rank

    c1 = run_c1(sa, args, oracle) if not args.no_extra else None

    # ---- CPU baseline: the reference's CPU SpMSpV (oracle port, 1 core)
    cpu, cpu_all = None, None
    model, cores = cpu_info()
    if not args.no_cpu:
        cp, ri, cv = rp, col, val  # symmetric: the CSR is the CSC
        # a bounded sample: whole sweeps of the 6 densities until about 8 s
        # of single-core CPU work, then the same calls on every core
        t_cpu, b_cpu, calls = 0.0, 0.0, 0
        while (t_cpu < 8.0 or calls % len(DENSITIES)) and calls < len(DENSITIES) * 5000:
            i = calls % len(DENSITIES)
            k, _, _, key, xv = xs[i]
            t0 = time.perf_counter()
            _, ck, _ = oracle.spmspv_csc_host_f32(n, cp, ri, cv, key, xv)
            t_cpu += time.perf_counter() - t0
            b_cpu += alg_bytes(k, meta[i], ck.size)
            calls += 1
        cpu = {"value": b_cpu / t_cpu / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"{calls} calls of SpmSpvCSCHost (FP32, single thread) cycling the 6 densities "
                         f"of the same R-MAT s{args.scale} workload, {t_cpu:.1f} s",
               "cpu_model": model}
        t_all, b_all = 0.0, 0.0
        for c in range(calls):
            i = c % len(DENSITIES)
            k, _, _, key, xv = xs[i]
            t0 = time.perf_counter()
            _, ck, _ = oracle.spmspv_csc_host_f32_omp(n, cp, ri, cv, key, xv, cores)
            t_all += time.perf_counter() - t0
            b_all += alg_bytes(k, meta[i], ck.size)
        cpu_all = {"value": round(b_all / t_all / 1e9, 3), "unit": "GB/s", "cores": cores,
                   "kind": "the build's OpenMP port (not the reference, which is single-threaded)",
                   "sample": f"the same {calls} calls, {t_all:.2f} s", "cpu_model": model}

    out = {
        "metric": "SpMSpV/BFS GB/s (% HBM peak) + GTEPS: learned selector vs best fixed kernel",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rmat{args.scale}_density_sweep", "scale": args.scale, "edge_factor": 16,
                   "nnz": int(M.nnz), "densities": DENSITIES, "selector": sel_kind,
                   "timing": "cold per call (512 MB MALL flush before each call, outside the events)"},
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 5),
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": sa.KERNEL_NAMES[dom_slot] + fmt,
                     "densities": [DENSITIES[i] for i in dom_is], "launches": len(launches),
                     "launch_ms": round(dom_ms, 5), "call_ms": round(call_ms, 5),
                     "timed": "tile kernel alone, during the timed steps" if timed_kernel else "whole call",
                     "alg_bytes": dom_bytes},
        "cpu_baseline": cpu,
        "warm_value": round(warm_value, 3),
        "per_density": [{"density": DENSITIES[i], "xnnz": xs[i][0], "bin_len": int(meta[i]), "ynnz": int(ynnzs[i]),
                         "slot": sa.KERNEL_NAMES[chosen[i]], "ms": round(float(np.median(per_density[i])), 5),
                         "gbs": round(alg_bytes(xs[i][0], meta[i], ynnzs[i]) / (np.median(per_density[i]) * 1e-3) / 1e9, 2)}
                        for i in range(len(DENSITIES))],
        "setup_s": round(setup_s, 2),
        "autotuned": tuned,
        "ceilings": {k: round(v, 2) for k, v in ceilings.items()},
        "errors_vs_fp64": {"bound": "|y - y64| <= 1e-5 * sum|a_ij x_j|", "per_density": errors,
                           "max_rel_err": max((e["max_rel_err"] or 0.0) for e in errors),
                           "all_index_sets_exact": all(e["index_sets_exact"] for e in errors)},
        "speculation": {"timed_steps": {"hits": spec1[0] - spec0[0], "misses": spec1[1] - spec0[1]},
                        "mixed_order_24_calls": mixed_res},
        "cpu_baseline_allcore": cpu_all,
    }
    if fixed is not None:
        out["fixed_ms"] = {sa.KERNEL_NAMES[s]: [round(float(v), 5) for v in fixed[s]] for s in range(sa.NUM_SLOTS)}
        out["best_fixed"] = {"kernel": sa.KERNEL_NAMES[best_fixed_slot],
                             "gbs": round(step_bytes / (fixed[best_fixed_slot].sum() * 1e-3) / 1e9, 3)}
        out["oracle_best_gbs"] = round(step_bytes / (best_per_density.sum() * 1e-3) / 1e9, 3)
        out["selector_regret_geomean"] = round(float(np.exp(np.mean(np.log1p(np.maximum(regret, 0))))) - 1, 4)
    if bfs is not None:
        out["bfs"] = bfs
    if target is not None:
        out["rmat22_target"] = target
    if pagerank is not None:
        out["pagerank"] = pagerank
    if c1 is not None:
        out["c1"] = c1
    if not args.no_batched and not args.no_extra:
        _BFS_GRAPH.clear()  # free the scale-22 graph before the scale-24 one
        del M, A
        out["batched"] = run_batched_bench(sa, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
