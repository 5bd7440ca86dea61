#!/usr/bin/env python3
"""Selector corpus, labels and training (the learned part of the engine).

The reference labels each (matrix, x) pair with the fastest of its kernels
(model/1-generate_matrix_class.py:36-38, 100) and fits sklearn trees offline
(model/1-test_dtree.py:170-196), scoring them by loss_ratio =
(t_pred - t_best) / t_best (model/1-test_dtree.py:300-317). Here:

  corpus  builds synthetic matrices on the host, autotunes each one (as a
          user would once per matrix), times all eight slots on the GPU as
          the selector dispatches them (spmspv_run_auto pinned to the slot:
          the selector launch and speculative queueing included; cold cache,
          median of reps) and records the 16 device features and the rule
          fallback's choice for each (matrix, x):
            uniform random (n = 2^14..2^21 log-uniform, mean nnz/col 2..64),
            R-MAT (a in [0.45, 0.65], d = 0.05, b = c),
            banded (bandwidth 3..201), block-diagonal (dense 32x32 blocks);
          4 x densities per matrix, log-uniform in [1e-5, 1]; matrix seeds
          0..1999 train, 2000..2399 held out. Sizes are capped at 2^25
          nonzeros per matrix so that a shard fits one GPU session.
  bfs-corpus  BFS level labels: symmetric R-MAT / uniform graphs, a few
          sources each, every level timed under the 8 fixed slots
          (spmspv_bfs_trace), features of each level's frontier; trained
          with `train` into the BFS levels' own weights.
  train   trains the MLP 16->64->64->8 on the GPU with the library's own
          kernels (spmspv_mlp_train) on soft targets
          p_k ~ exp(-(t_k / t_best - 1) / tau), then reports held-out
          regret against oracle-best, the best single fixed kernel and the
          rule fallback, and writes the weights.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import spmspv_amd as sa  # noqa: E402

NNZ_CAP = 1 << 25


def matrix_spec(seed):
    rng = np.random.default_rng(seed + 1_000_003)
    kind = ["uniform", "rmat", "banded", "blockdiag"][seed % 4]
    lg = rng.uniform(14, 21)
    n = int(2 ** lg)
    if kind == "uniform":
        mean = float(np.exp(rng.uniform(np.log(2), np.log(64))))
        mean = min(mean, NNZ_CAP / n)
        return dict(kind=kind, n=n, mean=mean)
    if kind == "rmat":
        scale = int(round(lg))
        a = float(rng.uniform(0.45, 0.65))
        ef = int(np.exp(rng.uniform(np.log(2), np.log(32))))
        ef = max(1, min(ef, NNZ_CAP // (2 << scale)))
        return dict(kind=kind, scale=scale, a=a, ef=ef)
    if kind == "banded":
        bw = int(rng.integers(3, 202))
        n = min(n, NNZ_CAP // (2 * bw + 1))
        return dict(kind=kind, n=n, bw=bw)
    n = min(n, NNZ_CAP // 32)
    return dict(kind=kind, n=n, bs=32)


def build_matrix(spec, seed):
    k = spec["kind"]
    if k == "uniform":
        rp, col, val = sa.gen_random_csr(spec["n"], spec["n"], spec["mean"], seed)
        n = spec["n"]
    elif k == "rmat":
        a = spec["a"]
        b = c = (0.95 - a) / 2
        rp, col, val = sa.gen_rmat_csr(spec["scale"], spec["ef"], a, b, c, seed)
        n = 1 << spec["scale"]
    elif k == "banded":
        rp, col, val = sa.gen_banded_csr(spec["n"], spec["bw"], seed)
        n = spec["n"]
    else:
        rp, col, val = sa.gen_blockdiag_csr(spec["n"], spec["bs"], seed)
        n = spec["n"]
    return n, rp, col, val


def time_slot(A, slot, yk, yv, reps, flush, dispatched=True, reset=None):
    """Median cold time of one slot. dispatched: the call as the selector
    dispatches it (spmspv_run_auto with the selector pinned to the slot, so
    the label carries the selector launch and the speculative queueing the
    slot gets in use), else a plain spmspv_run of the slot. reset (when
    given) sets x again before every call, so that each call converts x
    afresh (the time c of the reference's accounting); without it the x
    formats of the previous call are reused (its c_single)."""
    ts = []
    if dispatched:
        A.selector_mode(2, slot)
    for r in range(reps + 1):
        if reset:
            reset()
        if flush:
            A.flush_cache(flush)
        A.timer_start()
        if dispatched:
            A.run_auto(y_key=yk, y_val=yv)
        else:
            A.run(slot, y_key=yk, y_val=yv)
        ms = A.timer_stop()
        if r:
            ts.append(ms)
    return float(np.median(ts))


def cmd_corpus(args):
    feats, times, single, rule, meta = [], [], [], [], []
    t0 = time.time()
    flush = args.flush_mb << 20

    def save():
        np.savez(args.out, feats=np.array(feats, np.float32).reshape(-1, sa.MLP_F),
                 times=np.array(times, np.float64).reshape(-1, sa.NUM_SLOTS),
                 times_single=np.array(single, np.float64).reshape(-1, sa.NUM_SLOTS),
                 rule=np.array(rule, np.int32), meta=np.array(meta, np.int64).reshape(-1, 4))

    for seed in range(args.start, args.start + args.count):
        if args.budget_s and time.time() - t0 > args.budget_s:
            print(f"corpus stopped at seed {seed} (time budget)", flush=True)
            break
        spec = matrix_spec(seed)
        n, rp, col, val = build_matrix(spec, seed)
        if rp[-1] == 0:
            continue
        M = sa.Matrix(n, n, rp, col, val)
        A = M.A
        yk = sa.DeviceArray(n, np.int32)
        yv = sa.DeviceArray(n, np.float32)
        if args.autotune:
            # the per-matrix tuning a user runs once (team sizes, row tile
            # width, K5/K7 format), on a dense x, before any label is timed
            key = np.arange(n, dtype=np.int32)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(np.ones(n, np.float32), np.float32)
            A.set_sparse_x(dk, dv, n)
            A.autotune(reps=2)
        rng = np.random.default_rng(seed + 7)
        for j in range(4):
            d = float(np.exp(rng.uniform(np.log(1e-5), 0.0)))
            k = max(1, min(n, int(round(d * n))))
            key, xv = sa.gen_sparse_vector(n, k, seed * 16 + j)
            dk = sa.DeviceArray.from_host(key, np.int32)
            dv = sa.DeviceArray.from_host(xv, np.float32)
            A.set_sparse_x(dk, dv, k)
            A.selector_mode(1)
            rule.append(A.select())
            feats.append(A.features())
            def reset():
                A.set_sparse_x(dk, dv, k)
            times.append([time_slot(A, s, yk, yv, args.reps, flush, args.dispatched, reset)
                          for s in range(sa.NUM_SLOTS)])
            single.append([time_slot(A, s, yk, yv, args.reps, flush, args.dispatched)
                           for s in range(sa.NUM_SLOTS)] if args.single else times[-1])
            A.selector_mode(1)
            meta.append([seed, n, int(rp[-1]), k])
        del M, A
        if (seed - args.start) % 25 == 0:
            save()
            print(f"corpus seed {seed} {spec} {time.time() - t0:.1f}s", flush=True)
    save()
    print(f"corpus {len(feats)} samples -> {args.out} in {time.time() - t0:.1f}s", flush=True)


def bfs_graph_spec(seed):
    rng = np.random.default_rng(seed + 2_000_003)
    if seed % 2 == 0:
        scale = int(rng.integers(14, 22))
        a = float(rng.uniform(0.45, 0.65))
        ef = int(np.exp(rng.uniform(np.log(4), np.log(32))))
        ef = max(1, min(ef, NNZ_CAP // (2 << scale)))
        return dict(kind="rmat", scale=scale, a=a, ef=ef)
    n = int(2 ** rng.uniform(14, 21))
    mean = float(np.exp(rng.uniform(np.log(2), np.log(16))))
    return dict(kind="uniform", n=n, mean=min(mean, NNZ_CAP / (2 * n)))


def build_bfs_graph(spec, seed):
    """Symmetric pattern (BFS follows edges both ways)."""
    if spec["kind"] == "rmat":
        a = spec["a"]
        b = c = (0.95 - a) / 2
        rp, col, val = sa.gen_rmat_csr(spec["scale"], spec["ef"], a, b, c, seed)
        return 1 << spec["scale"], rp, col, val
    import scipy.sparse as sp
    n = spec["n"]
    rp, col, val = sa.gen_random_csr(n, n, spec["mean"], seed)
    A = sp.csr_matrix((np.ones(col.size, np.float32), col, rp), shape=(n, n))
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    A.sort_indices()
    return n, A.indptr.astype(np.int32), A.indices.astype(np.int32), np.ones(A.nnz, np.float32)


def rule_from_features(f):
    """The rule fallback (tree_select in csrc/selector.hpp) on recorded
    features: the inputs it reads are recovered from their log2 forms."""
    n = 2.0 ** f[1]
    nnz = 2.0 ** f[2] - 1.0
    avg = 2.0 ** f[4] - 1.0
    std = f[5] * (avg + 1e-9)
    xnnz = 2.0 ** f[8] - 1.0
    bl = 2.0 ** f[9] - 1.0
    mx = 2.0 ** f[10] - 1.0
    density = xnnz / max(n, 1.0)
    skewed = std > 2.0 * avg + 1.0 or mx > 8.0 * (avg + 1.0)
    if density >= 0.3:
        return 7 if skewed else 6
    if density >= 0.05:
        return 5 if skewed else 4
    if bl <= 3072:
        return 1
    return 2 if skewed else 0


def cmd_bfs_corpus(args):
    """BFS level labels: every level of a BFS from a source runs the same
    frontier whatever slot runs it (levels are unique), so the 8 fixed-slot
    runs give the 8 times of each level; its features are those the BFS loop's
    selector computes for that frontier."""
    feats, times, rule, meta = [], [], [], []
    t0 = time.time()

    def save():
        np.savez(args.out, feats=np.array(feats, np.float32).reshape(-1, sa.MLP_F),
                 times=np.array(times, np.float64).reshape(-1, sa.NUM_SLOTS),
                 rule=np.array(rule, np.int32), meta=np.array(meta, np.int64).reshape(-1, 4))

    for seed in range(args.start, args.start + args.count):
        if args.budget_s and time.time() - t0 > args.budget_s:
            print(f"bfs corpus stopped at seed {seed} (time budget)", flush=True)
            break
        spec = bfs_graph_spec(seed)
        n, rp, col, val = build_bfs_graph(spec, seed)
        if rp[-1] == 0:
            continue
        M = sa.Matrix(n, n, rp, col, val)
        A = M.A
        lev = sa.DeviceArray(n, np.int32)
        deg = np.diff(rp)
        cand = np.nonzero(deg > 0)[0]
        rng = np.random.default_rng(seed + 11)
        for src in rng.choice(cand, size=min(args.sources, cand.size), replace=False):
            per = []
            f0 = None
            for pol in range(sa.NUM_SLOTS):
                A.bfs_trace(int(src), lev, policy=pol, use_graphs=1)  # graphs captured, caches warm
                st, ms, ft = A.bfs_trace(int(src), lev, policy=pol, use_graphs=1)
                per.append(ms)
                if f0 is None:
                    f0 = ft
            nl = min(len(p) for p in per)
            for L in range(nl):
                feats.append(f0[L])
                times.append([float(per[p][L]) for p in range(sa.NUM_SLOTS)])
                rule.append(rule_from_features(f0[L]))
                meta.append([seed, n, int(rp[-1]), int(src)])
        del M, A
        save()
        print(f"bfs corpus seed {seed} {spec} {len(feats)} levels {time.time() - t0:.1f}s", flush=True)
    save()
    print(f"bfs corpus {len(feats)} levels -> {args.out} in {time.time() - t0:.1f}s", flush=True)


def soft_targets(times, tau):
    rel = times / times.min(axis=1, keepdims=True) - 1.0
    z = np.exp(-rel / tau)
    return (z / z.sum(axis=1, keepdims=True)).astype(np.float32)


def logits(w, X):
    F = sa.MLP_F
    mean, istd = w[sa.MLP_NPARAM:sa.MLP_NPARAM + F], w[sa.MLP_NPARAM + F:]
    Xn = np.ascontiguousarray((X - mean) * istd, np.float32)
    z = np.zeros((X.shape[0], sa.MLP_C), np.float32)
    rc = sa.lib().spmspv_mlp_forward(w.ctypes.data_as(ctypes.c_void_p), Xn.ctypes.data_as(ctypes.c_void_p),
                                     X.shape[0], z.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise sa.SpmspvError(rc, "mlp forward")
    return z


def predict(w, X):
    return logits(w, X).argmax(axis=1)


FAMILY = np.array([0, 0, 0, 0, 1, 1, 2, 2])  # column, row SpMSpV, SpMV (csrc/kernels.hpp slot_family)


def switching_choices(z, seq, penalty):
    """The device's switching-aware choice (csrc/selector.hpp) replayed on
    logits z in call order: keep the previous call's family when its best
    slot is within `penalty` of the top logit; seq marks the sequences."""
    out = np.zeros(len(z), np.int64)
    prev = -1
    for i in range(len(z)):
        if i and seq[i] != seq[i - 1]:
            prev = -1
        best = int(np.argmax(z[i]))
        if penalty > 0 and prev >= 0 and FAMILY[best] != FAMILY[prev]:
            same = np.nonzero(FAMILY == FAMILY[prev])[0]
            k = int(same[np.argmax(z[i][same])])
            if z[i][k] >= z[i][best] - penalty:
                best = k
        out[i] = best
        prev = best
    return out


def real_loss_ratio(times, single, seq, choice):
    """The reference's switching-aware accounting (model/1-test_dtree.py:
    266-287, 312-313): a call costs its plain time c when its kernel family
    differs from the previous call's in the sequence and c_single (x formats
    kept from the previous call) otherwise; per sequence (pred_real -
    best_real) / best with best the per-call oracle-best plain time, the
    oracle itself charged the same way; averaged over the sequences."""
    ratios = []
    for sq in np.unique(seq):
        idx = np.nonzero(seq == sq)[0]
        best_k = times[idx].argmin(axis=1)
        best_plain = float(times[idx, best_k].sum())
        pr, br = 0.0, 0.0
        for j, i in enumerate(idx):
            ck, bk = int(choice[i]), int(best_k[j])
            same_p = j > 0 and FAMILY[ck] == FAMILY[int(choice[idx[j - 1]])]
            same_b = j > 0 and FAMILY[bk] == FAMILY[int(best_k[j - 1])]
            pr += single[i, ck] if (j == 0 or same_p) else times[i, ck]
            br += single[i, bk] if (j == 0 or same_b) else times[i, bk]
        ratios.append((pr - br) / best_plain)
    return float(np.mean(ratios))


def geo_regret(times, choice):
    t = times[np.arange(len(choice)), choice]
    return float(np.exp(np.mean(np.log(t / times.min(axis=1)))) - 1.0)


def cmd_train(args):
    parts = [np.load(p, allow_pickle=False) for p in args.corpus]
    feats = np.concatenate([p["feats"] for p in parts])
    times = np.concatenate([p["times"] for p in parts])
    single = np.concatenate([p["times_single"] if "times_single" in p.files else p["times"] for p in parts])
    rule = np.concatenate([p["rule"] for p in parts])
    meta = np.concatenate([p["meta"] for p in parts])
    train = meta[:, 0] < args.holdout_from
    test = ~train
    X, T = feats[train], soft_targets(times[train], args.tau)
    w = np.zeros(sa.MLP_NWEIGHTS, np.float32)
    loss = ctypes.c_float(0)
    t0 = time.time()
    rc = sa.lib().spmspv_mlp_train(np.ascontiguousarray(X).ctypes.data_as(ctypes.c_void_p),
                                   T.ctypes.data_as(ctypes.c_void_p), int(X.shape[0]), args.epochs,
                                   args.batch, args.lr, args.seed, w.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(loss))
    if rc:
        raise sa.SpmspvError(rc, sa.lib().spmspv_last_error(None).decode())
    train_s = time.time() - t0
    rep = {"train_samples": int(train.sum()), "test_samples": int(test.sum()), "final_loss": loss.value,
           "train_seconds": round(train_s, 2), "epochs": args.epochs, "tau": args.tau}
    for name, mask in (("train", train), ("heldout", test)):
        if not mask.any():
            continue
        tm = times[mask]
        pred = predict(w, feats[mask])
        fixed = [geo_regret(tm, np.full(mask.sum(), s)) for s in range(sa.NUM_SLOTS)]
        best_fixed = int(np.argmin(fixed))
        rep[name] = {
            "mlp_regret": geo_regret(tm, pred),
            "rule_regret": geo_regret(tm, rule[mask]),
            "best_fixed_slot": sa.KERNEL_NAMES[best_fixed],
            "best_fixed_regret": fixed[best_fixed],
            "fixed_regrets": {sa.KERNEL_NAMES[s]: fixed[s] for s in range(sa.NUM_SLOTS)},
            "mlp_accuracy": float((pred == tm.argmin(axis=1)).mean()),
            "real_loss_ratio": {
                "mlp": real_loss_ratio(tm, single[mask], meta[mask, 0], pred),
                "best_fixed": real_loss_ratio(tm, single[mask], meta[mask, 0],
                                              np.full(mask.sum(), best_fixed)),
                "mlp_switching_aware": {
                    str(p): real_loss_ratio(tm, single[mask], meta[mask, 0],
                                            switching_choices(logits(w, feats[mask]), meta[mask, 0], p))
                    for p in (0.5, 1.0, 2.0, 4.0)}},
            "mlp_total_ms": float(tm[np.arange(len(pred)), pred].sum()),
            "best_fixed_total_ms": float(tm[:, best_fixed].sum()),
            "oracle_total_ms": float(tm.min(axis=1).sum()),
        }
    np.save(args.out, w)
    print(json.dumps(rep, indent=1), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("corpus")
    c.add_argument("--start", type=int, default=0)
    c.add_argument("--count", type=int, default=100)
    c.add_argument("--reps", type=int, default=3)
    c.add_argument("--flush-mb", type=int, default=512)
    c.add_argument("--out", required=True)
    c.add_argument("--no-single", dest="single", action="store_false",
                   help="skip the second timing with x's formats kept from the previous call (c_single)")
    c.add_argument("--plain-calls", dest="dispatched", action="store_false",
                   help="label with plain spmspv_run times instead of the selector-dispatched call's")
    c.add_argument("--no-autotune", dest="autotune", action="store_false",
                   help="time the slots with the default widths instead of the autotuned ones")
    c.add_argument("--budget-s", type=float, default=0.0,
                   help="stop starting new matrices after this many seconds (results are kept)")
    bc = sub.add_parser("bfs-corpus")
    bc.add_argument("--start", type=int, default=0)
    bc.add_argument("--count", type=int, default=50)
    bc.add_argument("--sources", type=int, default=4)
    bc.add_argument("--out", required=True)
    bc.add_argument("--budget-s", type=float, default=0.0,
                    help="stop starting new graphs after this many seconds (results are kept)")
    t = sub.add_parser("train")
    t.add_argument("--corpus", nargs="+", required=True)
    t.add_argument("--holdout-from", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--tau", type=float, default=0.1)
    t.add_argument("--seed", type=int, default=11)
    t.add_argument("--out", default=os.path.join(HERE, "selector_weights.npy"))
    t.add_argument("--report", default=None)
    args = ap.parse_args()
    {"corpus": cmd_corpus, "bfs-corpus": cmd_bfs_corpus, "train": cmd_train}[args.cmd](args)


if __name__ == "__main__":
    main()
