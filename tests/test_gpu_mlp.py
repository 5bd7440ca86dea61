"""The selector's GPU forward/backward kernels against the CPU FP64 oracle:
loss, logits and every gradient entry, plus FP64 finite differences of the
oracle loss at sampled coordinates; and a training run that must fit."""
import ctypes

import numpy as np
import pytest

import oracle
import spmspv_amd as sa

pytestmark = pytest.mark.gpu
F, H1, H2, C = 16, 64, 64, 8


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gpu_loss_grad(params, X, T):
    g = np.zeros(sa.MLP_NPARAM, np.float32)
    loss = ctypes.c_float(0)
    rc = sa.lib().spmspv_mlp_loss_grad(_p(params), _p(X), _p(T), X.shape[0], ctypes.byref(loss), _p(g))
    assert rc == 0, sa.lib().spmspv_last_error(None)
    return loss.value, g


def test_mlp_gradient_matches_oracle(gpu):
    rng = np.random.default_rng(3)
    B = 37
    params = rng.normal(0, 0.25, sa.MLP_NPARAM).astype(np.float32)
    X = rng.normal(0, 1, (B, F)).astype(np.float32)
    T = rng.dirichlet(np.ones(C), B).astype(np.float32)
    loss, g = gpu_loss_grad(params, X, T)
    l64, g64 = oracle.mlp_loss_grad(F, H1, H2, C, params.astype(np.float64), X.astype(np.float64),
                                    soft=T.astype(np.float64))
    assert abs(loss - l64) <= 1e-5 * max(1.0, abs(l64))
    scale = np.abs(g64).max()
    np.testing.assert_allclose(g, g64, rtol=0, atol=2e-5 * scale)
    # finite differences of the FP64 oracle loss at sampled coordinates
    for i in rng.choice(sa.MLP_NPARAM, 40, replace=False):
        q = params.astype(np.float64)
        eps = 1e-6
        q[i] += eps
        lp, _ = oracle.mlp_loss_grad(F, H1, H2, C, q, X.astype(np.float64), soft=T.astype(np.float64), want_grad=False)
        q[i] -= 2 * eps
        lm, _ = oracle.mlp_loss_grad(F, H1, H2, C, q, X.astype(np.float64), soft=T.astype(np.float64), want_grad=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(g[i] - fd) <= 2e-5 * scale + 1e-4 * abs(fd), (i, g[i], fd)


def test_mlp_forward_matches_oracle(gpu):
    rng = np.random.default_rng(4)
    B = 50
    params = rng.normal(0, 0.25, sa.MLP_NPARAM).astype(np.float32)
    X = rng.normal(0, 1, (B, F)).astype(np.float32)
    z = np.zeros((B, C), np.float32)
    assert sa.lib().spmspv_mlp_forward(_p(params), _p(X), B, _p(z)) == 0
    z64 = oracle.mlp_forward(F, H1, H2, C, params.astype(np.float64), X.astype(np.float64))
    np.testing.assert_allclose(z, z64, rtol=1e-4, atol=1e-4)


def test_mlp_training_fits_a_rule(gpu):
    # label = a simple threshold rule of two features; the trained network
    # must reproduce it on held-out points
    rng = np.random.default_rng(5)
    B = 4000
    X = rng.normal(0, 1, (B, F)).astype(np.float32)
    lab = (X[:, 3] > 0).astype(int) * 2 + (X[:, 11] > 0.5).astype(int)
    T = np.eye(C, dtype=np.float32)[lab]
    w = np.zeros(sa.MLP_NWEIGHTS, np.float32)
    loss = ctypes.c_float(0)
    rc = sa.lib().spmspv_mlp_train(_p(X[:3000]), _p(T[:3000]), 3000, 60, 128, 3e-3, 7, _p(w), ctypes.byref(loss))
    assert rc == 0
    assert loss.value < 0.1
    mean, istd = w[sa.MLP_NPARAM:sa.MLP_NPARAM + F], w[sa.MLP_NPARAM + F:]
    Xn = ((X[3000:] - mean) * istd).astype(np.float32)
    z = np.zeros((1000, C), np.float32)
    assert sa.lib().spmspv_mlp_forward(_p(w), _p(Xn), 1000, _p(z)) == 0
    acc = (z.argmax(1) == lab[3000:]).mean()
    assert acc > 0.95, acc


@pytest.mark.parametrize("which", ["spmspv", "bfs"])
def test_trained_selector_gain_on_heldout(gpu, which):
    """The committed networks, scored on their committed held-out corpora
    (SpMSpV: matrix seeds 2000-2399, every (matrix, x) timed on all eight
    slots; BFS: graph seeds 300-359, every BFS level timed under all eight
    slots): geomean regret against the fastest slot at most 5% and below
    both the rule fallback's and the best single fixed slot's."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmspv-adaptive_amd")
    wname, cname = (("selector_weights.npy", "heldout_seeds_2000_2399.npz") if which == "spmspv" else
                    ("bfs_selector_weights.npy", "bfs_heldout_seeds_300_359.npz"))
    w = np.load(os.path.join(here, wname)).astype(np.float32)
    d = np.load(os.path.join(here, "selector_corpus", cname), allow_pickle=False)
    X, T, rule = d["feats"], d["times"], d["rule"]
    mean, istd = w[sa.MLP_NPARAM:sa.MLP_NPARAM + F], w[sa.MLP_NPARAM + F:]
    Xn = np.ascontiguousarray((X - mean) * istd, np.float32)
    z = np.zeros((X.shape[0], C), np.float32)
    assert sa.lib().spmspv_mlp_forward(_p(w), _p(Xn), X.shape[0], _p(z)) == 0
    pred = z.argmax(axis=1)

    def regret(choice):
        t = T[np.arange(len(choice)), choice]
        return float(np.exp(np.mean(np.log(t / T.min(axis=1)))) - 1.0)

    mlp = regret(pred)
    rules = regret(rule)
    fixed = min(regret(np.full(len(pred), s)) for s in range(C))
    print(f"{which} held-out regret: mlp {mlp:.4f} rules {rules:.4f} best fixed {fixed:.4f}")
    assert mlp <= 0.05
    assert mlp < rules and mlp < fixed


def test_rule_restatement_matches_device_rule(gpu):
    """train_selector.rule_from_features (the rule fallback recomputed from
    recorded features, used to score the BFS corpus) picks the slot the
    device's rule fallback picks, on an R-MAT s12 matrix over a density
    sweep."""
    import train_selector as ts
    rp, col, val = sa.gen_rmat_csr(12, 16, seed=91)
    M = sa.Matrix(1 << 12, 1 << 12, rp, col, val)
    M.A.selector_mode(1)
    for k in (1, 7, 60, 400, 1500, 4096):
        key, v = sa.gen_sparse_vector(M.n, k, 300 + k)
        dk = sa.DeviceArray.from_host(key, np.int32)
        dv = sa.DeviceArray.from_host(v, np.float32)
        M.A.set_sparse_x(dk, dv, k)
        dev = M.A.select()
        f = M.A.features().astype(np.float64)
        assert ts.rule_from_features(f) == dev, f"xnnz {k}: python rule {ts.rule_from_features(f)} device {dev}"

