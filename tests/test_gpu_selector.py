"""Selector on the GPU: the HIP MLP kernels against the float64 oracle, and
the trained selector against exhaustive timing and against the fixed
default variant on the held-out seeds."""
import os

import numpy as np
import pytest
import torch

import mlp_oracle as mo
from helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("batch,din,dout,relu", [(1, 3, 2, False), (37, 25, 64, True), (1600, 64, 17, False),
                                                 (5, 64, 64, True)])
def test_linear_layers_vs_oracle(ext, batch, din, dout, relu):
    rng = np.random.default_rng(batch + din)
    x = rng.standard_normal((batch, din)).astype(np.float32)
    w = rng.standard_normal((din, dout)).astype(np.float32) * 0.3
    b = rng.standard_normal(dout).astype(np.float32) * 0.1
    dy = rng.standard_normal((batch, dout)).astype(np.float32)
    say("mlp layer", batch, din, dout, relu)
    y = ext.mlp_linear_forward(T(x), T(w), T(b), relu=relu)
    y64 = mo.linear_forward(x, w, b, relu)
    assert np.allclose(y.cpu().numpy(), y64, rtol=1e-5, atol=1e-5)
    dx, dw, db = ext.mlp_linear_backward(T(x), T(w), y, T(dy), relu=relu)
    dx64, dw64, db64 = mo.linear_backward(x, w, y.cpu().numpy().astype(np.float64), dy, relu)
    assert np.allclose(dx.cpu().numpy(), dx64, rtol=1e-4, atol=1e-4)
    assert np.allclose(dw.cpu().numpy(), dw64, rtol=1e-4, atol=1e-3)
    assert np.allclose(db.cpu().numpy(), db64, rtol=1e-4, atol=1e-3)


def test_mse_and_adam_vs_oracle(ext):
    rng = np.random.default_rng(3)
    pred = rng.standard_normal((100, 17)).astype(np.float32)
    tgt = rng.standard_normal((100, 17)).astype(np.float32)
    mask = (rng.random((100, 17)) > 0.1).astype(np.float32)
    inv = 1.0 / mask.sum()
    say("mlp mse/adam")
    d, loss = ext.mlp_mse(T(pred), T(tgt), T(mask), inv)
    l64, d64 = mo.mse(pred, tgt, mask, inv)
    assert abs(float(loss.item()) - l64) <= 1e-5 * max(1, l64)
    assert np.allclose(d.cpu().numpy(), d64, rtol=1e-5, atol=1e-7)
    p = rng.standard_normal(50).astype(np.float32)
    g = rng.standard_normal(50).astype(np.float32)
    pd, gd, md, vd = T(p), T(g), torch.zeros(50, device="cuda"), torch.zeros(50, device="cuda")
    m64, v64, p64 = np.zeros(50), np.zeros(50), p.astype(np.float64)
    for step in range(1, 6):
        ext.adam_step(pd, gd, md, vd, step, 1e-2)
        p64, m64, v64 = mo.adam(p64, g.astype(np.float64), m64, v64, step, 1e-2)
    assert np.allclose(pd.cpu().numpy(), p64, rtol=1e-5, atol=1e-6)


def test_training_matches_oracle_steps(ext):
    """Five full training steps on the GPU equal five float64 oracle steps."""
    from dtc.selector import MLPSelector
    rng = np.random.default_rng(4)
    X = rng.standard_normal((64, 25)).astype(np.float32)
    t = np.exp(rng.standard_normal((64, 17))).astype(np.float32)
    say("mlp training steps")
    m = MLPSelector(25, 17, seed=1)
    params64 = [(w.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)) for w, b in m.params]
    m.fit(X, t, steps=5, lr=1e-3, log_every=1)
    Xn = (X - X.mean(0)) / (X.std(0) + 1e-6)
    tgt, mask = MLPSelector.targets(t)
    state = [[np.zeros_like(p) for p in pair] for pair in params64]
    vstate = [[np.zeros_like(p) for p in pair] for pair in params64]
    params64 = [list(p) for p in params64]
    for step in range(1, 6):
        _, grads = mo.mlp_loss_and_grads(Xn, tgt, mask, params64)
        for li in range(3):
            for pi in range(2):
                params64[li][pi], state[li][pi], vstate[li][pi] = mo.adam(
                    params64[li][pi], grads[li][pi], state[li][pi], vstate[li][pi], step, 1e-3)
    for li in range(3):
        for pi in range(2):
            assert np.allclose(m.params[li][pi].cpu().numpy(), params64[li][pi], rtol=1e-3, atol=1e-5)


def test_selector_regret_and_gain(ext):
    """Train on the training seeds with the HIP kernels; on the held-out seeds the
    selector must beat the library's default options (the method without it:
    the "auto" schedule rule) and the best single fixed variant."""
    from dtc.selector import DEFAULT_DATA, MLPSelector, evaluate, split
    if not os.path.exists(DEFAULT_DATA):
        pytest.skip("no committed selector dataset")
    data = np.load(DEFAULT_DATA)
    tr, te = split(data)
    say("selector training", int(tr.sum()), "train /", int(te.sum()), "held out")
    model = MLPSelector(data["features"].shape[1], data["times_ms"].shape[1], seed=0)
    hist = model.fit(data["features"][tr], data["times_ms"][tr], steps=3000)
    assert hist[-1][1] < hist[0][1]
    rep = evaluate(model, data)
    print("[selector]", rep, flush=True)
    assert rep["selector_mean_regret"] < rep["default_mean_regret"]
    assert rep["selector_mean_regret"] < rep["best_fixed_mean_regret"]
    assert rep["mean_speedup_over_best_fixed"] > 1.0
    # timing the network's top choices on the matrix itself only removes regret
    assert rep["top3_timed_mean_regret"] <= rep["selector_mean_regret"]


def test_pick_variant_times_the_top_choices(ext):
    """dtc.selector.pick_variant on a matrix inside the training range times
    the network's top 3 predictions and returns the fastest of them; with
    top_k=1 it returns the network's first choice."""
    from dtc.matrix import DTCMatrix
    from dtc.selector import DEFAULT_MODEL, MLPSelector, TOP_K, features, make_matrix, pick_variant
    if not os.path.exists(DEFAULT_MODEL):
        pytest.skip("no committed selector model")
    model = MLPSelector.load(DEFAULT_MODEL)
    # a training matrix (seed 6: R-MAT), inside the training range by construction
    _, family, N, rowptr, col, vals, M, K = make_matrix(6)
    A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=K)
    B = torch.randn((K, N), device="cuda")
    x = features(A, N)
    assert model.out_of_range(x) == [], family
    say("selector pick_variant")
    top = [int(i) for i in model.top_k(x[None, :], TOP_K)[0]]
    assert top[0] == int(model.choose(x[None, :])[0]) and len(set(top)) == TOP_K
    pick, how = pick_variant(A, B, model)
    assert pick in top and "top" in how
    assert pick_variant(A, B, model, top_k=1) == (top[0], "selector")
