"""CPU side of the selector: the float64 MLP oracle is pinned by central
finite differences, and the committed timing dataset / model are checked
for shape and consistency with the variant list."""
import os

import numpy as np
import pytest

import mlp_oracle as mo


def _params(rng, dims):
    return [(rng.standard_normal((dims[i], dims[i + 1])) * 0.5, rng.standard_normal(dims[i + 1]) * 0.1)
            for i in range(3)]


def test_mlp_oracle_gradients_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((7, 5))
    target = rng.standard_normal((7, 4))
    mask = (rng.random((7, 4)) > 0.2).astype(np.float64)
    params = _params(rng, [5, 6, 6, 4])
    loss, grads = mo.mlp_loss_and_grads(x, target, mask, params)
    h = 1e-6
    for li in range(3):
        for pi in range(2):
            p = params[li][pi]
            g = grads[li][pi]
            it = np.nditer(p, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = p[idx]
                p[idx] = old + h
                lp, _ = mo.mlp_loss_and_grads(x, target, mask, params)
                p[idx] = old - h
                lm, _ = mo.mlp_loss_and_grads(x, target, mask, params)
                p[idx] = old
                assert abs((lp - lm) / (2 * h) - g[idx]) <= 1e-6 * max(1.0, abs(g[idx])), (li, pi, idx)


def test_adam_oracle_first_step():
    # With bias correction the first Adam step moves each weight by lr * sign(g).
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -0.25, 1e-3])
    new, m, v = mo.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    assert np.allclose(new, p - 0.1 * np.sign(g), atol=1e-6)


def test_dataset_and_model_consistent():
    from dtc.selector import DEFAULT_DATA, DEFAULT_MODEL, FEATURE_NAMES, VARIANTS, variant_name
    if not os.path.exists(DEFAULT_DATA):
        pytest.skip("no committed selector dataset")
    z = np.load(DEFAULT_DATA)
    n = z["features"].shape[0]
    assert z["features"].shape == (n, len(FEATURE_NAMES))
    assert z["times_ms"].shape == (n, len(VARIANTS))
    assert [str(s) for s in z["variants"]] == [variant_name(v) for v in VARIANTS]
    assert np.isfinite(z["features"]).all()
    assert (np.isfinite(z["times_ms"]) & (z["times_ms"] > 0)).mean() > 0.99
    assert len(np.unique(z["seed"])) == n
    if os.path.exists(DEFAULT_MODEL):
        m = np.load(DEFAULT_MODEL)
        assert [str(s) for s in m["variants"]] == [variant_name(v) for v in VARIANTS]
        assert int(m["members"]) >= 1
        for k in range(int(m["members"])):
            assert m[f"m{k}_w0"].shape[0] == len(FEATURE_NAMES) and m[f"m{k}_w2"].shape[1] == len(VARIANTS)


def test_model_range_guard():
    """The shipped network carries its training feature range: the training
    matrices and all but a few of the held-out ones lie inside it, and so does
    the 232,965-row block model (121M entries, 10.4M TC blocks, covered by the
    block-model seeds of that size); a matrix four times larger lies outside
    and goes to exhaustive timing."""
    from dtc import selector as sel
    z = np.load(sel.DEFAULT_MODEL)
    d = np.load(sel.DEFAULT_DATA)
    train = sel.is_train_seed(d["seed"])
    X = d["features"].astype(np.float32)
    assert np.array_equal(z["lo"], X[train].min(0)) and np.array_equal(z["hi"], X[train].max(0))
    assert all(sel.out_of_range(x, z["lo"], z["hi"]) == [] for x in X[train])
    assert sum(1 for x in X[~train] if sel.out_of_range(x, z["lo"], z["hi"])) <= 4
    xl = X[d["seed"] == sel.XL_START][0].copy()
    xl[sel.FEATURE_NAMES.index("log2_nnz")] = np.log2(1.21e8)
    xl[sel.FEATURE_NAMES.index("log2_blocks")] = np.log2(1.04e7)
    assert sel.out_of_range(xl, z["lo"], z["hi"]) == []
    xl[sel.FEATURE_NAMES.index("log2_nnz")] = np.log2(4 * 1.21e8)
    xl[sel.FEATURE_NAMES.index("log2_blocks")] = np.log2(4 * 1.04e7)
    assert sel.out_of_range(xl, z["lo"], z["hi"]) == ["log2_nnz", "log2_blocks"]
    assert sel.out_of_range(xl, None, None) == []


def test_default_schedule_rule():
    """The default options' schedule: unbalanced up to AUTO_WINDOW_BLOCKS
    blocks in the longest window, balanced past it; on the held-out timings
    the rule beats balanced-only defaults."""
    from dtc import selector as sel
    from dtc.selector import FEATURE_NAMES
    assert sel.default_schedule(0) == "windows"
    assert sel.default_schedule(sel.AUTO_WINDOW_BLOCKS) == "windows"
    assert sel.default_schedule(sel.AUTO_WINDOW_BLOCKS + 1) == "balanced"
    assert sel.VARIANTS[sel.default_variant(10_000, 12)] == ("windows", "fp32", 0, 4, 1, False, 0)
    assert sel.VARIANTS[sel.default_variant(10_000_000, 5_000)] == ("balanced", "fp32", 512, 2, 1, False, 0)
    x = np.zeros(len(FEATURE_NAMES), np.float32)
    x[FEATURE_NAMES.index("log2_max_blocks_per_window")] = np.log2(161.0)
    assert sel.max_window_blocks_of(x) == 160
    z = np.load(sel.DEFAULT_DATA)
    test = ~sel.is_train_seed(z["seed"])
    t, X, blocks = z["times_ms"][test], z["features"][test], z["blocks"][test]
    best = np.nanmin(t, axis=1)
    rows = np.arange(t.shape[0])
    rule = np.array([sel.default_variant(b, sel.max_window_blocks_of(x)) for b, x in zip(blocks, X)])
    balanced_only = np.array([sel.VARIANTS.index(("balanced", "fp32", sel.default_chunk_blocks(int(b)), 2, 1, False, 0))
                              for b in blocks])
    assert (t[rows, rule] / best).mean() < (t[rows, balanced_only] / best).mean()
