"""The reorder advisor trained with the HIP MLP kernels on the committed
timings: on the held-out matrices, following its advice must beat both
always and never reordering."""
import os

import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


def test_advice_beats_fixed_policies():
    from dtc import reorder_advice as ra
    if not os.path.exists(ra.DEFAULT_DATA):
        pytest.skip("no committed reorder timings")
    data = np.load(ra.DEFAULT_DATA)
    tr = data["seed"] < 1600
    say("reorder advisor training", int(tr.sum()))
    model = ra.ReorderAdvisor(data["features"].shape[1], seed=0)
    hist = model.fit(data["features"][tr], ra.gains(data)[tr], steps=1000)
    assert hist[-1][1] < hist[0][1]
    rep = ra.evaluate(model, data)
    print("[reorder advisor]", rep, flush=True)
    assert rep["mean_speedup_following_advice"] > max(1.0, rep["mean_speedup_always_reorder"])
    assert rep["pred_gain_corr"] > 0.5


def test_shipped_model_loads_and_predicts():
    from dtc import reorder_advice as ra
    from dtc import selector as sel
    if not os.path.exists(ra.DEFAULT_MODEL):
        pytest.skip("no committed advisor")
    m = ra.ReorderAdvisor.load()
    x = np.zeros((3, len(sel.FEATURE_NAMES)), np.float32)
    g = m.predict_gain(x)
    assert g.shape == (3,) and np.isfinite(g).all()
