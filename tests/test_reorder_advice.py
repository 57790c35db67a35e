"""The reorder advisor's committed timings (no GPU needed): every family is
present in both splits, the stored gains follow from the stored times, and
the reordering changed the TC-block counts of structured matrices."""
import numpy as np
import pytest

from dtc import reorder_advice as ra
from dtc import selector as sel


@pytest.fixture(scope="module")
def data():
    import os
    if not os.path.exists(ra.DEFAULT_DATA):
        pytest.skip("no committed reorder timings")
    return np.load(ra.DEFAULT_DATA)


def test_dataset_shapes_and_families(data):
    n = data["features"].shape[0]
    assert data["features"].shape == (n, len(sel.FEATURE_NAMES))
    # timed over the selector's variants of the time of collection: a subset
    # of today's (gains are the best time over whichever were timed)
    nv = len(data["variants"])
    assert data["times_orig_ms"].shape == data["times_reordered_ms"].shape == (n, nv)
    assert set(str(s) for s in data["variants"]) <= set(sel.variant_name(v) for v in sel.VARIANTS)
    assert len(np.unique(data["seed"])) == n
    for split in (data["seed"] < 1600, data["seed"] >= 1600):
        assert set(np.unique(data["family"][split]).tolist()) == set(range(len(sel.FAMILIES)))
    # the family of a seed is seed mod 5 (dtc.selector.matrix_spec)
    assert np.array_equal(data["family"], data["seed"] % len(sel.FAMILIES))


def test_gains_from_times(data):
    g = ra.gains(data)
    want = np.log(np.nanmin(data["times_orig_ms"], 1) / np.nanmin(data["times_reordered_ms"], 1))
    assert np.allclose(g, want)
    assert np.isfinite(g).all()


def test_reordering_changes_structured_matrices(data):
    # a permutation keeps the entries; the block count moves for most matrices
    moved = data["blocks_reordered"] != data["blocks"]
    assert moved.mean() > 0.5
