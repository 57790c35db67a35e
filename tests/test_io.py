"""Graph npz, reorder npz and ME-TCF cache formats (CPU parts)."""
import numpy as np
import pytest
import scipy.sparse as sp


def test_graph_npz_matches_reference_loader_rules(tmp_path):
    from dtc.io import load_graph_npz, save_graph_npz
    rng = np.random.default_rng(0)
    n = 300
    src = rng.integers(0, n, 4000)
    dst = rng.integers(0, n, 4000)
    src = np.concatenate([src, src[:100]])  # duplicate edges
    dst = np.concatenate([dst, dst[:100]])
    p = str(tmp_path / "g.npz")
    save_graph_npz(p, src, dst, n)
    a = load_graph_npz(p)
    ref = sp.coo_matrix((np.ones(src.shape[0]), (src, dst)), shape=(n, n)).tocsr()
    ref.sort_indices()
    assert a.rowptr.dtype == np.int32 and a.col.dtype == np.int32
    assert np.array_equal(a.rowptr, ref.indptr) and np.array_equal(a.col, ref.indices)
    assert np.array_equal(a.values, ref.data.astype(np.float32))
    for r in range(n):
        assert np.all(np.diff(a.col[a.rowptr[r]:a.rowptr[r + 1]]) > 0)
    with pytest.raises(ValueError):
        load_graph_npz(str(tmp_path / "g.txt"))


def test_apply_reorder_relabels_both_ends(tmp_path):
    from dtc.io import apply_reorder, load_reorder_id
    reorder = np.array([3, 0, 2, 1])
    p = str(tmp_path / "x.reorder_id.npz")
    np.savez(p, reorder_id=reorder)
    r = load_reorder_id(p)
    s, d = apply_reorder([3, 0, 1], [1, 2, 3], r)
    # node 3 -> 0, node 0 -> 1, node 2 -> 2, node 1 -> 3
    assert list(s) == [0, 1, 3] and list(d) == [3, 2, 0]


def test_reorder_files_round_trip_and_relabel_agree(tmp_path):
    """reorder_id.npz and reorder.npz written as the reference's script writes
    them (TCA_reorder.py:317, :326); the relabelled CSR equals P A P^T."""
    from dtc.graphs import rmat
    from dtc.io import load_graph_npz, load_reorder_id, save_reorder_id, save_reordered_graph
    from dtc.reorder import relabel
    a = rmat(9, 8, seed=2)
    M = a.num_rows
    order = np.random.default_rng(1).permutation(M)
    save_reorder_id(str(tmp_path / "g.reorder_id.npz"), order)
    assert np.array_equal(load_reorder_id(str(tmp_path / "g.reorder_id.npz")), order)
    src = np.repeat(np.arange(M), np.diff(a.rowptr))
    save_reordered_graph(str(tmp_path / "g.reorder.npz"), src, a.col, M, order)
    g = load_graph_npz(str(tmp_path / "g.reorder.npz"))
    b = relabel(a, order)
    assert np.array_equal(g.rowptr, b.rowptr) and np.array_equal(g.col, b.col)
    P = sp.csr_matrix((np.ones(M), (np.arange(M), order)), shape=(M, M))
    dense = (P @ a.to_scipy() @ P.T).toarray()
    assert np.array_equal(b.to_scipy().toarray(), dense)
    with pytest.raises(ValueError):
        relabel(a, np.zeros(M, np.int64))
