"""GPU TCA reordering against the CPU oracle: signatures, LSH pairs and
Jaccard similarities bit-exact, the whole two-level reorder identical, and
the reordered matrix's ME-TCF no larger on planted and power-law inputs."""
import numpy as np
import pytest
import torch

import reorder_oracle as ro
from helpers import say
from test_reorder import _csr, planted_blocks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def _cases():
    from dtc.graphs import block_diagonal, rmat, sbm_hubs
    out = []
    a = rmat(10, 16, seed=3)
    out.append(("rmat10", a.rowptr, a.col, a.num_rows))
    a = rmat(12, 8, seed=4)
    out.append(("rmat12", a.rowptr, a.col, a.num_rows))
    lists, _ = planted_blocks(40, 16, 48, seed=2)
    p, i = _csr(lists)
    out.append(("planted", p, i, len(lists)))
    a = block_diagonal(3000, 64, fill=0.2, noise=0.01, seed=5)
    rng = np.random.default_rng(0)
    from dtc.reorder import relabel
    a = relabel(a, rng.permutation(a.num_rows))
    out.append(("blockdiag_shuffled", a.rowptr, a.col, a.num_rows))
    a = sbm_hubs(num_rows=4000, communities=8, avg_per_row=24, hubs=3, hub_min=300, hub_max=600, seed=6)
    out.append(("sbm", a.rowptr, a.col, a.num_rows))
    p, i = _csr([[], [1, 2], [], [1, 2], [5]])
    out.append(("tiny_with_empty", p, i, 5))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_stages_bit_exact(ext, case):
    name, rowptr, col, M = case
    say("reorder stages", name, M, col.shape[0])
    rp = torch.from_numpy(np.asarray(rowptr, np.int32)).cuda()
    cc = torch.from_numpy(np.asarray(col, np.int32)).cuda()
    for K, seed in ((56, 2022), (128, 9), (3, 1)):
        sig = ext.minhash(rp, cc, K, seed).cpu().numpy().view(np.uint32)
        assert np.array_equal(sig, ro.minhash_signatures(rowptr, col, M, K, seed)), (name, K)
    sig_np = ro.minhash_signatures(rowptr, col, M, 56, 2022)
    sig = ext.minhash(rp, cc, 56, 2022)
    live = np.diff(rowptr) > 0
    for bands, r, D in ((28, 2, 8), (56, 1, 2), (14, 3, 4)):
        got = ext.lsh_pairs(sig, rp, bands, r, D, 0).cpu().numpy().view(np.uint64)
        want = ro.lsh_pairs(sig_np, live, bands, r, D)
        assert np.array_equal(got, want), (name, bands, r, D)
    pairs = ext.lsh_pairs(sig, rp, 28, 2, 8, 0)
    sims = ext.jaccard_pairs(rp, cc, pairs).cpu().numpy()
    assert np.array_equal(sims, ro.jaccard_pairs(rowptr, col, pairs.cpu().numpy().view(np.uint64)))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_reorder_matches_oracle(ext, case):
    from dtc.reorder import tca_reorder
    name, rowptr, col, M = case
    say("tca_reorder", name)
    got, stats = tca_reorder(rowptr, col, M)
    want = ro.tca_reorder(rowptr, col, M)
    assert np.array_equal(got, want), name
    assert sorted(got.tolist()) == list(range(M))
    assert stats["clusters"] >= stats["super_clusters"] >= 1


def test_planted_blocks_recovered(ext):
    from dtc.reorder import tca_reorder
    lists, _ = planted_blocks(64, 16, 40, seed=7)
    p, i = _csr(lists)
    order, _ = tca_reorder(p, i, len(lists))
    p2, i2 = _csr([lists[k] for k in order])
    assert ro.sum_unique_columns(p2, i2, len(lists)) == 64 * 20


def test_rmat_sum_unique_columns_drops(ext):
    """R-MAT scale 14: the reorder lowers sum_w U_w (fewer distinct B rows
    per window) and the ME-TCF build and SpMM on the relabelled graph agree
    with the float64 product of the relabelled matrix."""
    import oracle
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix
    from dtc.reorder import relabel, tca_reorder
    a = rmat(14, 16, seed=1)
    order, stats = tca_reorder(a.rowptr, a.col, a.num_rows)
    b = relabel(a, order)
    before = ro.sum_unique_columns(a.rowptr, a.col, a.num_rows)
    after = ro.sum_unique_columns(b.rowptr, b.col, b.num_rows)
    say("rmat14 sum U", before, "->", after, stats)
    assert after < 0.95 * before
    A = DTCMatrix(b.rowptr, b.col, b.values, num_rows=b.num_rows, num_cols=b.num_cols)
    assert A.sum_unique_columns() == after
    B = torch.randn((b.num_rows, 64), device="cuda")
    C = A.spmm(B).cpu().numpy()
    C64, Cabs = oracle.spmm_csr_f64(b.rowptr, b.col, b.values, B.cpu().numpy(), b.num_rows, with_abs=True)
    m = Cabs > 0
    assert float((np.abs(C - C64)[m] / Cabs[m]).max()) <= 4e-6


def test_reorder_rejects_unsorted_rows(ext):
    rp = torch.tensor([0, 2], dtype=torch.int32, device="cuda")
    cc = torch.tensor([3, 1], dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="strictly increasing"):
        ext.tca_reorder(rp, cc, 16, 128, 28, 2, 8, 2022, 0)


def test_lsh_overflow_is_an_error(ext):
    from dtc.graphs import rmat
    a = rmat(10, 16, seed=3)
    rp = torch.from_numpy(a.rowptr).cuda()
    sig = ext.minhash(rp, torch.from_numpy(a.col).cuda(), 56, 2022)
    with pytest.raises(RuntimeError, match="max_pairs"):
        ext.lsh_pairs(sig, rp, 28, 2, 8, 3)
