"""TCA row reordering: the oracle pinned by hand-derived and mathematical
checks, and the host greedy merge of libdtc against it (no GPU needed).

The reference's own signatures and candidate sets come from libraries that
are absent here (datasketch, libMHCUDA, cugraph), so the reorder as a whole
has no reference golden vectors; the clustering loop is pinned by a
hand-derived fixture, the hashing and LSH stages by their defining
properties, and the end result by planted-structure recovery."""
import itertools

import numpy as np
import pytest
import torch

import reorder_oracle as ro


def _csr(lists):
    ptr = np.zeros(len(lists) + 1, np.int32)
    ptr[1:] = np.cumsum([len(x) for x in lists])
    idx = np.concatenate([np.asarray(x, np.int32) for x in lists]) if ptr[-1] else np.zeros(0, np.int32)
    return ptr, idx


def _pack(pairs):
    return np.array([(a << 32) | b for a, b in pairs], np.uint64)


# Hand simulation of TCA_reorder.py:152-187 with ties popped in (p1, p2)
# order: pairs (0,1) 1.0, (1,2) 2/3, (3,4) 2/3, (2,3) 0.0, thres 3.  Pops:
# (0,1) merges (equal sizes: 0 absorbs 1); of the tied pairs (1,2) comes
# first and is stale -> re-queued as (0,2) with jd({1,2,3},{1,2}) = 2/3;
# (0,2) ties with (3,4) and comes first: merges and freezes cluster 0 at
# size 3; (3,4) merges; (2,3) is stale and its root 0 is frozen -> dropped.
GREEDY_LISTS = [[1, 2, 3], [1, 2, 3], [1, 2], [7, 8], [7, 8, 9], []]
GREEDY_PAIRS = [(0, 1, 1.0), (1, 2, 2 / 3), (3, 4, 2 / 3), (2, 3, 0.0)]
GREEDY_ROOTS = [0, 0, 0, 3, 3, 5]


def test_greedy_oracle_hand_fixture():
    pk = _pack([(a, b) for a, b, _ in GREEDY_PAIRS])
    s = np.array([x for _, _, x in GREEDY_PAIRS], np.float32)
    roots = ro.greedy_cluster(6, pk, s, GREEDY_LISTS, 3)
    assert roots.tolist() == GREEDY_ROOTS
    assert ro.clusters_in_order(roots) == [[0, 1, 2], [3, 4], [5]]


def test_greedy_oracle_freeze_threshold():
    # thres 2: (0,1) merges and freezes at once, (3,4) likewise; the stale
    # pairs then meet frozen roots and nothing else merges.
    pk = _pack([(a, b) for a, b, _ in GREEDY_PAIRS])
    s = np.array([x for _, _, x in GREEDY_PAIRS], np.float32)
    roots = ro.greedy_cluster(6, pk, s, GREEDY_LISTS, 2)
    assert roots.tolist() == [0, 0, 2, 3, 3, 5]


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def _run_host_greedy(ext, n, pk, s, lists, thres):
    ptr, idx = _csr(lists)
    return ext.tca_greedy(torch.from_numpy(pk.view(np.int64)), torch.from_numpy(s),
                          torch.from_numpy(ptr), torch.from_numpy(idx), thres).numpy()


def test_host_greedy_hand_fixture(ext):
    pk = _pack([(a, b) for a, b, _ in GREEDY_PAIRS])
    s = np.array([x for _, _, x in GREEDY_PAIRS], np.float32)
    assert _run_host_greedy(ext, 6, pk, s, GREEDY_LISTS, 3).tolist() == GREEDY_ROOTS


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("thres", [2, 3, 16])
def test_host_greedy_matches_oracle(ext, seed, thres):
    """libdtc's C++ greedy (pre-sorted stream + heap of re-queued pairs)
    against the oracle's single heapq, with many tied similarities
    (quantised to quarters) so that the tie order matters."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    lists = [np.unique(rng.integers(0, 60, rng.integers(0, 12))).tolist() for _ in range(n)]
    P = int(rng.integers(1, 6 * n))
    u, v = rng.integers(0, n, P), rng.integers(0, n, P)
    keep = u != v
    pairs = sorted({(int(min(a, b)), int(max(a, b))) for a, b in zip(u[keep], v[keep])})
    rng.shuffle(pairs)
    pk = _pack(pairs)
    ptr, idx = _csr(lists)
    s = np.round(ro.jaccard_pairs(ptr, idx, pk) * 4) / 4
    s = s.astype(np.float32)
    want = ro.greedy_cluster(n, pk, s, lists, thres)
    got = _run_host_greedy(ext, n, pk, s, lists, thres)
    assert np.array_equal(want, got)


def test_host_greedy_rejects_out_of_range(ext):
    ptr, idx = _csr([[1], [2]])
    with pytest.raises(RuntimeError, match="out of range"):
        ext.tca_greedy(torch.tensor([(0 << 32) | 5], dtype=torch.int64), torch.tensor([0.5]),
                       torch.from_numpy(ptr), torch.from_numpy(idx), 16)


def test_minhash_oracle_properties():
    rng = np.random.default_rng(3)
    lists = [np.unique(rng.integers(0, 400, rng.integers(1, 80))) for _ in range(60)]
    ptr, idx = _csr(lists)
    K = 128
    sig = ro.minhash_signatures(ptr, idx, 60, K, 7).astype(np.int64)
    # entry j is the minimum of an explicit hash over the set
    seeds = ro.hash_seeds(K, 7)
    for r in (0, 17, 59):
        h = ro.fmix32(np.asarray(lists[r], np.uint64)[:, None] ^ seeds[None, :])
        assert np.array_equal(sig[r], h.min(axis=0).astype(np.int64))
    # MinHash of a union = entrywise minimum of the members' signatures
    up, ui = _csr([np.union1d(lists[0], lists[1])])
    assert np.array_equal(ro.minhash_signatures(up, ui, 1, K, 7)[0].astype(np.int64),
                          np.minimum(sig[0], sig[1]))
    # fmix32 is MurmurHash3's finaliser: known value fmix32(1) = 0x514E28B7
    assert int(ro.fmix32(np.uint64(1))) == 0x514E28B7
    # the fraction of agreeing entries estimates Jaccard (std <= 0.5/sqrt(K))
    errs = []
    for a, b in itertools.combinations(range(12), 2):
        J = len(np.intersect1d(lists[a], lists[b])) / len(np.union1d(lists[a], lists[b]))
        errs.append(abs(float((sig[a] == sig[b]).mean()) - J))
    assert np.mean(errs) < 0.05 and max(errs) < 0.25


def test_lsh_oracle_vs_brute_force():
    """With max_neighbors >= n every pair sharing a band is a candidate, and
    only those."""
    rng = np.random.default_rng(5)
    n, K, bands, r = 40, 8, 4, 2
    sig = rng.integers(0, 3, (n, K)).astype(np.uint32)  # tiny alphabet: many collisions
    live = rng.random(n) > 0.1
    got = set(ro.lsh_pairs(sig, live, bands, r, n).tolist())
    want = set()
    for a, b in itertools.combinations(range(n), 2):
        if live[a] and live[b] and any(np.array_equal(sig[a, r * t:r * t + r], sig[b, r * t:r * t + r])
                                       for t in range(bands)):
            want.add((a << 32) | b)
    assert got == want


def test_jaccard_oracle_brute_force():
    rng = np.random.default_rng(9)
    lists = [np.unique(rng.integers(0, 50, rng.integers(0, 20))) for _ in range(30)]
    ptr, idx = _csr(lists)
    pairs = _pack(list(itertools.combinations(range(30), 2)))
    s = ro.jaccard_pairs(ptr, idx, pairs)
    for p, x in zip(pairs.tolist(), s.tolist()):
        A, B = set(lists[p >> 32].tolist()), set(lists[p & 0xFFFFFFFF].tolist())
        want = 0.0 if not A or not B else len(A & B) / len(A | B)
        assert x == pytest.approx(want, rel=1e-7, abs=0)


def planted_blocks(num_blocks, block, width, seed):
    """Block-diagonal pattern whose rows inside a block share one column set
    (block b owns columns [b*width, (b+1)*width), each row taking the same
    random subset), rows shuffled."""
    rng = np.random.default_rng(seed)
    lists = []
    for b in range(num_blocks):
        cols = np.sort(rng.choice(width, size=width // 2, replace=False)) + b * width
        lists += [cols] * block
    perm = rng.permutation(len(lists))
    return [lists[i] for i in perm], perm


def test_oracle_recovers_planted_blocks():
    """Rows of identical sets (Jaccard 1 inside a block, 0 across) shuffled:
    the reorder must bring every block back into whole 16-row windows, so
    sum_w U_w returns to its minimum, num_blocks * width / 2."""
    lists, _ = planted_blocks(24, 16, 32, seed=1)
    ptr, idx = _csr(lists)
    M = len(lists)
    before = ro.sum_unique_columns(ptr, idx, M)
    order = ro.tca_reorder(ptr, idx, M)
    assert sorted(order.tolist()) == list(range(M))
    new = [lists[i] for i in order]
    p2, i2 = _csr(new)
    assert ro.sum_unique_columns(p2, i2, M) == 24 * 16 < before
