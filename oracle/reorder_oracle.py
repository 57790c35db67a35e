"""CPU oracle for the TCA row reordering (test infrastructure only: imported by
tests, never by the product path).

The reference's reordering (reordering/TCA_reorder.py) is a two-level greedy
clustering of rows:
  1. TCU-aware level: MinHash/LSH (threshold 0.2, 128 permutations,
     :22-24, :55-80) proposes row pairs, their Jaccard similarity
     (cugraph.jaccard, :122-146) fills a priority queue, and rows are merged
     greedily, most similar pair first, into clusters frozen once they reach
     `thres` rows (:152-187).
  2. Cache-aware level: the same over the clusters, each represented by the
     union of its rows' columns, frozen at `c_thres` clusters (:203-286).
  3. Output: rows cluster by cluster, clusters super-cluster by super-cluster
     (:300-315), as `reorder_id` (original ids in their new order).

The reference's hashing (datasketch MinHash over SHA-1 of decimal strings,
libMHCUDA weighted MinHash) and Jaccard (cugraph) come from libraries absent
here, so the signatures and LSH candidate sets are this implementation's own
and cannot be compared with the reference's; what is restated line by line is
the greedy clustering and the output order, given the candidate pairs.
Signature family: h_j(c) = fmix32(c XOR s_j), s_j = fmix32(j * 0x9E3779B9 +
seed), fmix32 being MurmurHash3's 32-bit finaliser; a signature entry is the
minimum of h_j over the row's columns.  LSH: b bands of r entries; two rows
are candidates when some band agrees and, in that band's stable order of rows
sorted by band key, they lie at most D positions apart.
"""
from __future__ import annotations

import heapq

import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def fmix32(h):
    h = np.asarray(h).astype(np.uint64) & M32
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & M32
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & M32
    h ^= h >> np.uint64(16)
    return h


def fmix64(k):
    k = np.asarray(k).astype(np.uint64)
    k ^= k >> np.uint64(33)
    k *= np.uint64(0xFF51AFD7ED558CCD)
    k ^= k >> np.uint64(33)
    k *= np.uint64(0xC4CEB9FE1A85EC53)
    k ^= k >> np.uint64(33)
    return k


def hash_seeds(num_hashes, seed):
    j = np.arange(num_hashes, dtype=np.uint64)
    return fmix32((j * np.uint64(0x9E3779B9) + np.uint64(seed)) & M32)


def minhash_signatures(rowptr, col, num_rows, num_hashes, seed):
    """uint32 [num_rows, num_hashes]; empty rows hold 0xFFFFFFFF."""
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64).astype(np.uint64)
    rows = np.repeat(np.arange(num_rows), np.diff(rowptr))
    sig = np.full((num_rows, num_hashes), 0xFFFFFFFF, np.uint64)
    seeds = hash_seeds(num_hashes, seed)
    for j in range(num_hashes):
        np.minimum.at(sig[:, j], rows, fmix32(col ^ seeds[j]))
    return sig.astype(np.uint32)


def band_keys(sig, band, rows_per_band):
    """64-bit key of one band: the two entries side by side when r = 2, the
    entry itself when r = 1, else an fmix64 chain over the r entries."""
    s = np.asarray(sig, np.uint32).astype(np.uint64)
    c0 = band * rows_per_band
    if rows_per_band == 1:
        return s[:, c0]
    if rows_per_band == 2:
        return (s[:, c0] << np.uint64(32)) | s[:, c0 + 1]
    k = np.zeros(s.shape[0], np.uint64)
    for t in range(rows_per_band):
        k = fmix64(k ^ (s[:, c0 + t] + np.uint64(0x9E3779B97F4A7C15 & 0xFFFFFFFFFFFFFFFF)))
    return k


def lsh_pairs(sig, live, bands, rows_per_band, max_neighbors):
    """Sorted unique candidate pairs (u < v) as packed uint64 u << 32 | v."""
    n = sig.shape[0]
    live = np.asarray(live, bool)
    out = []
    for b in range(bands):
        key = band_keys(sig, b, rows_per_band)
        order = np.argsort(key, kind="stable")
        ks = key[order]
        for d in range(1, max_neighbors + 1):
            if d >= n:
                break
            u, v = order[:-d], order[d:]
            m = (ks[:-d] == ks[d:]) & live[u] & live[v]
            a, c = np.minimum(u[m], v[m]), np.maximum(u[m], v[m])
            out.append((a.astype(np.uint64) << np.uint64(32)) | c.astype(np.uint64))
    if not out:
        return np.zeros(0, np.uint64)
    return np.unique(np.concatenate(out))


def jaccard_pairs(rowptr, col, pairs):
    """float32 |S_u ∩ S_v| / |S_u ∪ S_v| over sorted duplicate-free rows."""
    rowptr = np.asarray(rowptr, np.int64)
    out = np.zeros(pairs.shape[0], np.float32)
    for i, p in enumerate(pairs.tolist()):
        u, v = p >> 32, p & 0xFFFFFFFF
        a = col[rowptr[u]:rowptr[u + 1]]
        b = col[rowptr[v]:rowptr[v + 1]]
        if a.shape[0] == 0 or b.shape[0] == 0:
            continue
        inter = np.intersect1d(a, b, assume_unique=True).shape[0]
        out[i] = np.float32(inter) / np.float32(a.shape[0] + b.shape[0] - inter)
    return out


def _jd(l1, l2):
    # TCA_reorder.py:95-100
    if len(l1) == 0 or len(l2) == 0:
        return 0.0
    s1, s2 = set(l1), set(l2)
    return float(len(s1 & s2)) / len(s1 | s2)


def greedy_cluster(n, pairs, sims, lists, thres):
    """Root of every item after the greedy merge (TCA_reorder.py:152-187 for
    rows, :250-286 for clusters).  The queue pops the most similar pair first
    and, among equal similarities, the smaller (p1, p2) first (a total order;
    the reference's PriorityQueue breaks such ties by its heap's internal
    arrangement).  Of repeated pairs the first given is kept, pairs (a, a)
    are skipped (:139); `lists[i]` are the column sets used when a stale pair
    is re-queued with the Jaccard of the two roots' own lists (:182-187)."""
    que = []
    sset = set()

    def makenum(a, b):
        return min(a, b) * n + max(a, b)

    for p, s in zip(pairs.tolist(), sims.tolist()):
        a, b = p >> 32, p & 0xFFFFFFFF
        if a == b or makenum(a, b) in sset:
            continue
        heapq.heappush(que, (-float(s), a, b))
        sset.add(makenum(a, b))
    cid = list(range(n))
    csz = [1] * n
    deleted = [0] * n
    num_cluster = n

    def root(i):
        while i != cid[i]:
            cid[i] = cid[cid[i]]
            i = cid[i]
        return i

    while que and num_cluster > 0:
        _, p1, p2 = heapq.heappop(que)
        sset.remove(makenum(p1, p2))
        if p1 == cid[p1] and p2 == cid[p2]:
            if deleted[p1] or deleted[p2]:
                continue
            if csz[p1] < csz[p2]:
                cid[p1] = p2
                num_cluster -= 1
                csz[p2] += csz[p1]
                if csz[p2] >= thres:
                    deleted[p2] = 1
                    num_cluster -= 1
            else:
                cid[p2] = p1
                num_cluster -= 1
                csz[p1] += csz[p2]
                if csz[p1] >= thres:
                    deleted[p1] = 1
                    num_cluster -= 1
        else:
            p1, p2 = root(p1), root(p2)
            if deleted[p1] or deleted[p2]:
                continue
            if p1 != p2 and makenum(p1, p2) not in sset:
                heapq.heappush(que, (-_jd(lists[p1], lists[p2]), p1, p2))
                sset.add(makenum(p1, p2))
    return np.array([root(i) for i in range(n)], np.int64)


def clusters_in_order(roots):
    """Clusters keyed by root in first-occurrence order, members ascending
    (TCA_reorder.py:189-196)."""
    clusters = {}
    for i, r in enumerate(roots.tolist()):
        clusters.setdefault(r, []).append(i)
    return list(clusters.values())


def tca_reorder(rowptr, col, num_rows, thres=16, cluster_thres=128, bands=28,
                rows_per_band=2, max_neighbors=8, seed=2022):
    """reorder_id (int64 [num_rows]): original row ids in their new order."""
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64)
    K = bands * rows_per_band
    deg = np.diff(rowptr)
    sig = minhash_signatures(rowptr, col, num_rows, K, seed)
    pairs = lsh_pairs(sig, deg > 0, bands, rows_per_band, max_neighbors)
    sims = jaccard_pairs(rowptr, col, pairs)
    lists = [col[rowptr[i]:rowptr[i + 1]].tolist() for i in range(num_rows)]
    clusters = clusters_in_order(greedy_cluster(num_rows, pairs, sims, lists, thres))
    C = len(clusters)
    # cluster unions (TCA_reorder.py:222-231); MinHash of a union is the
    # entrywise minimum of its members' signatures
    uptr = np.zeros(C + 1, np.int64)
    ucol = []
    sig_c = np.empty((C, K), np.uint32)
    for k, rows in enumerate(clusters):
        u = np.unique(np.concatenate([col[rowptr[r]:rowptr[r + 1]] for r in rows]))
        ucol.append(u)
        uptr[k + 1] = uptr[k] + u.shape[0]
        sig_c[k] = sig[rows].min(axis=0)
    ucol = np.concatenate(ucol) if ucol else np.zeros(0, np.int64)
    pairs_c = lsh_pairs(sig_c, np.diff(uptr) > 0, bands, rows_per_band, max_neighbors)
    sims_c = jaccard_pairs(uptr, ucol, pairs_c)
    lists_c = [ucol[uptr[k]:uptr[k + 1]].tolist() for k in range(C)]
    supers = clusters_in_order(greedy_cluster(C, pairs_c, sims_c, lists_c, cluster_thres))
    # TCA_reorder.py:300-305
    return np.array([r for s in supers for k in s for r in clusters[k]], np.int64)


def sum_unique_columns(rowptr, col, num_rows, blk_h=16):
    """sum_w U_w: distinct columns per row window."""
    rows = np.repeat(np.arange(num_rows, dtype=np.int64), np.diff(np.asarray(rowptr, np.int64)))
    col = np.asarray(col, np.int64)
    if col.shape[0] == 0:
        return 0
    key = (rows // blk_h) * (int(col.max()) + 1) + col
    return int(np.unique(key).shape[0])
