"""Shared test helpers: fixture loading, random CSR cases and an independent
brute-force restatement of ME-TCF used to pin the C oracle."""
from __future__ import annotations

import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture_18():
    with open(os.path.join(GOLDEN, "metcf_fixture_18.json")) as f:
        fx = json.load(f)
    M = fx["num_nodes"]
    rows = {int(k): v for k, v in fx["rows"].items()}
    rowptr = [0]
    col = []
    for r in range(M):
        col.extend(rows.get(r, []))
        rowptr.append(len(col))
    return np.array(rowptr, np.int32), np.array(col, np.int32), M, fx


def random_csr(rng, M, K, avg, sorted_rows=True, empty_frac=0.0, dup_frac=0.0):
    """Random CSR with optional unsorted rows, empty rows and duplicates."""
    rowptr = [0]
    col = []
    for r in range(M):
        n = 0 if rng.random() < empty_frac else int(rng.poisson(avg))
        n = min(n, K)
        c = list(rng.choice(K, n, replace=False)) if n else []
        if c and dup_frac > 0 and rng.random() < dup_frac:
            c.append(c[rng.integers(len(c))])
        if sorted_rows:
            c.sort()
        else:
            rng.shuffle(c)
        col.extend(int(x) for x in c)
        rowptr.append(len(col))
    return np.array(rowptr, np.int32), np.array(col, np.int32)


def naive_metcf(rowptr, col, M, bh=16, bw=8, values=None):
    """ME-TCF straight from its definition with Python sets and lists.

    S_w = sorted set of the window's columns; block k of the window holds the
    columns S_w[k*bw:(k+1)*bw]; its stored elements are the window's edges
    whose column falls in that slice, kept in CSR order; tile id =
    (row % bh) * bw + slot, as uint8.
    """
    W = math.ceil(M / bh) if M else 0
    rowptr = [int(x) for x in rowptr]
    col = [int(x) for x in col]
    bp = [0] * W
    e2r = [0] * len(col)
    e2c = [0] * len(col)
    for r in range(M):
        for e in range(rowptr[r], rowptr[r + 1]):
            e2r[e] = r
    rwo = [0]
    rowid, off, tile, atox, tcv, tce = [], [0], [], [], [], []
    for w in range(W):
        s, e = rowptr[w * bh], rowptr[min(w * bh + bh, M)]
        edges = list(range(s, e))
        S = sorted(set(col[x] for x in edges))
        rank = {c: i for i, c in enumerate(S)}
        nb = -(-len(S) // bw) if edges else 0
        bp[w] = nb
        for x in edges:
            e2c[x] = rank[col[x]]
        for k in range(nb):
            rowid.append(w)
            members = [x for x in edges if rank[col[x]] // bw == k]
            for x in members:
                tile.append(((e2r[x] % bh) * bw + rank[col[x]] % bw) % 256)
                tcv.append(1.0 if values is None else float(values[x]))
                tce.append(x)
            off.append(off[-1] + len(members))
            sl = S[k * bw:(k + 1) * bw]
            atox.extend(sl + [0] * (bw - len(sl)))
        rwo.append(rwo[-1] + nb)
    return dict(blockPartition=np.array(bp, np.int32), RowWindowOffset=np.array(rwo, np.int32),
                TCblockRowid=np.array(rowid, np.int32), TCblockoffset=np.array(off, np.int32),
                TCblocktileId=np.array(tile, np.uint8), SparseAToXindex=np.array(atox, np.int32),
                edgeToRow=np.array(e2r, np.int32), edgeToColumn=np.array(e2c, np.int32),
                block_count=len(rowid), TCblockValues=np.array(tcv, np.float32),
                TCblockEdge=np.array(tce, np.int32))


METCF_KEYS = ["blockPartition", "RowWindowOffset", "TCblockRowid", "TCblockoffset", "TCblocktileId",
              "SparseAToXindex", "edgeToRow", "edgeToColumn"]


def assert_metcf_equal(a, b, keys=METCF_KEYS, ctx=""):
    assert int(a["block_count"]) == int(b["block_count"]), f"{ctx} block_count"
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, f"{ctx} {k} shape {x.shape} vs {y.shape}"
        if not np.array_equal(x, y):
            bad = np.nonzero(x != y)[0][:10]
            raise AssertionError(f"{ctx} {k} differs at {bad}: {x[bad]} vs {y[bad]}")


def decode_metcf(m, M, bh=16, bw=8):
    """(row, col) pairs stored in an ME-TCF dict, in TC-block order."""
    rwo, off = m["RowWindowOffset"], m["TCblockoffset"]
    tile, atox = m["TCblocktileId"], m["SparseAToXindex"]
    out = []
    for w in range(len(rwo) - 1):
        for b in range(rwo[w], rwo[w + 1]):
            for p in range(off[b], off[b + 1]):
                t = int(tile[p])
                out.append((w * bh + t // bw, int(atox[b * bw + t % bw])))
    return out


# ----------------------------------------------------------------- GPU side

def say(*args):
    """Names the case before a GPU step (flushed, so a crash still shows it)."""
    print("[case]", *args, flush=True)


def gpu_metcf(ext, rowptr, col, M, bh=16, bw=8, values=None, block_partition=None,
              edge_order=True):
    import torch
    dev = torch.device("cuda")
    W = -(-M // bh) if M else 0
    colt = torch.from_numpy(np.ascontiguousarray(col, np.int32)).to(dev)
    rpt = torch.from_numpy(np.ascontiguousarray(rowptr, np.int32)).to(dev)
    bp = (torch.zeros(W, dtype=torch.int32, device=dev) if block_partition is None
          else torch.from_numpy(np.ascontiguousarray(block_partition, np.int32)).to(dev))
    e2c = torch.zeros(col.shape[0], dtype=torch.int32, device=dev)
    e2r = torch.zeros(col.shape[0], dtype=torch.int32, device=dev)
    v = None if values is None else torch.from_numpy(np.ascontiguousarray(values, np.float32)).to(dev)
    out = ext.preprocess_gpu(colt, rpt, M, bh, bw, bp, e2c, e2r, values=v, return_edge_order=edge_order)
    names = ["RowWindowOffset", "TCblockRowid", "TCblocktileId", "TCblockoffset", "SparseAToXindex"]
    m = {k: out[i].cpu().numpy() for i, k in enumerate(names)}
    m["block_count"] = int(out[5])
    i = 6
    if values is not None:
        m["TCblockValues"] = out[i].cpu().numpy()
        i += 1
    if edge_order:
        m["TCblockEdge"] = out[i].cpu().numpy()
    m["blockPartition"] = bp.cpu().numpy()
    m["edgeToColumn"] = e2c.cpu().numpy()
    m["edgeToRow"] = e2r.cpu().numpy()
    m["_dev"] = out
    return m
