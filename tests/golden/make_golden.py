"""Writes tests/golden/c1_metcf_digest.json: SHA-256 digests of the CPU
oracle's ME-TCF arrays for the 4096 x 4096, 16-per-row uniform matrix
(PCG64 seed 0).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "dtc-spmm_asplos24_amd")]

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from dtc.graphs import uniform_rows  # noqa: E402

KEYS = ["RowWindowOffset", "TCblockRowid", "TCblocktileId", "TCblockoffset", "SparseAToXindex",
        "edgeToColumn", "edgeToRow", "blockPartition"]


def main():
    a = uniform_rows(4096, 4096, 16, seed=0)
    m = oracle.metcf(a.rowptr, a.col, a.num_rows)
    out = {
        "matrix": "uniform_rows(4096, 4096, per_row=16, seed=0), ME-TCF 16x8",
        "nnz": a.nnz,
        "block_count": int(m["block_count"]),
        "sha256": {k: hashlib.sha256(np.ascontiguousarray(m[k]).tobytes()).hexdigest() for k in KEYS},
    }
    with open(os.path.join(HERE, "c1_metcf_digest.json"), "w") as f:
        json.dump(out, f, indent=2)
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
