"""Whether TCA reordering pays for a matrix (reordering/TCA_reorder.py is an
offline step the reference applies to every benchmarked dataset,
scripts/DTCSpMM/run_DTC_SpMM.sh:2).

``ReorderAdvisor`` is an ensemble of small MLPs, trained with the same HIP
kernels as the variant selector, that predicts from the features of the
matrix as given (``dtc.selector.features``) the log speedup of the best SpMM
variant after the TCA row reordering:

    gain = log(min_v t_v(A) / min_v t_v(P A P^T))

The training set is a fifth of the selector's seeded matrices (every fifth
group of five consecutive seeds, so all five families), timed on the
original and on the reordered matrix (``collect``); seeds from 1600 on are
held out.  ``advise`` recommends reordering when the predicted gain
exceeds a threshold (5% by default).  The reorder's own cost (seconds on
the host for the greedy merge) is not in the gain: the reference treats it
as a one-off preprocessing step.

    python -m dtc.reorder_advice collect [--every 5] [--out ...]
    python -m dtc.reorder_advice train [--data ...] [--out ...]
    python -m dtc.reorder_advice eval
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import selector as sel

DEFAULT_DATA = os.path.join(sel.DATA_DIR, "reorder_timings.npz")
DEFAULT_MODEL = os.path.join(sel.DATA_DIR, "reorder_advisor.npz")
THRESHOLD = math.log(1.05)


def _reorder_sample(seed: int):
    """Pool worker: the matrix of `seed` (dtc.selector.make_matrix)."""
    return sel.make_matrix(seed)


def collect(count: int, out: str, start: int = 0, every: int = 5, workers: int = 8,
            iters: int = 3) -> None:
    """Times every selector variant on one group of five seeds in `every`
    (all five families) before and after the TCA reordering; writes an npz
    (rewritten every 25 matrices, so an interrupted run keeps its progress)."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")  # workers fork before this process touches the GPU
    pool = ctx.Pool(max(1, workers))
    import torch

    from .graphs import CSR
    from .matrix import DTCMatrix
    from .reorder import relabel, tca_reorder
    # groups of five consecutive seeds (one matrix of each family), every
    # `every`-th group
    seeds = [s for s in range(start, start + count) if (s // len(sel.FAMILIES)) % every == 0]
    feats, t_orig, t_reord, meta = [], [], [], []
    t0 = time.time()

    def write():
        m = np.asarray(meta, np.int64)
        np.savez_compressed(out, features=np.stack(feats), times_orig_ms=np.stack(t_orig),
                            times_reordered_ms=np.stack(t_reord), seed=m[:, 0], family=m[:, 1],
                            N=m[:, 2], rows=m[:, 3], nnz=m[:, 4], blocks=m[:, 5],
                            blocks_reordered=m[:, 6], reorder_ms=m[:, 7],
                            variants=np.array([sel.variant_name(v) for v in sel.VARIANTS]),
                            feature_names=np.array(sel.FEATURE_NAMES))

    for k, (seed, family, N, rowptr, col, vals, M, K) in enumerate(
            pool.imap(_reorder_sample, seeds, chunksize=1)):
        if M != K:
            continue  # the symmetric relabelling needs a square matrix
        A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=K)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        B = torch.randn((K, N), generator=gen, device="cuda")
        f = sel.features(A, N)
        ta = sel.time_variants(A, B, iters=iters)
        r0 = time.time()
        order, _ = tca_reorder(rowptr, col, M)
        r_s = time.time() - r0
        a2 = relabel(CSR(rowptr, col, vals, M, K), order)
        A2 = DTCMatrix(a2.rowptr, a2.col, a2.values, num_rows=M, num_cols=K)
        B2 = B[torch.from_numpy(order).to(B.device)].contiguous()  # row i of B2 = row order[i] of B
        tb = sel.time_variants(A2, B2, iters=iters)
        feats.append(f)
        t_orig.append(ta)
        t_reord.append(tb)
        meta.append((seed, sel.FAMILIES.index(family), N, M, A.nnz, A.block_count, A2.block_count,
                     int(round(r_s * 1e3))))
        del A, A2, B, B2
        if (k + 1) % 10 == 0:
            print(f"[collect-reorder] {k + 1}/{len(seeds)} matrices, {time.time() - t0:.0f}s", flush=True)
        if (k + 1) % 25 == 0:
            write()  # progress survives an interrupted run
    pool.close()
    write()
    print(f"[collect-reorder] wrote {out} ({len(feats)} matrices, {time.time() - t0:.0f}s)", flush=True)


def load_data(paths):
    """One dataset from several collect outputs (concatenated)."""
    parts = [np.load(p) for p in paths]
    keys = [k for k in parts[0].files if k not in ("variants", "feature_names")]
    out = {k: np.concatenate([z[k] for z in parts]) for k in keys}
    out["variants"], out["feature_names"] = parts[0]["variants"], parts[0]["feature_names"]
    return out


def gains(data) -> np.ndarray:
    """log(best original time / best reordered time) per matrix."""
    return (np.log(np.nanmin(data["times_orig_ms"], axis=1))
            - np.log(np.nanmin(data["times_reordered_ms"], axis=1)))


class ReorderAdvisor(sel.MLPSelector):
    """features -> 16 -> 16 -> predicted log speedup from reordering (an
    ensemble mean, as the variant selector; 320 training matrices are few,
    and a float64 replay of the training generalised best at this width and
    1,000 steps)."""

    def __init__(self, in_dim: int, hidden: int = 16, seed: int = 0, members: int = 5):
        super().__init__(in_dim, 1, hidden=hidden, seed=seed, members=members)

    @staticmethod
    def targets(gain):
        g = np.asarray(gain, np.float32).reshape(-1, 1)
        return g, np.isfinite(g).astype(np.float32)

    def predict_gain(self, X) -> np.ndarray:
        return self.predict(X)[:, 0]

    def advise(self, x, threshold: float = THRESHOLD) -> bool:
        return bool(self.predict_gain(np.atleast_2d(x))[0] > threshold)

    def save(self, path: str):
        flat = {}
        for k, params in enumerate(self.members):
            flat.update({f"m{k}_w{i}": p[0].cpu().numpy() for i, p in enumerate(params)})
            flat.update({f"m{k}_b{i}": p[1].cpu().numpy() for i, p in enumerate(params)})
        np.savez(path, mu=self.mu, sigma=self.sigma, lo=self.lo, hi=self.hi,
                 members=np.int32(len(self.members)), feature_names=np.array(sel.FEATURE_NAMES), **flat)

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL) -> "ReorderAdvisor":
        import torch
        z = np.load(path)
        if [str(s) for s in z["feature_names"]] != sel.FEATURE_NAMES:
            raise ValueError(f"{path} was trained on different features")
        obj = cls.__new__(cls)
        obj.members = [[[torch.from_numpy(z[f"m{k}_w{i}"]).cuda(), torch.from_numpy(z[f"m{k}_b{i}"]).cuda()]
                        for i in range(3)] for k in range(int(z["members"]))]
        obj.mu, obj.sigma, obj.lo, obj.hi = z["mu"], z["sigma"], z["lo"], z["hi"]
        return obj


def evaluate(model: ReorderAdvisor, data, holdout_from: int = 1600, threshold: float = THRESHOLD) -> dict:
    """Held-out quality of the advice: decision accuracy against the measured
    gain, and the mean speedup of following it against never and always
    reordering and against the oracle decision."""
    test = data["seed"] >= holdout_from
    g = gains(data)[test]
    pred = model.predict_gain(data["features"][test])
    want = g > threshold
    do = pred > threshold
    realised = np.where(do, g, 0.0)          # log speedup obtained by following the advice
    best = np.maximum(g, 0.0)                 # with perfect advice
    return {"heldout_matrices": int(test.sum()),
            "accuracy": float((do == want).mean()),
            "reorder_pays_fraction": float(want.mean()),
            "advised_fraction": float(do.mean()),
            "mean_speedup_following_advice": float(np.exp(realised.mean())),
            "mean_speedup_always_reorder": float(np.exp(g.mean())),
            "mean_speedup_perfect_advice": float(np.exp(best.mean())),
            "pred_gain_corr": float(np.corrcoef(pred, g)[0, 1]) if len(g) > 1 else float("nan")}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dtc.reorder_advice")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--count", type=int, default=2000)
    c.add_argument("--start", type=int, default=0)
    c.add_argument("--every", type=int, default=5)
    c.add_argument("--out", default=DEFAULT_DATA)
    c.add_argument("--workers", type=int, default=8)
    t = sub.add_parser("train")
    t.add_argument("--data", default=[DEFAULT_DATA], nargs="+")
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--steps", type=int, default=1000)
    t.add_argument("--holdout-from", type=int, default=1600)
    e = sub.add_parser("eval")
    e.add_argument("--data", default=[DEFAULT_DATA], nargs="+")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--holdout-from", type=int, default=1600)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        collect(a.count, a.out, start=a.start, every=a.every, workers=a.workers)
    elif a.cmd == "train":
        data = load_data(a.data)
        tr = data["seed"] < a.holdout_from
        model = ReorderAdvisor(data["features"].shape[1])
        hist = model.fit(data["features"][tr], gains(data)[tr], steps=a.steps)
        model.save(a.out)
        print(json.dumps({"loss_history": hist, **evaluate(model, data, a.holdout_from)}))
    else:
        data = load_data(a.data)
        print(json.dumps(evaluate(ReorderAdvisor.load(a.model), data, a.holdout_from)))


if __name__ == "__main__":
    sys.exit(main())
