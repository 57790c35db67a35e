"""Learned SpMM variant selector.

The reference picks its variant from a hand-written table keyed by dataset
and N (`ExecutionPlan`, scripts/DTCSpMM/run_DTC_SpMM.py:10-30).  Here a small
MLP predicts, from structural features of the matrix and N, the relative
log-time of every variant in ``VARIANTS`` (schedule, precision, chunk, tile
shape); the selector runs the variant with the smallest prediction.  Training
data are timings this library measures itself over seeded matrices of five
families (``collect``); training and inference use the HIP kernels of
csrc/selector.hip (``mlp_linear_forward``/``_backward``, ``mlp_mse``,
``adam_step``).  Exhaustive timing (``exhaustive``) is the fallback and the
yardstick: regret = t(chosen) / t(best) - 1.

    python -m dtc.selector collect --count 2000 --out timings.npz   # on a GPU
    python -m dtc.selector train --data timings.npz --out mlp.npz
    python -m dtc.selector eval --data timings.npz --model mlp.npz
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_DATA = os.path.join(DATA_DIR, "selector_timings.npz")
DEFAULT_MODEL = os.path.join(DATA_DIR, "selector_mlp.npz")

# (schedule, precision, chunk_blocks, vec, subtiles, persistent,
# cut_span_blocks; 0 = the chunk).  Every precision here meets a 1e-5
# componentwise error bound against float64.  The cut-span variants keep
# windows of up to 256 (512) blocks whole: fewer windows merged with atomic
# adds, at the price of longer work units.
VARIANTS = [
    ("windows", "fp32", 0, 4, 1, False, 0),
    ("windows", "fp32", 0, 4, 2, False, 0),
    ("windows", "fp32", 0, 2, 1, False, 0),
    ("windows", "bf16x6", 0, 4, 1, False, 0),
    ("balanced", "fp32", 32, 2, 1, False, 0),
    ("balanced", "fp32", 64, 1, 1, False, 0),
    ("balanced", "fp32", 64, 2, 1, False, 0),
    ("balanced", "fp32", 64, 4, 1, False, 0),
    ("balanced", "fp32", 128, 2, 1, False, 0),
    ("balanced", "fp32", 128, 4, 1, False, 0),
    ("balanced", "fp32", 256, 2, 1, False, 0),
    ("balanced", "fp32", 256, 4, 1, False, 0),
    ("balanced", "fp32", 512, 2, 1, False, 0),
    ("balanced", "fp32", 512, 4, 1, False, 0),
    ("balanced", "fp32", 256, 4, 2, False, 0),
    ("balanced", "fp32", 1024, 4, 2, False, 0),
    ("balanced", "fp32", 256, 4, 1, True, 0),
    ("balanced", "bf16x6", 64, 2, 1, False, 0),
    ("balanced", "bf16x6", 64, 4, 1, False, 0),
    ("balanced", "bf16x6", 256, 2, 1, False, 0),
    ("balanced", "bf16x6", 256, 4, 1, False, 0),
    ("balanced", "fp32", 64, 4, 1, False, 256),
    ("balanced", "fp32", 128, 2, 1, False, 256),
    ("balanced", "fp32", 128, 4, 1, False, 256),
    ("balanced", "fp32", 256, 4, 1, False, 512),
]


def default_chunk_blocks(num_blocks: int) -> int:
    """The library's balanced chunk when options leave it at 0: the largest
    power of two in [32, 512] at most num_blocks / 8192 (csrc/spmm.hip)."""
    c = 32
    while c < 512 and c * 2 * 8192 <= num_blocks:
        c *= 2
    return c


# Longest row window (TC blocks) up to which the default schedule is the
# unbalanced one: one wave per window beats the balanced chunks when no
# window is long (uniform, banded, block-diagonal matrices); past it, hub
# windows would serialise.  Chosen on the training split of the selector's
# timings (128-192 blocks are equivalent there).
AUTO_WINDOW_BLOCKS = 160


def default_schedule(max_window_blocks: int) -> str:
    """The schedule of ``DTCMatrix.spmm(schedule="auto")``."""
    return "windows" if int(max_window_blocks) <= AUTO_WINDOW_BLOCKS else "balanced"


def default_variant(num_blocks: int, max_window_blocks: int) -> int:
    """Index in VARIANTS of what the library runs with default options (the
    method without the selector): FP32, unbalanced with 64-column tiles when
    no window is longer than AUTO_WINDOW_BLOCKS, else balanced with 32-column
    tiles and the default chunk."""
    if default_schedule(max_window_blocks) == "windows":
        return VARIANTS.index(("windows", "fp32", 0, 4, 1, False, 0))
    return VARIANTS.index(("balanced", "fp32", default_chunk_blocks(int(num_blocks)), 2, 1, False, 0))


def max_window_blocks_of(features_row) -> int:
    """The longest window's block count from a feature vector."""
    return int(round(2.0 ** float(features_row[FEATURE_NAMES.index("log2_max_blocks_per_window")]) - 1.0))

FEATURE_NAMES = [
    "log2_rows", "log2_nnz", "log2_N", "log2_avg_deg", "deg_cv", "log2_max_deg",
    "h_deg0", "h_deg1", "h_deg2_3", "h_deg4_7", "h_deg8_15", "h_deg16_63", "h_deg64_255",
    "h_deg256p", "empty_window_frac", "log2_blocks", "tc_fill", "unique_per_nnz",
    "log2_mean_blocks_per_window", "log2_max_blocks_per_window", "blocks_per_window_cv",
    "frac_blocks_win_gt64", "frac_blocks_win_gt256", "frac_blocks_win_gt1024", "log2_B_bytes",
    "has_values",
]

# Scale features whose training range bounds where the network is trusted:
# a matrix more than RANGE_SLACK (log2 units, x1.41) outside the training
# extremes of any of them goes to exhaustive timing instead.
RANGE_CHECKED = ["log2_rows", "log2_nnz", "log2_N", "log2_avg_deg", "log2_max_deg", "log2_blocks",
                 "log2_mean_blocks_per_window", "log2_max_blocks_per_window", "log2_B_bytes"]
RANGE_SLACK = 0.5


def out_of_range(x, lo, hi) -> list:
    """Names of the range-checked features of x outside [lo - slack, hi + slack]."""
    if lo is None or hi is None:
        return []
    x = np.asarray(x, np.float64).reshape(-1)
    bad = []
    for name in RANGE_CHECKED:
        i = FEATURE_NAMES.index(name)
        if x[i] < lo[i] - RANGE_SLACK or x[i] > hi[i] + RANGE_SLACK:
            bad.append(name)
    return bad


def variant_name(v) -> str:
    s, p, c, vec, sub, pers, span = v
    return f"{s}:{p}:c{c}:v{vec}x{sub}{':persistent' if pers else ''}{f':s{span}' if span else ''}"


# ------------------------------------------------------------------ features

def features(A, N: int) -> np.ndarray:
    """Structural features of an ME-TCF matrix (``dtc.matrix.DTCMatrix``) at width N."""
    import torch
    deg = (A.rowptr[1:] - A.rowptr[:-1]).double()
    M, nnz = A.num_rows, A.nnz
    avg = float(deg.mean()) if M else 0.0
    std = float(deg.std()) if M > 1 else 0.0
    edges = torch.tensor([0, 1, 2, 4, 8, 16, 64, 256, 1 << 40], dtype=torch.float64, device=deg.device)
    hist = torch.histogram(deg.cpu(), bins=edges.cpu())[0].numpy() / max(M, 1)
    nb = (A.row_window_offset[1:] - A.row_window_offset[:-1]).double()
    W = max(A.num_windows, 1)
    T = max(A.block_count, 1)
    nonempty = nb[nb > 0]
    mean_b = float(nonempty.mean()) if nonempty.numel() else 0.0
    cv_b = float(nonempty.std() / nonempty.mean()) if nonempty.numel() > 1 else 0.0
    sum_u = A.sum_unique_columns()
    f = [
        math.log2(max(M, 1)), math.log2(max(nnz, 1)), math.log2(N), math.log2(avg + 1.0),
        std / avg if avg > 0 else 0.0, math.log2(float(deg.max()) + 1.0 if M else 1.0),
        *hist.tolist(),
        float((nb == 0).sum()) / W, math.log2(T), nnz / (T * 128.0), sum_u / max(nnz, 1),
        math.log2(mean_b + 1.0), math.log2(float(nb.max()) + 1.0), cv_b,
        float(nb[nb > 64].sum()) / T, float(nb[nb > 256].sum()) / T, float(nb[nb > 1024].sum()) / T,
        math.log2(4.0 * A.num_cols * N), 1.0 if A.values_tc is not None else 0.0,
    ]
    return np.asarray(f, np.float32)


# ---------------------------------------------------------------- timing

def time_variants(A, B, variants=VARIANTS, iters: int = 3) -> np.ndarray:
    """Mean time (ms) of every variant, HIP events on the current stream."""
    import torch
    C = torch.empty((A.num_rows, B.shape[1]), device=B.device, dtype=torch.float32)
    stream = torch.cuda.current_stream()
    out = np.full(len(variants), np.nan)
    for i, (sched, prec, chunk, vec, sub, pers, span) in enumerate(variants):
        def run():
            A.spmm(B, schedule=sched, precision=prec, chunk_blocks=chunk, vec=vec,
                   subtiles=sub, persistent=pers, out=C, cut_span_blocks=span)
        run()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(iters):
            run()
        e1.record(stream)
        e1.synchronize()
        out[i] = e0.elapsed_time(e1) / iters
    return out


def exhaustive(A, B, variants=VARIANTS, iters: int = 3) -> int:
    """Index of the fastest variant by timing all of them (the fallback)."""
    return int(np.nanargmin(time_variants(A, B, variants, iters)))


# ------------------------------------------------------- training matrices

FAMILIES = ["uniform", "rmat", "banded", "block_diagonal", "sbm"]
WIDTHS = [32, 64, 128, 256, 512]


# Seeds [LARGE_START, XL_START): power-law matrices of 2^20..2^21 rows
# (R-MAT and the block model with hubs, alternating); seeds [XL_START,
# XL_END): the block model at the size of the 232,965-row configuration
# (150K..400K rows, 200..520 entries per row, 200 hub rows of 20K..50K).
LARGE_START, XL_START, XL_END = 2000, 2120, 2140


def matrix_spec(seed: int, max_nnz: float = 3.2e7):
    """Family, rows (log-uniform 1e4..1e6), average row length (log-uniform
    4..512, capped so nnz <= max_nnz), N and value presence of training matrix
    `seed` (deterministic); the large seeds as described above."""
    rng = np.random.default_rng(10_000 + seed)
    if LARGE_START <= seed < XL_START:
        family = "rmat" if seed % 2 == 0 else "sbm"
        rows = int(round(2 ** rng.uniform(20.0, 21.0)))
        avg = min(10 ** rng.uniform(math.log10(8), math.log10(32)), max_nnz / rows)
    elif XL_START <= seed < XL_END:
        family = "sbm"
        rows = int(round(10 ** rng.uniform(math.log10(1.5e5), math.log10(4e5))))
        avg = min(10 ** rng.uniform(math.log10(200), math.log10(520)), 1.25e8 / rows)
    else:
        family = FAMILIES[seed % len(FAMILIES)]
        rows = int(round(10 ** rng.uniform(4.0, 6.0)))
        avg = 10 ** rng.uniform(math.log10(4), math.log10(512))
        avg = min(avg, max_nnz / rows)
    N = WIDTHS[int(rng.integers(len(WIDTHS)))]
    has_values = bool(rng.random() < 0.5)
    return family, rows, avg, N, has_values, int(rng.integers(1 << 30))


def make_matrix(seed: int):
    from . import graphs
    family, rows, avg, N, has_values, s = matrix_spec(seed)
    rng = np.random.default_rng(s)
    if family == "uniform":
        a = graphs.uniform_random(rows, rows, avg, seed=s)
    elif family == "rmat":
        scale = max(10, int(round(math.log2(rows))))
        a = graphs.rmat(scale, max(1, int(round(avg / 2))), seed=s, values=None)
    elif family == "banded":
        hw = int(rng.integers(1, 65))
        a = graphs.banded(rows, hw, fill=min(1.0, avg / (2 * hw + 1)), seed=s)
    elif family == "block_diagonal":
        blk = int(rng.integers(16, 257))
        a = graphs.block_diagonal(rows, blk, fill=min(1.0, avg / blk), noise=0.01, seed=s)
    elif XL_START <= seed < XL_END:
        a = graphs.sbm_hubs(rows, communities=int(rng.integers(30, 71)), avg_per_row=avg,
                            hubs=200, hub_min=20000, hub_max=50000, seed=s, values=None)
    else:
        a = graphs.sbm_hubs(rows, communities=int(rng.integers(10, 51)), avg_per_row=avg,
                            hubs=max(1, min(200, rows // 5000)), hub_min=max(1, rows // 50),
                            hub_max=max(2, rows // 20), seed=s, values=None)
    vals = rng.uniform(-1.0, 1.0, a.nnz).astype(np.float32) if has_values else None
    return seed, family, N, a.rowptr, a.col, vals, a.num_rows, a.num_cols


def collect(count: int, out: str, start: int = 0, workers: int = 8, iters: int = 3) -> None:
    """Times every variant on `count` seeded matrices; writes an npz."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")  # workers fork before this process touches the GPU
    pool = ctx.Pool(max(1, workers))
    import torch
    from .matrix import DTCMatrix
    feats, times, meta = [], [], []
    t0 = t_last = time.time()
    for k, (seed, family, N, rowptr, col, vals, M, K) in enumerate(
            pool.imap(make_matrix, range(start, start + count), chunksize=1)):
        A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=K)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        B = torch.randn((K, N), generator=gen, device="cuda")
        feats.append(features(A, N))
        times.append(time_variants(A, B, iters=iters))
        meta.append((seed, FAMILIES.index(family), N, M, A.nnz, A.block_count))
        del A, B
        if (k + 1) % 25 == 0 or time.time() - t_last > 30:
            t_last = time.time()
            print(f"[collect] {k + 1}/{count} matrices, {t_last - t0:.0f}s", flush=True)
    pool.close()
    meta = np.asarray(meta, np.int64)
    np.savez_compressed(out, features=np.stack(feats), times_ms=np.stack(times), seed=meta[:, 0],
                        family=meta[:, 1], N=meta[:, 2], rows=meta[:, 3], nnz=meta[:, 4],
                        blocks=meta[:, 5], variants=np.array([variant_name(v) for v in VARIANTS]),
                        feature_names=np.array(FEATURE_NAMES))
    print(f"[collect] wrote {out} ({count} matrices, {time.time() - t0:.0f}s)", flush=True)


# ----------------------------------------------------------------- the MLP

class MLPSelector:
    """An ensemble of `members` networks features -> 64 -> 64 -> one relative
    log-time per variant (ReLU), trained alike from different initial
    weights; the prediction is the members' mean (held-out regret 2.7% for
    five members against 3.0-3.9% for each alone, float64 replay of the
    training on the shipped timings)."""

    def __init__(self, in_dim: int, out_dim: int, hidden: int = 64, seed: int = 0, members: int = 5):
        import torch
        dims = [in_dim, hidden, hidden, out_dim]
        self.members = []
        for k in range(members):
            rng = np.random.default_rng(seed + k)
            params = []
            for i in range(3):
                lim = math.sqrt(6.0 / (dims[i] + dims[i + 1]))  # Glorot-uniform
                w = rng.uniform(-lim, lim, (dims[i], dims[i + 1])).astype(np.float32)
                b = np.zeros(dims[i + 1], np.float32)
                params.append([torch.from_numpy(w).cuda(), torch.from_numpy(b).cuda()])
            self.members.append(params)
        self.mu = np.zeros(in_dim, np.float32)
        self.sigma = np.ones(in_dim, np.float32)
        self.lo = self.hi = None  # training feature range

    @property
    def params(self):
        """The first member's [[w, b]] x 3 (the one seeded with `seed`)."""
        return self.members[0]

    # --- helpers
    def _norm(self, X):
        import torch
        Xn = (np.asarray(X, np.float32) - self.mu) / self.sigma
        return torch.from_numpy(np.ascontiguousarray(Xn, np.float32)).cuda()

    @staticmethod
    def _forward(params, Xd):
        from . import load_extension
        ext = load_extension()
        (w1, b1), (w2, b2), (w3, b3) = params
        h1 = ext.mlp_linear_forward(Xd, w1, b1, relu=True)
        h2 = ext.mlp_linear_forward(h1, w2, b2, relu=True)
        y = ext.mlp_linear_forward(h2, w3, b3, relu=False)
        return h1, h2, y

    @staticmethod
    def targets(times_ms):
        """Relative log-times: log t - mean over the sample's timed variants."""
        t = np.asarray(times_ms, np.float64)
        mask = np.isfinite(t) & (t > 0)
        lt = np.where(mask, np.log(np.where(mask, t, 1.0)), 0.0)
        mean = (lt * mask).sum(1, keepdims=True) / np.maximum(mask.sum(1, keepdims=True), 1)
        return ((lt - mean) * mask).astype(np.float32), mask.astype(np.float32)

    def fit(self, X, times_ms, steps: int = 3000, lr: float = 3e-3, log_every: int = 500):
        import torch
        from . import load_extension
        ext = load_extension()
        X = np.asarray(X, np.float32)
        self.mu = X.mean(0)
        self.sigma = X.std(0) + 1e-6
        self.lo, self.hi = X.min(0), X.max(0)
        Xd = self._norm(X)
        tgt, mask = self.targets(times_ms)
        Td = torch.from_numpy(tgt).cuda()
        Md = torch.from_numpy(mask).cuda()
        inv = 1.0 / max(1.0, float(mask.sum()))
        logged = {}
        for params in self.members:
            state = [[torch.zeros_like(p), torch.zeros_like(p)] for pair in params for p in pair]
            (w1, b1), (w2, b2), (w3, b3) = params
            for step in range(1, steps + 1):
                h1, h2, y = self._forward(params, Xd)
                dy, loss = ext.mlp_mse(y, Td, Md, inv)
                dh2, dw3, db3 = ext.mlp_linear_backward(h2, w3, y, dy, relu=False)
                dh1, dw2, db2 = ext.mlp_linear_backward(h1, w2, h2, dh2, relu=True)
                _, dw1, db1 = ext.mlp_linear_backward(Xd, w1, h1, dh1, relu=True, need_dx=False)
                grads = [dw1, db1, dw2, db2, dw3, db3]
                flat = [p for pair in params for p in pair]
                for p, g, (m, v) in zip(flat, grads, state):
                    ext.adam_step(p, g, m, v, step, lr)
                if step % log_every == 0 or step == 1:
                    logged.setdefault(step, []).append(float(loss.item()))
        # mean training loss over the members at each logged step
        return [(k, float(np.mean(v))) for k, v in sorted(logged.items())]

    def predict(self, X) -> np.ndarray:
        Xd = self._norm(np.atleast_2d(X))
        return np.mean([self._forward(p, Xd)[2].cpu().numpy() for p in self.members], axis=0)

    def choose(self, X) -> np.ndarray:
        return np.argmin(self.predict(X), axis=1)

    def top_k(self, X, k: int) -> np.ndarray:
        """Per row of X, the k variants with the lowest predicted time, best first."""
        return np.argsort(self.predict(X), axis=1, kind="stable")[:, :k]

    def out_of_range(self, x) -> list:
        return out_of_range(x, self.lo, self.hi)

    def save(self, path: str):
        flat = {}
        for k, params in enumerate(self.members):
            flat.update({f"m{k}_w{i}": p[0].cpu().numpy() for i, p in enumerate(params)})
            flat.update({f"m{k}_b{i}": p[1].cpu().numpy() for i, p in enumerate(params)})
        extra = {} if self.lo is None else {"lo": self.lo, "hi": self.hi}
        np.savez(path, mu=self.mu, sigma=self.sigma, variants=np.array([variant_name(v) for v in VARIANTS]),
                 members=np.int32(len(self.members)), **extra, **flat)

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL) -> "MLPSelector":
        import torch
        z = np.load(path)
        names = [str(s) for s in z["variants"]]
        if names != [variant_name(v) for v in VARIANTS]:
            raise ValueError(f"{path} was trained for a different variant list")
        obj = cls.__new__(cls)
        obj.members = [[[torch.from_numpy(z[f"m{k}_w{i}"]).cuda(), torch.from_numpy(z[f"m{k}_b{i}"]).cuda()]
                        for i in range(3)] for k in range(int(z["members"]))]
        obj.mu, obj.sigma = z["mu"], z["sigma"]
        obj.lo = z["lo"] if "lo" in z else None
        obj.hi = z["hi"] if "hi" in z else None
        return obj


# Variants the selector times on the matrix itself: the network's TOP_K best
# predictions, the fastest of which is kept (model-guided search).  On the
# held-out matrices the mean regret is 2.6% for the network's first choice
# alone, 0.87% with two timed, 0.35% with three (the exact best for 91%),
# at the cost of timing three variants once per matrix.
TOP_K = 3


def pick_variant(A, B, model: MLPSelector | None = None, top_k: int = TOP_K):
    """(variant index, how): when A lies inside the network's training range,
    the fastest of its top_k predicted variants timed on A and B (top_k = 1:
    its first choice, untimed); else the fastest variant by exhaustive timing."""
    N = int(B.shape[1])
    x = features(A, N)
    bad = ["no model"] if model is None else model.out_of_range(x)
    if not bad:
        cand = [int(i) for i in model.top_k(x[None, :], max(1, top_k))[0]]
        if len(cand) == 1:
            return cand[0], "selector"
        t = time_variants(A, B, [VARIANTS[i] for i in cand])
        return cand[int(np.nanargmin(t))], f"selector (top {len(cand)} timed)"
    t = time_variants(A, B)
    return int(np.nanargmin(t)), "exhaustive (" + ", ".join(bad) + " outside the training range)"


def regret(times_ms, chosen) -> np.ndarray:
    t = np.asarray(times_ms, np.float64)
    best = np.nanmin(t, axis=1)
    return t[np.arange(t.shape[0]), chosen] / best - 1.0


def is_train_seed(seed, holdout_from: int = 1600):
    """Training split: the seeds below holdout_from, and three in four of the
    large seeds (LARGE_START on; every fourth is held out)."""
    seed = np.asarray(seed)
    return (seed < holdout_from) | ((seed >= LARGE_START) & (seed % 4 != 3))


def split(data, holdout_from: int = 1600):
    train = is_train_seed(data["seed"], holdout_from)
    return train, ~train


def best_fixed_variant(times_ms) -> int:
    """The single variant with the lowest geometric-mean slowdown over a set."""
    t = np.asarray(times_ms, np.float64)
    slow = t / np.nanmin(t, axis=1, keepdims=True)
    return int(np.nanargmin(np.nanmean(np.log(slow), axis=0)))


def evaluate(model: MLPSelector, data, holdout_from: int = 1600) -> dict:
    """Held-out regret of the selector, of the library's default options
    (their variant depends on the block count) and of the best single
    variant chosen on the training split (the strongest fixed method)."""
    train, test = split(data, holdout_from)
    X, t = data["features"][test], data["times_ms"][test]
    chosen = model.choose(X)
    fixed = best_fixed_variant(data["times_ms"][train])
    dflt = np.array([default_variant(b, max_window_blocks_of(x))
                     for b, x in zip(data["blocks"][test], X)])
    r_sel = regret(t, chosen)
    top = model.top_k(X, TOP_K)
    timed_pick = np.array([row[np.nanargmin(ti[row])] for row, ti in zip(top, t)])
    r_top = regret(t, timed_pick)
    r_def = regret(t, dflt)
    r_fix = regret(t, np.full(t.shape[0], fixed))
    picked = t[np.arange(t.shape[0]), chosen]
    fallback = np.array([bool(model.out_of_range(x)) for x in X])
    r_guard = np.where(fallback, 0.0, r_sel)  # exhaustive timing has no regret
    return {"heldout_matrices": int(test.sum()),
            "fallback_to_exhaustive": int(fallback.sum()),
            "guarded_mean_regret": float(r_guard.mean()),
            "selector_mean_regret": float(r_sel.mean()), "selector_p90_regret": float(np.quantile(r_sel, 0.9)),
            f"top{TOP_K}_timed_mean_regret": float(r_top.mean()),
            f"top{TOP_K}_timed_p90_regret": float(np.quantile(r_top, 0.9)),
            "default_variant": (f"windows:fp32:c0:v4x1 if the longest window has <= {AUTO_WINDOW_BLOCKS} "
                                "blocks, else balanced:fp32:c<T/8192 in 32..512>:v2x1"),
            "default_mean_regret": float(r_def.mean()), "default_p90_regret": float(np.quantile(r_def, 0.9)),
            "best_fixed_variant": variant_name(VARIANTS[fixed]),
            "best_fixed_mean_regret": float(r_fix.mean()), "best_fixed_p90_regret": float(np.quantile(r_fix, 0.9)),
            "mean_speedup_over_default": float(np.exp(np.log(t[np.arange(t.shape[0]), dflt] / picked).mean())),
            "mean_speedup_over_best_fixed": float(np.exp(np.log(t[:, fixed] / picked).mean())),
            "exact_pick_rate": float((r_sel < 1e-9).mean())}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dtc.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--count", type=int, default=2000)
    c.add_argument("--start", type=int, default=0)
    c.add_argument("--out", default=DEFAULT_DATA)
    c.add_argument("--workers", type=int, default=8)
    t = sub.add_parser("train")
    t.add_argument("--data", default=DEFAULT_DATA)
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--holdout-from", type=int, default=1600)
    e = sub.add_parser("eval")
    e.add_argument("--data", default=DEFAULT_DATA)
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--holdout-from", type=int, default=1600)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        collect(a.count, a.out, start=a.start, workers=a.workers)
    elif a.cmd == "train":
        data = np.load(a.data)
        tr, _ = split(data, a.holdout_from)
        model = MLPSelector(data["features"].shape[1], data["times_ms"].shape[1])
        hist = model.fit(data["features"][tr], data["times_ms"][tr], steps=a.steps)
        model.save(a.out)
        print(json.dumps({"loss_history": hist, **evaluate(model, data, a.holdout_from)}))
    else:
        data = np.load(a.data)
        print(json.dumps(evaluate(MLPSelector.load(a.model), data, a.holdout_from)))


if __name__ == "__main__":
    sys.exit(main())
