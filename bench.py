"""DTC-SpMM benchmark on one MI355X.

Workload (default): the power-law configuration -- R-MAT (a, b, c, d) =
(0.57, 0.19, 0.19, 0.05), scale 20 (1,048,576 nodes), edge factor 16,
symmetrised, self loops and duplicates removed, values 1/sqrt(deg_i deg_j)
float32; B = 1,048,576 x 256 float32 N(0, 1).

A step is one SpMM launch C = A * B over the resident ME-TCF (built once on
the GPU and timed separately, as the reference times preprocess and SpMM
apart: DTCSpMM.cpp:337-348 vs DTCSpMM_kernel.cu:841-856).  The variant is the
learned selector's pick (dtc/selector.py, weights in dtc/data/); every
variant is also timed (exhaustive timing: the selector's yardstick and fallback)
and its componentwise error against the float64 CSR product is recorded.
The plain-FP32 path is the fastest FP32 variant of that table.  A matrix
outside the selector's training range goes to exhaustive timing instead.
For the R-MAT workload the chosen variant is also timed on the TCA-reordered
graph (field "reordered"; --no-reorder skips it); the headline value stays
on the graph as generated.

Prints one JSON line.  --gpus N > 1 exits: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "dtc-spmm_asplos24_amd")
sys.path[:0] = [PKG]

ERR_TARGET = 1e-5
HBM_PEAK_GBS = 8000.0
EXTRA_PRECISIONS = ["bf16x3", "fp16", "bf16"]  # timed and reported, fail the error gate
# Chunks below the selector's range (32-column tiles): many more, shorter
# waves, which pays on small matrices such as the 4096-row uniform case.
EXTRA_SMALL_CHUNKS = [16, 8]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed(stream, fn, iters):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(iters):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def timed_steps(stream, fn, iters):
    """(mean, median) ms per step: one event before every step and one after
    the last, all on the launch stream; the mean is first to last event."""
    import torch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(iters + 1)]
    ev[0].record(stream)
    for i in range(iters):
        fn()
        ev[i + 1].record(stream)
    ev[-1].synchronize()
    steps = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(iters))
    mid = len(steps) // 2
    median = steps[mid] if len(steps) % 2 else 0.5 * (steps[mid - 1] + steps[mid])
    return ev[0].elapsed_time(ev[-1]) / iters, median


def copy_probe():
    """Three copies of a 1 GiB float32 tensor: each reads and writes exactly
    2**30 bytes, the reference point for FETCH_SIZE / WRITE_SIZE."""
    import torch
    src = torch.ones(1 << 28, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 1 << 30, "launches": 3}), flush=True)


def bench_gcn(args):
    """GCN layer H = A (X W): R-MAT scale 18 (262,144 nodes, edge factor 16,
    symmetric normalisation), X 262,144 x 256 and W 256 x 256 (Xavier) in
    bf16, float32 accumulation.  Times one layer's forward and forward +
    backward, a 2-layer model A (relu(A (X W1)) W2) forward + backward, and the
    SpMM alone, against the same computations with float32 X, W and B."""
    import numpy as np
    import torch

    from dtc.gcn import gcn_layer, gcn_two_layer, transform
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix, spmm_autograd
    from dtc.metrics import spmm_flops

    scale = 18 if args.scale == 20 else args.scale
    F = args.n
    a = rmat(scale, 16, seed=1)
    M = a.num_rows
    A = DTCMatrix(a.rowptr, a.col, None, num_rows=M, num_cols=M)
    A.transpose()
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(0)
    lim = np.sqrt(6.0 / (2 * F))
    X32 = torch.from_numpy(rng.standard_normal((M, F)).astype(np.float32)).cuda()
    W32 = torch.from_numpy(rng.uniform(-lim, lim, (F, F)).astype(np.float32)).cuda()
    Xb, Wb = X32.to(torch.bfloat16), W32.to(torch.bfloat16)
    v = torch.from_numpy(a.values).cuda()
    dH = torch.randn((M, F), device="cuda")

    def leaf(t):
        return t.detach().clone().requires_grad_(True)

    vb, Xbl, Wbl = leaf(v), leaf(Xb), leaf(Wb)
    v32, X32l, W32l = leaf(v), leaf(X32), leaf(W32)
    W2_32 = torch.from_numpy(rng.uniform(-lim, lim, (F, F)).astype(np.float32)).cuda()
    W2bl, W2_32l = leaf(W2_32.to(torch.bfloat16)), leaf(W2_32)

    def two_layer_bf16():  # the second transform applies the ReLU and the bf16 cast
        return gcn_two_layer(vb, Xbl, Wbl, W2bl, A)

    def two_layer_fp32():
        H1 = torch.relu(spmm_autograd(v32, X32l @ W32l, A))
        return spmm_autograd(v32, H1 @ W2_32l, A)

    runs = {
        "bf16_forward": lambda: gcn_layer(vb, Xbl, Wbl, A),
        "bf16_forward_backward": lambda: gcn_layer(vb, Xbl, Wbl, A).backward(dH),
        "fp32_forward": lambda: spmm_autograd(v32, X32l @ W32l, A),
        "fp32_forward_backward": lambda: spmm_autograd(v32, X32l @ W32l, A).backward(dH),
        "bf16_2layer_forward_backward": lambda: two_layer_bf16().backward(dH),
        "fp32_2layer_forward_backward": lambda: two_layer_fp32().backward(dH),
    }
    vals_tc = A.values_tc if A.values_tc is not None else None
    XWb = (Xb @ Wb).contiguous()
    XW32 = (X32 @ W32).contiguous()
    from dtc import load_extension
    vtc = load_extension().gather(v, A.tc_edge)
    runs["spmm_bf16_B"] = lambda: A.spmm(XWb, values_tc=vtc)
    runs["spmm_fp32_B"] = lambda: A.spmm(XW32, values_tc=vtc)
    # the dense half: libdtc's GCN transform against the torch path it
    # replaces (bf16 GEMM; ReLU, cast and GEMM of a float32 input)
    H1f = torch.randn((M, F), device="cuda")
    ext_ = load_extension()
    runs["transform_bf16_x"] = lambda: ext_.gcn_transform(Xb, ext_.gcn_pack_weights(Wb), F, False)
    runs["torch_matmul_bf16_x"] = lambda: torch.matmul(Xb, Wb)
    runs["transform_relu_f32_x"] = lambda: transform(H1f, Wb, relu=True)
    runs["torch_relu_cast_matmul_f32_x"] = lambda: torch.matmul(torch.relu(H1f).to(torch.bfloat16), Wb)
    del vals_tc
    out = {}
    for name, fn in runs.items():
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        out[name] = round(timed(stream, fn, args.steps), 4)
        log(f"  {name:24s} {out[name]:8.3f} ms")
    flops = spmm_flops(a.nnz, F)
    print(json.dumps({
        "metric": "GCN layer time (ms), A (X W) with bf16 X, W vs float32",
        "workload": f"R-MAT scale {scale}, edge factor 16, symmetric normalisation; X {M}x{F}, W {F}x{F}",
        "nnz": a.nnz, "ms": out,
        "spmm_gflops": {"bf16_B": round(flops / out["spmm_bf16_B"] / 1e6, 1),
                        "fp32_B": round(flops / out["spmm_fp32_B"] / 1e6, 1)},
        "speedup_bf16_layer_fwd_bwd": round(out["fp32_forward_backward"] / out["bf16_forward_backward"], 3),
        "speedup_bf16_2layer_fwd_bwd": round(out["fp32_2layer_forward_backward"]
                                             / out["bf16_2layer_forward_backward"], 3),
        # bytes of X read once and B written once (packing W, K x F bf16, included)
        "transform_roofline": {
            key: {"bytes": nbytes, "gbps": round(nbytes / out[key] / 1e6, 1),
                  "frac_of_8tbps": round(nbytes / out[key] / 1e6 / 8000.0, 3)}
            for key, nbytes in (("transform_bf16_x", M * F * 2 + M * F * 2 + 2 * F * F * 2),
                                ("transform_relu_f32_x", M * F * 4 + M * F * 2 + 2 * F * F * 2))},
    }), flush=True)


def gather_ceiling(A, B, chunk, vec, stream, steps):
    """Time of the gather-only probe (libdtc_probe.so): the SpMM's B-row loads
    for this ME-TCF, chunk size and column-tile width (vec 2 or 4), nothing
    else -- the practical ceiling of the access pattern.  None when N is not
    a multiple of the tile width or the probe library is not built."""
    import ctypes

    import torch
    path = os.path.join(PKG, "libdtc_probe.so")
    N = int(B.shape[1])
    if vec not in (2, 4) or N % (16 * vec) or not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.dtc_gather_ceiling_out_size.restype = ctypes.c_int64
    T = A.block_count
    atox = A.sparse_atox
    # every slot must index a row of B before the probe reads it
    if T == 0 or int(atox.min()) < 0 or int(atox.max()) >= B.shape[0] or atox.numel() != 8 * T:
        return None
    out = torch.empty(lib.dtc_gather_ceiling_out_size(T, chunk, N, vec), device=B.device)

    def run():
        rc = lib.dtc_gather_ceiling(ctypes.c_void_p(atox.data_ptr()), T, chunk, ctypes.c_void_p(B.data_ptr()), N,
                                    vec, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
        if rc != 0:
            raise RuntimeError(f"gather probe failed ({rc})")
    run()
    torch.cuda.synchronize()
    return timed(stream, run, steps)


def bench_reordered(a, B, C64, Cabs, mask, v, ms, flops, sum_u, W, N, stream, steps):
    """The chosen variant on the TCA-reordered graph (rows and columns
    relabelled as the reference's reordered datasets are): reorder time, the
    new ME-TCF's size, the SpMM time and its error against the float64 product
    of the original graph (rows permuted back)."""
    import numpy as np
    import torch

    from dtc.matrix import DTCMatrix
    from dtc.metrics import gather_model_bytes
    from dtc.reorder import relabel, tca_reorder
    dev = B.device
    M, nnz = a.num_rows, a.nnz
    torch.cuda.synchronize()
    t = time.perf_counter()
    order, rstats = tca_reorder(a.rowptr, a.col, M)
    t_reorder = time.perf_counter() - t
    b = relabel(a, order)
    Ar = DTCMatrix(torch.from_numpy(b.rowptr).to(dev), torch.from_numpy(b.col).to(dev),
                   torch.from_numpy(b.values).to(dev), num_rows=M, num_cols=M)
    back = torch.from_numpy(np.argsort(order)).to(dev)
    Br = torch.empty_like(B)
    Br[back] = B  # row i of the relabelled B is node reorder_id[i]
    Cr = torch.empty((M, N), device=dev, dtype=torch.float32)

    def fr():
        return Ar.spmm(Br, schedule=v[0], precision=v[1], chunk_blocks=v[2], vec=v[3], subtiles=v[4],
                       persistent=v[5], out=Cr, cut_span_blocks=v[6])
    for _ in range(3):
        fr()
    r_ms = timed(stream, fr, steps)
    Crn = Cr[back].cpu().numpy().astype(np.float64)  # rows back in the original order
    dr = np.abs(Crn - C64)
    r_sum_u = Ar.sum_unique_columns()
    log(f"reordered: sumU {sum_u} -> {r_sum_u}, {ms:.3f} -> {r_ms:.3f} ms, reorder {t_reorder:.1f}s")
    return {"reorder_s": round(t_reorder, 2), "stats": rstats, "sum_unique_columns": r_sum_u,
            "tc_blocks": Ar.block_count, "ms": round(r_ms, 4), "gflops": round(flops / r_ms / 1e6, 2),
            "speedup": round(ms / r_ms, 3),
            "gather_model_frac": round(gather_model_bytes(r_sum_u, M, N, nnz, Ar.block_count, W)
                                       / (r_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "err": float((dr[mask] / Cabs[mask]).max())}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--n", type=int, default=0,
                    help="columns of B (default: 256 for rmat and gcn, 128 for uniform and sbm)")
    ap.add_argument("--variant", default="selector",
                    help="selector (learned pick), exhaustive, or "
                         "schedule:precision:chunk:vec:subtiles[:cut_span]")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-table", action="store_true", help="skip the exhaustive variant table")
    ap.add_argument("--traffic-file", default=os.path.join(REPO, "profiles", "traffic_rmat20_n256.json"))
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration)")
    ap.add_argument("--workload", default="rmat", choices=["rmat", "uniform", "sbm", "gcn"],
                    help="rmat: power-law SpMM (default); uniform: 4096^2, 16 per row; "
                         "sbm: 232,965-row block model with hubs; gcn: layer A (X W), bf16 X and W")
    ap.add_argument("--no-reorder", action="store_true",
                    help="skip timing the chosen variant on the TCA-reordered graph (R-MAT only)")
    ap.add_argument("--chunks", default="",
                    help="extra balanced FP32 variants for the table: chunk sizes, optionally "
                         "with the lane width as chunk:vec, e.g. 16,4096 or 8:2,16:2 (vec 4 if omitted)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.n <= 0:
        args.n = 128 if args.workload in ("uniform", "sbm") else 256
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: --gpus {args.gpus}: this DTC-SpMM library runs on one GPU only; "
              "multi-GPU runs are not part of it.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "gcn":
        bench_gcn(args)
        return

    import numpy as np
    import torch

    from dtc import selector as sel
    from dtc.graphs import rmat, sbm_hubs, uniform_rows
    from dtc.matrix import DTCMatrix
    from dtc.metrics import compulsory_bytes, gather_model_bytes, spmm_flops, spmm_source_sha16

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    N = args.n
    t0 = time.time()
    if args.workload == "rmat":
        a = rmat(args.scale, 16, seed=1)
        wdesc = (f"R-MAT scale {args.scale} (a=.57,b=.19,c=.19), edge factor 16, symmetrised, "
                 f"1/sqrt(deg_i deg_j) values")
    elif args.workload == "uniform":
        a = uniform_rows(4096, 4096, 16, seed=0)
        wdesc = "uniform 4096x4096, 16 distinct columns per row (PCG64 seed 0), U[-1,1] values"
    else:
        a = sbm_hubs()
        wdesc = ("block model, 232,965 rows, 50 Zipf(1.1) communities, ~490 per row, 5% cross, "
                 "200 hub rows of 20k-50k, U[0,1] values")
    M, nnz = a.num_rows, a.nnz
    log(f"{args.workload}: M={M} nnz={nnz} generated in {time.time() - t0:.1f}s")

    # ---- ME-TCF on the GPU (timed apart from the SpMM)
    rowptr_d = torch.from_numpy(a.rowptr).to(dev)
    col_d = torch.from_numpy(a.col).to(dev)
    vals_d = torch.from_numpy(a.values).to(dev)
    A = DTCMatrix(rowptr_d, col_d, vals_d, num_rows=M, num_cols=M)  # warm-up build
    pre_ms = []
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        A = DTCMatrix(rowptr_d, col_d, vals_d, num_rows=M, num_cols=M)
        torch.cuda.synchronize()
        pre_ms.append((time.perf_counter() - t) * 1e3)
    T, W = A.block_count, A.num_windows
    sum_u = A.sum_unique_columns()
    log(f"ME-TCF: T={T} W={W} sumU={sum_u} build {np.median(pre_ms):.2f} ms")

    if args.workload == "uniform":
        # the small plain case: B U[-1, 1] from PCG64(1)
        rng_b = np.random.Generator(np.random.PCG64(1))
        B = torch.from_numpy(rng_b.uniform(-1.0, 1.0, (M, N)).astype(np.float32)).to(dev)
        bdesc = "U[-1,1] (PCG64 seed 1)"
    else:
        gen = torch.Generator(device=dev)
        gen.manual_seed(2)
        B = torch.randn((M, N), generator=gen, device=dev, dtype=torch.float32)
        bdesc = "N(0,1) (torch seed 2)"
    C = torch.empty((M, N), device=dev, dtype=torch.float32)

    # ---- float64 reference for the error gate
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    t = time.time()
    Bn = B.cpu().numpy()
    C64, Cabs = oracle.spmm_csr_f64(a.rowptr, a.col, a.values, Bn, M, with_abs=True)
    log(f"float64 reference in {time.time() - t:.1f}s")
    mask = Cabs > 0

    def errors(Cd):
        Ch = Cd.cpu().numpy().astype(np.float64)
        d = np.abs(Ch - C64)
        comp = float((d[mask] / Cabs[mask]).max()) if mask.any() else 0.0
        norm = float(d.max() / np.abs(C64).max())
        return comp, norm

    def runner(v):
        sched, prec, chunk, vec, sub, pers, span = v
        if sched == "tcgnn":  # TC-GNN's SpMM (no TC-block compaction), float32
            from dtc import load_extension
            ext = load_extension()
            return lambda: C.copy_(ext.run_TCGNNSpMM(B, A.rowptr, A.col, A.block_partition, A.edge_to_column,
                                                     A.edge_to_row, values=A.values)[0][:M])
        return lambda: A.spmm(B, schedule=sched, precision=prec, chunk_blocks=chunk, vec=vec,
                              subtiles=sub, persistent=pers, out=C, cut_span_blocks=span)

    # ---- the learned selector's pick
    # the network's first choice, and the fastest of its top TOP_K timed on
    # this matrix (the selector's pick, dtc.selector.pick_variant)
    feats = sel.features(A, N)
    pick = pick1 = None
    ood = []
    if os.path.exists(sel.DEFAULT_MODEL):
        model = sel.MLPSelector.load(sel.DEFAULT_MODEL)
        ood = model.out_of_range(feats)
        pick1 = int(model.choose(feats[None, :])[0])
        if not ood and args.variant == "selector":  # the timing runs only when it decides
            pick, _ = sel.pick_variant(A, B, model)
        log(f"selector: network's choice {sel.variant_name(sel.VARIANTS[pick1])}, "
            + (f"pick {sel.variant_name(sel.VARIANTS[pick])} (top {sel.TOP_K} timed)" if pick is not None
               else f"outside its training range: {', '.join(ood)}" if ood else "no timed pick (fixed variant)"))

    # ---- exhaustive table (the selector's yardstick and fallback)
    table_variants = list(sel.VARIANTS)
    if not args.no_table:
        table_variants += [("balanced", p, 256, 4, 1, False, 0) for p in EXTRA_PRECISIONS]
        table_variants += [("balanced", "fp32", c, 2, 1, False, 0) for c in EXTRA_SMALL_CHUNKS]
        for c in filter(None, args.chunks.split(",")):
            chunk, _, vec = c.partition(":")
            table_variants.append(("balanced", "fp32", int(chunk), int(vec or 4), 1, False, 0))
    table = []
    err_cache = {}
    if not args.no_table:
        for v in table_variants:
            fn = runner(v)
            fn()
            torch.cuda.synchronize()
            if v[1] not in err_cache:
                err_cache[v[1]] = errors(C)
            ms = timed(stream, fn, 5)
            table.append(dict(variant=sel.variant_name(v), precision=v[1], ms=ms,
                              gflops=spmm_flops(nnz, N) / ms / 1e6, err=err_cache[v[1]][0],
                              err_norm=err_cache[v[1]][1]))
            log(f"  {sel.variant_name(v):42s} {ms:8.3f} ms  err={err_cache[v[1]][0]:.2e}")
        # TC-GNN's SpMM as one more row of the exhaustive search (the baseline
        # without TC-block compaction; seconds per call on power-law graphs,
        # minutes on the block model, where it is left out)
        if args.workload != "sbm":
            v = ("tcgnn", "fp32", 0, 0, 0, False, 0)
            fn = runner(v)
            fn()
            torch.cuda.synchronize()
            e_t = errors(C)
            ms = timed(stream, fn, 1)
            table_variants.append(v)
            table.append(dict(variant="tcgnn:fp32", precision="fp32", ms=ms,
                              gflops=spmm_flops(nnz, N) / ms / 1e6, err=e_t[0], err_norm=e_t[1]))
            log(f"  {'tcgnn:fp32':42s} {ms:8.3f} ms  err={e_t[0]:.2e}")

    if args.variant == "selector" and pick is not None and not ood:
        chosen_v = sel.VARIANTS[pick]
        how = f"selector (top {sel.TOP_K} timed)"
    elif args.variant in ("selector", "exhaustive"):
        # exhaustive timing: the selector's fallback (no model, or a matrix
        # outside its training range) and the explicit choice
        if table:
            ok = [(r["ms"], i) for i, r in enumerate(table) if r["err"] <= ERR_TARGET]
            chosen_v = table_variants[min(ok)[1]]
        else:
            chosen_v = sel.VARIANTS[int(np.nanargmin(sel.time_variants(A, B)))]
        how = "exhaustive" + (f" (selector fallback: {', '.join(ood)} outside its training range)"
                              if ood and args.variant == "selector" else "")
    else:
        s_, p_, c_, vec_, sub_, *span_ = args.variant.split(":")
        chosen_v = (s_, p_, int(c_), int(vec_), int(sub_), False, int(span_[0]) if span_ else 0)
        how = "fixed"

    # ---- timed region: K steps of the chosen variant
    fn = runner(chosen_v)
    for _ in range(args.warmup):
        fn()
    torch.cuda.synchronize()
    ms, ms_median = timed_steps(stream, fn, args.steps)
    torch.cuda.synchronize()
    comp_err, norm_err = errors(C)

    flops = spmm_flops(nnz, N)
    gflops = flops / ms / 1e6
    bytes_model = gather_model_bytes(sum_u, M, N, nnz, T, W)
    achieved = bytes_model / (ms * 1e-3) / 1e9
    # HBM traffic measured by counters (FETCH_SIZE + WRITE_SIZE, calibrated on a
    # plain copy) for this workload and variant, used only if it was taken on
    # the current SpMM sources (a stamp of their hash)
    traffic = None
    if args.workload == "rmat" and args.scale == 20 and N == 256 and os.path.exists(args.traffic_file):
        with open(args.traffic_file) as f:
            tj = json.load(f)
        if tj.get("source_sha16") == spmm_source_sha16():
            traffic = tj.get("variants", {}).get(sel.variant_name(chosen_v), {}).get("hbm_bytes_per_launch")

    # ---- plain FP32 path: fastest FP32 variant, timed like the chosen one
    fp32 = None
    fp32_rows = [(r["ms"], i) for i, r in enumerate(table) if r["precision"] == "fp32"]
    if fp32_rows:
        v32 = table_variants[min(fp32_rows)[1]]
        fr = runner(v32)
        for _ in range(3):
            fr()
        fp32_ms, _ = timed_steps(stream, fr, args.steps)  # per-step events, as the chosen one
        fp32 = {"variant": sel.variant_name(v32), "ms": round(fp32_ms, 4),
                "gflops": round(flops / fp32_ms / 1e6, 2), "speedup_of_chosen": round(fp32_ms / ms, 3)}

    # ---- mixed-precision path: the fastest non-fp32 variant within the error
    # gate, timed like the others, and its speed against the fp32 path --
    # reported whatever the pick is
    mixed = None
    mixed_rows = [(r["ms"], i) for i, r in enumerate(table)
                  if r["precision"] != "fp32" and r["err"] <= ERR_TARGET]
    if mixed_rows and fp32:
        i_m = min(mixed_rows)[1]
        fm = runner(table_variants[i_m])
        for _ in range(3):
            fm()
        m_ms, _ = timed_steps(stream, fm, args.steps)
        mixed = {"variant": table[i_m]["variant"], "ms": round(m_ms, 4), "gflops": round(flops / m_ms / 1e6, 2),
                 "err": table[i_m]["err"], "speedup_vs_fp32": round(fp32["ms"] / m_ms, 3), "target_speedup": 1.5,
                 "fastest_failing_gate": min(((r["ms"], r["variant"], r["err"]) for r in table
                                              if r["err"] > ERR_TARGET), default=None)}

    selector_info = None
    if table:
        best = min(range(len(sel.VARIANTS)), key=lambda i: table[i]["ms"])  # the selector's variants
        selector_info = {"pick": None if pick is None else sel.variant_name(sel.VARIANTS[pick]),
                         "how": f"network's top {sel.TOP_K} timed on this matrix",
                         "out_of_training_range": ood,
                         "exhaustive_best": sel.variant_name(sel.VARIANTS[best]),
                         "regret": None if pick is None else round(table[pick]["ms"] / table[best]["ms"] - 1, 4),
                         "network_choice": None if pick1 is None else sel.variant_name(sel.VARIANTS[pick1]),
                         "network_choice_regret": (None if pick1 is None
                                                   else round(table[pick1]["ms"] / table[best]["ms"] - 1, 4))}

    # ---- CPU baseline: the oracle's ME-TCF SpMM port on the host cores
    cpu = None
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        m_host = {"RowWindowOffset": A.row_window_offset.cpu().numpy(),
                  "TCblocktileId": A.tcblocktile_id.cpu().numpy(),
                  "TCblockoffset": A.tcblock_offset.cpu().numpy(),
                  "SparseAToXindex": A.sparse_atox.cpu().numpy()}
        tcv = A.values_tc.cpu().numpy()
        Cc = np.empty((M, N), np.float32)
        reps = 3 if (args.workload == "sbm" and N >= 512) else 5

        def median_time(fn):  # one warm-up, then the median of `reps` runs
            fn()
            ts = []
            for _ in range(reps):
                t = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t)
            return float(np.median(ts))
        t_cpu = median_time(lambda: oracle.spmm_metcf_f32(m_host, Bn, M, tc_values=tcv, nthreads=cores, out=Cc))
        t_build = median_time(lambda: oracle.metcf(a.rowptr, a.col, M, values=a.values, nthreads=cores))
        # the host SpMM of the baseline plan: torch.sparse CSR x dense on the same cores
        torch.set_num_threads(cores)
        As = torch.sparse_csr_tensor(torch.from_numpy(a.rowptr.astype(np.int64)),
                                     torch.from_numpy(a.col.astype(np.int64)),
                                     torch.from_numpy(a.values), size=(M, M))
        Bt = torch.from_numpy(Bn)
        t_ts = median_time(lambda: torch.sparse.mm(As, Bt))
        cpu = {"value": round(flops / t_cpu / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
               "sample": f"full SpMM of this workload ({args.workload}, N={N}) with the CPU "
                         f"oracle's float32 ME-TCF port, median of {reps} runs after a warm-up; host "
                         f"ME-TCF build {t_build * 1e3:.0f} ms and torch.sparse CSR SpMM "
                         f"{flops / t_ts / 1e9:.1f} GFLOP/s timed the same way", "cpu_model": _cpu_model(),
               "host_metcf_build_ms": round(t_build * 1e3, 1),
               "torch_sparse_csr_gflops": round(flops / t_ts / 1e9, 2),
               "torch_version": torch.__version__, "runs": reps}

    # ---- TC-GNN's SpMM on the same matrix (its row of the exhaustive table)
    tcgnn = None
    t_rows = [r for r in table if r["variant"] == "tcgnn:fp32"]
    if t_rows:
        r = t_rows[0]
        tcgnn = {"ms": round(r["ms"], 3), "gflops": round(r["gflops"], 2), "err": r["err"],
                 "chosen_speedup_over_tcgnn": round(r["ms"] / ms, 2)}

    # ---- practical ceiling: the same B-row loads with nothing else
    ceiling = None
    if chosen_v[0] == "balanced" and chosen_v[3] in (2, 4) and chosen_v[4] == 1:
        g_ms = gather_ceiling(A, B, chosen_v[2], chosen_v[3], stream, args.steps)
        if g_ms:
            ceiling = {"gather_only_ms": round(g_ms, 4), "spmm_frac_of_gather_ceiling": round(g_ms / ms, 4)}

    # ---- the same SpMM on the TCA-reordered graph (rows and columns relabelled)
    reordered = None
    if args.workload == "rmat" and not args.no_reorder:
        try:
            reordered = bench_reordered(a, B, C64, Cabs, mask, chosen_v, ms, flops, sum_u, W, N, stream,
                                        args.steps)
        except Exception as exc:  # reported, never fatal to the headline measurement
            reordered = {"error": f"{type(exc).__name__}: {exc}"}
            log(f"reordered run failed: {exc}")
        # what the reorder advisor (dtc/reorder_advice.py) predicted from the
        # original matrix's features, beside the measured speedup
        try:
            from dtc.reorder_advice import THRESHOLD, ReorderAdvisor
            pred = float(ReorderAdvisor.load().predict_gain(feats[None, :])[0])
            reordered["advisor"] = {"predicted_speedup": round(float(np.exp(pred)), 3),
                                    "advises_reorder": bool(pred > THRESHOLD)}
        except Exception as exc:
            reordered["advisor"] = {"error": f"{type(exc).__name__}: {exc}"}

    dtypes = {"fp32": "f32", "bf16x6": "bf16x6 (f32 accumulate)", "bf16x3": "bf16x3 (f32 accumulate)",
              "fp16": "f16 (f32 accumulate)", "bf16": "bf16 (f32 accumulate)"}
    result = {
        "metric": "SpMM GFLOP/s (2*nnz*N) + % HBM roofline, mixed vs FP32 at <=1e-5 rel err",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "ms_median_step": round(ms_median, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes[chosen_v[1]],
        "data": f"synthetic (seeded {args.workload} matrix, B {bdesc})",
        "config": {"workload": f"{wdesc}; B {M}x{N} f32",
                   "num_rows": M, "nnz": nnz, "N": N, "tc_blocks": T, "row_windows": W,
                   "sum_unique_columns": sum_u, "variant": sel.variant_name(chosen_v), "chosen_by": how,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_frac": (None if traffic is None
                                      else round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
                     "model": "gather: 4 N sum_w U_w + 4 M N + 5 nnz + 36 T + 4 (W + 1) bytes per launch"},
        "cpu_baseline": cpu,
        "error_vs_fp64": {"componentwise": comp_err, "normwise": norm_err, "target": ERR_TARGET},
        "fp32_path": fp32,
        "mixed_path": mixed,
        "selector": selector_info,
        "gather_ceiling": ceiling,
        "tcgnn_baseline": tcgnn,
        "reordered": reordered,
        "preprocess_ms": round(float(np.median(pre_ms)), 3),
        "bytes_model_per_launch": bytes_model,
        "compulsory_bytes": compulsory_bytes(M, M, N, nnz),
        "variants": [{k: (round(v, 4) if isinstance(v, float) and not k.startswith("err") else v)
                      for k, v in r.items()} for r in table],
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
