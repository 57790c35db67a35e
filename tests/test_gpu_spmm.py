"""GPU SpMM vs the float64 CSR oracle, for both schedules, every precision,
vector shapes and N, plus the reference's own call sequence.

Tolerances (componentwise: max |C - C64| / (|A| |B|), float64 denominators):
  fp32    4e-6   float32 operands and accumulation
  bf16x6  4e-6   3-limb bf16 split, products of limb order <= 2
  bf16x3  1e-4   2-limb split, hi*hi + hi*lo + lo*hi (dropped terms ~3 * 2^-16)
  fp16    4e-3   single fp16 pass, unit roundoff 2^-11 per operand
  bf16    3e-2   single bf16 pass, unit roundoff 2^-8 per operand
"""
import numpy as np
import pytest
import torch

import oracle
from helpers import gpu_metcf, load_fixture_18, random_csr, say

pytestmark = pytest.mark.gpu

TOL = {"fp32": 4e-6, "bf16x6": 4e-6, "bf16x3": 1e-4, "fp16": 4e-3, "bf16": 3e-2}
PLANS = ["float_nonsplit", "float2_nonsplit", "float2_split", "float4_nonsplit", "float4_split"]


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def componentwise(C, C64, Cabs):
    C = C.astype(np.float64)
    d = np.abs(C - C64)
    m = Cabs > 0
    e = float((d[m] / Cabs[m]).max()) if m.any() else 0.0
    z = float(d[~m].max()) if (~m).any() else 0.0
    return max(e, z)


class Case:
    def __init__(self, ext, rowptr, col, M, vals=None, name=""):
        self.rowptr, self.col, self.M, self.vals, self.name = rowptr, col, M, vals, name
        say("build", name, f"M={M} nnz={col.shape[0]}")
        self.m = gpu_metcf(ext, rowptr, col, M, values=vals)
        self.dev = self.m["_dev"]
        self.tcv = self.dev[6] if vals is not None else None

    def run(self, ext, B, schedule, precision, use_vals=True, **kw):
        d = self.dev
        idx = d[0] if schedule == "windows" else d[1]
        v = self.tcv if use_vals else None
        return ext.spmm(B, idx, d[2], d[3], d[4], self.M, schedule=schedule, values=v,
                        precision=precision, **kw)

    def ref(self, Bn, use_vals=True):
        v = self.vals if use_vals else None
        return oracle.spmm_csr_f64(self.rowptr, self.col, v, Bn, self.M, with_abs=True)


@pytest.fixture(scope="module")
def c1(ext):
    from dtc.graphs import uniform_rows
    a = uniform_rows(4096, 4096, 16, seed=0)
    return Case(ext, a.rowptr, a.col, a.num_rows, a.values, "C1")


@pytest.fixture(scope="module")
def rmat12(ext):
    from dtc.graphs import rmat
    a = rmat(12, 16, seed=1)
    return Case(ext, a.rowptr, a.col, a.num_rows, a.values, "rmat12")


@pytest.fixture(scope="module")
def ragged(ext):
    rng = np.random.default_rng(21)
    M = 1000 + 7  # partial last window
    rowptr, col = random_csr(rng, M, M, 9.0, sorted_rows=False, empty_frac=0.25)
    # a few very long rows so windows span many chunks
    extra_rows = [3, 500, 1006]
    rows = [list(col[rowptr[r]:rowptr[r + 1]]) for r in range(M)]
    for r in extra_rows:
        rows[r] = sorted(set(rows[r]) | set(rng.choice(M, 700, replace=False).tolist()))
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.array(r, np.int32) for r in rows]).astype(np.int32)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    return Case(ext, rowptr, col, M, vals, "ragged")


@pytest.mark.parametrize("precision", ["fp32", "bf16x6", "bf16x3", "fp16", "bf16"])
@pytest.mark.parametrize("schedule", ["windows", "balanced"])
@pytest.mark.parametrize("N", [16, 128, 256])
def test_precisions_c1(ext, c1, precision, schedule, N):
    rng = np.random.default_rng(N)
    Bn = rng.uniform(-1, 1, (c1.M, N)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    for use_vals in (True, False):
        say("spmm C1", precision, schedule, N, use_vals)
        C = c1.run(ext, B, schedule, precision, use_vals=use_vals).cpu().numpy()
        C64, Cabs = c1.ref(Bn, use_vals)
        err = componentwise(C, C64, Cabs)
        assert err <= TOL[precision], (precision, schedule, N, use_vals, err)


@pytest.mark.parametrize("N", [1, 3, 5, 8, 17, 32, 64, 100, 128, 260, 512])
@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_shapes_ragged(ext, ragged, N, schedule):
    rng = np.random.default_rng(1000 + N)
    Bn = rng.standard_normal((ragged.M, N)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    C64, Cabs = ragged.ref(Bn)
    for vec, sub in [(0, 0), (1, 1), (2, 1), (4, 1), (4, 2)]:
        say("spmm ragged", schedule, N, vec, sub)
        C = ragged.run(ext, B, schedule, "fp32", vec=vec, subtiles=sub, chunk_blocks=5).cpu().numpy()
        err = componentwise(C, C64, Cabs)
        assert err <= TOL["fp32"], (schedule, N, vec, sub, err)


@pytest.mark.parametrize("chunk", [1, 2, 3, 4, 5, 8, 64, 1000])
@pytest.mark.parametrize("persistent", [False, True])
def test_balanced_chunks(ext, ragged, rmat12, chunk, persistent):
    for case in (ragged, rmat12):
        rng = np.random.default_rng(chunk)
        Bn = rng.standard_normal((case.M, 64)).astype(np.float32)
        B = torch.from_numpy(Bn).cuda()
        say("spmm chunks", case.name, chunk, persistent)
        C = case.run(ext, B, "balanced", "fp32", chunk_blocks=chunk, persistent=persistent).cpu().numpy()
        C64, Cabs = case.ref(Bn)
        assert componentwise(C, C64, Cabs) <= TOL["fp32"], (case.name, chunk, persistent)


@pytest.fixture(scope="module")
def gappy(ext):
    """Long runs of empty row windows (at the start, in the middle, at the
    end) around windows short and long, so that chunks begin and end inside
    gaps, inside windows cut at chunk boundaries and on window starts."""
    rng = np.random.default_rng(33)
    M = 3000 + 5
    rows = [[] for _ in range(M)]
    for r in list(range(400, 1000)) + list(range(2200, 2900)):
        if rng.random() < 0.7:
            rows[r] = sorted(rng.choice(M, int(rng.integers(1, 12)), replace=False).tolist())
    for r in (401, 950, 2201, 2899):  # long rows: windows of many blocks
        rows[r] = sorted(rng.choice(M, 900, replace=False).tolist())
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.array(r, np.int32) for r in rows]).astype(np.int32)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    return Case(ext, rowptr, col, M, vals, "gappy")


@pytest.mark.parametrize("chunk", [1, 3, 8, 64, 113, 1000, 100000])
@pytest.mark.parametrize("persistent", [False, True])
def test_balanced_empty_windows(ext, gappy, chunk, persistent):
    # The output starts as NaN: every element must be written (zeros for the
    # empty windows, sums elsewhere), whatever the chunking.
    rng = np.random.default_rng(chunk)
    Bn = rng.standard_normal((gappy.M, 64)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    C64, Cabs = gappy.ref(Bn)
    out = torch.full((gappy.M, 64), float("nan"), device="cuda")
    say("spmm gappy", chunk, persistent)
    gappy.run(ext, B, "balanced", "fp32", chunk_blocks=chunk, persistent=persistent, out=out)
    C = out.cpu().numpy()
    assert np.isfinite(C).all()
    assert (C[:400] == 0).all() and (C[1000:2200] == 0).all() and (C[2900:] == 0).all()
    assert componentwise(C, C64, Cabs) <= TOL["fp32"], (chunk, persistent)


@pytest.mark.parametrize("M", [1, 17, 100])
def test_empty_matrix_writes_zeros(ext, M):
    # No entries, so no TC blocks: every output row is an empty window and
    # must be written as zeros over a NaN-filled output, by every schedule and
    # precision, and by the reference's operator names.
    rowptr = np.zeros(M + 1, np.int32)
    case = Case(ext, rowptr, np.zeros(0, np.int32), M, np.zeros(0, np.float32), f"empty M={M}")
    assert case.m["block_count"] == 0
    B = torch.randn((M, 64), device="cuda")
    for schedule in ("windows", "balanced"):
        for precision in ("fp32", "bf16x6"):
            say("spmm empty", M, schedule, precision)
            out = torch.full((M, 64), float("nan"), device="cuda")
            case.run(ext, B, schedule, precision, out=out)
            assert (out == 0).all(), (M, schedule, precision)
    d = case.dev
    (C,) = ext.run_DTCSpMM(B, d[0], d[2], d[3], d[4], M, 0, "float4_split")
    assert C.shape == (M, 64) and (C == 0).all()
    (C,) = ext.run_DTCSpMM_balance(B, d[1], d[2], d[3], d[4], M, "float4_split")
    assert C.shape == (M, 64) and (C == 0).all()


def test_dtc_bench_timed_loop(ext, ragged, tmp_path, monkeypatch):
    """DTC_BENCH=1 reproduces the reference's 1000 timed launches and its CSV
    record "DTCSpMM:,<ms>,<GFLOP/s>," (DTCSpMM_kernel.cu:931-937), appended
    without a newline, for both operator names; the result is that of one
    plain launch."""
    rng = np.random.default_rng(77)
    B = torch.from_numpy(rng.standard_normal((ragged.M, 64)).astype(np.float32)).cuda()
    d = ragged.dev
    nnz = int(ragged.col.shape[0])
    (Cw1,) = ext.run_DTCSpMM(B, d[0], d[2], d[3], d[4], ragged.M, nnz, "float4_split")
    (Cb1,) = ext.run_DTCSpMM_balance(B, d[1], d[2], d[3], d[4], ragged.M, "float4_split")
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("DTC_BENCH", "1")
    say("DTC_BENCH loop")
    (Cw2,) = ext.run_DTCSpMM(B, d[0], d[2], d[3], d[4], ragged.M, nnz, "float4_split")
    (Cb2,) = ext.run_DTCSpMM_balance(B, d[1], d[2], d[3], d[4], ragged.M, "float4_split")
    fields = (tmp_path / "DTCSpMM_exe_time_and_throughput.csv").read_text().split(",")
    assert len(fields) == 7 and fields[0] == fields[3] == "DTCSpMM:" and fields[6] == ""
    for ms, gflops in ((float(fields[1]), float(fields[2])), (float(fields[4]), float(fields[5]))):
        assert ms > 0 and abs(gflops - 2.0 * nnz * 64 / (ms * 1e6)) <= 1e-3 * gflops
    assert torch.equal(Cw1, Cw2)
    assert torch.allclose(Cb1, Cb2, rtol=1e-5, atol=1e-5)  # cut windows add in run order


def test_auto_schedule(ext):
    """DTCMatrix.spmm's default schedule: unbalanced when no row window holds
    more than AUTO_WINDOW_BLOCKS TC blocks (here 2 per window), balanced when
    one does (a 3,000-column row: 375 blocks); same result as the forced
    schedule, within float32 against float64."""
    from dtc import selector as sel
    from dtc.matrix import DTCMatrix
    rng = np.random.default_rng(12)
    M = 4096
    rowptr, col = random_csr(rng, M, M, 6.0)
    short = (rowptr, col)
    rows = [list(col[rowptr[r]:rowptr[r + 1]]) for r in range(M)]
    rows[77] = sorted(set(rows[77]) | set(rng.choice(M, 3000, replace=False).tolist()))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    long_ = (rp, np.concatenate([np.array(r, np.int32) for r in rows]).astype(np.int32))
    for (rowptr_, col_), want in ((short, "windows"), (long_, "balanced")):
        vals = rng.uniform(-1, 1, col_.shape[0]).astype(np.float32)
        A = DTCMatrix(rowptr_, col_, vals, num_rows=M, num_cols=M)
        assert sel.default_schedule(A.max_window_blocks) == want, A.max_window_blocks
        Bn = rng.standard_normal((M, 64)).astype(np.float32)
        B = torch.from_numpy(Bn).cuda()
        say("auto schedule", want, A.max_window_blocks)
        C = A.spmm(B)
        assert torch.allclose(C, A.spmm(B, schedule=want), rtol=1e-5, atol=1e-5)
        C64, Cabs = oracle.spmm_csr_f64(rowptr_, col_, vals, Bn, M, with_abs=True)
        assert componentwise(C.cpu().numpy(), C64, Cabs) <= TOL["fp32"]


def test_schedules_agree_rmat(ext, rmat12):
    rng = np.random.default_rng(5)
    Bn = rng.standard_normal((rmat12.M, 128)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    say("spmm rmat12 schedules")
    Cw = rmat12.run(ext, B, "windows", "fp32")
    Cb = rmat12.run(ext, B, "balanced", "fp32", chunk_blocks=16)
    C64, Cabs = rmat12.ref(Bn)
    assert componentwise(Cw.cpu().numpy(), C64, Cabs) <= TOL["fp32"]
    assert componentwise(Cb.cpu().numpy(), C64, Cabs) <= TOL["fp32"]
    assert torch.allclose(Cw, Cb, rtol=1e-5, atol=1e-6)


def test_repeated_launches_identical(ext, ragged):
    # The reference's balanced op accumulates into its output on every one of
    # its 1000 timed launches; every launch here starts from the zeroed rows.
    # Windows merged from three or more chunks add in a run-dependent order,
    # so repeated results agree to float32 rounding, not bitwise.
    rng = np.random.default_rng(8)
    Bn = rng.standard_normal((ragged.M, 32)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    C64, Cabs = ragged.ref(Bn)
    say("spmm repeat")
    out = torch.full((ragged.M, 32), 7.0, device="cuda")
    first = ragged.run(ext, B, "balanced", "fp32", chunk_blocks=3).clone()
    for _ in range(3):
        ragged.run(ext, B, "balanced", "fp32", chunk_blocks=3, out=out)
        assert torch.allclose(out, first, rtol=1e-5, atol=1e-5)
        assert componentwise(out.cpu().numpy(), C64, Cabs) <= TOL["fp32"]


@pytest.mark.parametrize("persistent", [False, True])
def test_deterministic_merge_bitwise(ext, ragged, rmat12, persistent):
    """The deterministic merge: windows cut across chunks are summed in chunk
    order from partial tiles, so repeated calls are bitwise identical (output
    pre-filled with NaN: every element must be written), for chunk sizes that
    cut windows into many pieces, every tile shape and bf16 B."""
    for case in (ragged, rmat12):
        rng = np.random.default_rng(90)
        for N in (1, 17, 64, 130):
            Bn = rng.standard_normal((case.M, N)).astype(np.float32)
            B = torch.from_numpy(Bn).cuda()
            C64, Cabs = case.ref(Bn)
            for chunk in (1, 3, 8, 64):
                for vec, sub in [(0, 0), (1, 1), (2, 1), (4, 1), (4, 2)]:  # narrowed to what N allows
                    say("deterministic", case.name, persistent, N, chunk, vec, sub)
                    first = None
                    for _ in range(3):
                        out = torch.full((case.M, N), float("nan"), device="cuda")
                        case.run(ext, B, "balanced", "fp32", chunk_blocks=chunk, vec=vec, subtiles=sub,
                                 persistent=persistent, deterministic=True, out=out)
                        if first is None:
                            first = out.clone()
                            assert componentwise(out.cpu().numpy(), C64, Cabs) <= TOL["fp32"], (N, chunk, vec)
                        else:
                            assert torch.equal(out, first), (case.name, N, chunk, vec, sub)
    # bf16 B (gathered as stored) through the same merge
    Bb = torch.randn((rmat12.M, 256), device="cuda").to(torch.bfloat16)
    Bn = Bb.float().cpu().numpy()
    C64, Cabs = rmat12.ref(Bn)
    say("deterministic bf16 B", persistent)
    outs = [rmat12.run(ext, Bb, "balanced", "fp32", chunk_blocks=4, persistent=persistent, deterministic=True)
            for _ in range(3)]
    assert componentwise(outs[0].cpu().numpy(), C64, Cabs) <= TOL["fp32"]
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_reference_call_sequence_degree(ext):
    """The sequence of scripts/DTCSpMM/run_DTC_SpMM.py:45-79 with X = ones:
    every plan returns C[r, :] = deg(r) exactly."""
    from dtc.graphs import uniform_rows
    for name, (rowptr, col, M) in {"fixture18": load_fixture_18()[:3],
                                   "C1": (lambda a: (a.rowptr, a.col, a.num_rows))(
                                       uniform_rows(4096, 4096, 16, seed=0, values=None))}.items():
        dev = "cuda"
        W = (M + 15) // 16
        column_index = torch.from_numpy(col).to(dev)
        row_pointers = torch.from_numpy(rowptr).to(dev)
        bp = torch.zeros(W, dtype=torch.int32, device=dev)
        e2c = torch.zeros(col.shape[0], dtype=torch.int32, device=dev)
        e2r = torch.zeros(col.shape[0], dtype=torch.int32, device=dev)
        say("reference sequence", name)
        RowWindowOffset, TCblockRowid, TCblocktileId, TCblockoffset, SparseAToXindex, block_count = \
            ext.preprocess_gpu(column_index, row_pointers, M, 16, 8, bp, e2c, e2r)
        deg = torch.from_numpy(np.diff(rowptr).astype(np.float32)).to(dev)
        for feat in (1, 7, 128, 256):
            X = torch.ones((M, feat), device=dev)
            expect = deg[:, None].expand(M, feat)
            for plan in PLANS:
                out = ext.run_DTCSpMM(X, RowWindowOffset, TCblocktileId, TCblockoffset, SparseAToXindex, M,
                                      col.shape[0], plan)
                assert isinstance(out, list) and len(out) == 1
                assert torch.equal(out[0], expect), (name, feat, plan)
            for plan in ("float_nonsplit", "float4_split"):
                out = ext.run_DTCSpMM_balance(X, TCblockRowid, TCblocktileId, TCblockoffset, SparseAToXindex, M,
                                              plan)[0]
                assert torch.equal(out, expect), (name, feat, plan, "balance")
            out = ext.DTCSpMM_gcn(X, RowWindowOffset, TCblocktileId, TCblockoffset, SparseAToXindex, M,
                                  col.shape[0])[0]
            assert torch.equal(out, expect), (name, feat, "gcn")
            for prec in ("bf16", "fp16", "bf16x3", "bf16x6"):
                out = ext.run_DTCSpMM(X, RowWindowOffset, TCblocktileId, TCblockoffset, SparseAToXindex, M,
                                      col.shape[0], "float4_split", precision=prec)[0]
                assert torch.equal(out, expect), (name, feat, prec)


def test_b_row_guard(ext):
    # Columns at or beyond the B row bound contribute nothing, as the
    # reference's dense_bound check (DTCSpMM_kernel.cu:2735).
    rowptr = np.array([0, 2, 3], np.int32)
    col = np.array([0, 5, 1], np.int32)
    case = Case(ext, rowptr, col, 2, None, "guard")
    B = torch.ones((3, 16), device="cuda")  # rows 0..2 only
    say("spmm B row guard")
    C = case.run(ext, B, "windows", "fp32")
    assert torch.equal(C[:, 0].cpu(), torch.tensor([1.0, 1.0]))


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_duplicate_entries_add(ext, schedule):
    """A CSR with repeated (row, column) pairs, values and A = 1: the stored
    elements that share a tile id are added, so the ME-TCF product equals the
    CSR product (and the degree answer counts every copy)."""
    rng = np.random.default_rng(77)
    M = 1500
    rowptr, col = random_csr(rng, M, M, 10.0, sorted_rows=False, empty_frac=0.1, dup_frac=0.7)
    # a row holding one column many times, and a hub with repeats
    rows = [list(col[rowptr[r]:rowptr[r + 1]]) for r in range(M)]
    rows[4] = [9] * 70
    rows[700] = list(rng.integers(0, M, 600))
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    case = Case(ext, rowptr, col, M, vals, "duplicates")
    Bn = rng.standard_normal((M, 64)).astype(np.float32)
    B = torch.from_numpy(Bn).cuda()
    for use_vals in (True, False):
        C64, Cabs = case.ref(Bn, use_vals)
        for precision in ("fp32", "bf16x6"):
            for chunk in (1, 4, 64):
                say("spmm duplicates", schedule, use_vals, precision, chunk)
                C = case.run(ext, B, schedule, precision, use_vals=use_vals, chunk_blocks=chunk).cpu().numpy()
                assert componentwise(C, C64, Cabs) <= TOL[precision], (use_vals, precision, chunk)
    ones = torch.ones((M, 16), device="cuda")
    deg = torch.from_numpy(np.diff(rowptr).astype(np.float32)).cuda()
    C = case.run(ext, ones, schedule, "fp32", use_vals=False)
    assert torch.equal(C, deg[:, None].expand_as(C))


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("deterministic", [False, True])
def test_cut_span(ext, ragged, gappy, rmat12, persistent, deterministic):
    """Windows of up to cut_span_blocks TC blocks are kept whole and longer
    ones are cut at its multiples: every chunk/span pair gives the FP64
    product (output pre-filled with NaN: every element written), and the
    deterministic merge stays bitwise repeatable with pieces longer than a
    chunk."""
    for case in (ragged, gappy, rmat12):
        rng = np.random.default_rng(7)
        N = 40
        Bn = rng.standard_normal((case.M, N)).astype(np.float32)
        B = torch.from_numpy(Bn).cuda()
        C64, Cabs = case.ref(Bn)
        for chunk in (1, 3, 8, 64):
            for mult in (1, 2, 5, 16, 64):
                span = chunk * mult
                say("cut span", case.name, persistent, deterministic, chunk, span)
                first = None
                for _ in range(2 if deterministic else 1):
                    out = torch.full((case.M, N), float("nan"), device="cuda")
                    case.run(ext, B, "balanced", "fp32", chunk_blocks=chunk, cut_span_blocks=span,
                             persistent=persistent, deterministic=deterministic, out=out)
                    if first is None:
                        first = out.clone()
                        C = out.cpu().numpy()
                        assert np.isfinite(C).all(), (case.name, chunk, span)
                        assert componentwise(C, C64, Cabs) <= TOL["fp32"], (case.name, chunk, span)
                    else:
                        assert torch.equal(out, first), (case.name, chunk, span)


def test_cut_span_must_be_a_chunk_multiple(ext, rmat12):
    B = torch.randn((rmat12.M, 16), device="cuda")
    say("cut span argument check")
    with pytest.raises(RuntimeError, match="cut_span_blocks"):
        rmat12.run(ext, B, "balanced", "fp32", chunk_blocks=8, cut_span_blocks=12)
