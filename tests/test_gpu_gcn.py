"""bfloat16-B SpMM and the GCN layer H = A (X W).

bf16 B is taken exactly, so the float64 reference multiplies the same bf16
values: with A split into three bf16 limbs the componentwise error bound is
the float32 one (tolerance 4e-6); with A as one limb (precision "bf16") and
A = 1 the result is exact for small integer data.  The GCN layer's bf16 X W
rounds to bf16 (unit roundoff 2^-8), so its comparisons against float64
use 2e-2 relative tolerances."""
import numpy as np
import pytest
import torch

import oracle
from helpers import random_csr, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def componentwise(C, C64, Cabs):
    d = np.abs(C.astype(np.float64) - C64)
    m = Cabs > 0
    return max(float((d[m] / Cabs[m]).max()) if m.any() else 0.0,
               float(d[~m].max()) if (~m).any() else 0.0)


@pytest.mark.parametrize("N", [4, 8, 12, 64, 128, 256, 260])
@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_bf16_b_vs_oracle(ext, N, schedule):
    from dtc.matrix import DTCMatrix
    rng = np.random.default_rng(N)
    M = 1500
    rowptr, col = random_csr(rng, M, M, 12.0, sorted_rows=False, empty_frac=0.2)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=M)
    Bb = torch.randn((M, N), device="cuda").to(torch.bfloat16)
    Bn = Bb.float().cpu().numpy()
    C64, Cabs = oracle.spmm_csr_f64(rowptr, col, vals, Bn, M, with_abs=True)
    say("bf16 B spmm", schedule, N)
    C = A.spmm(Bb, schedule=schedule, chunk_blocks=7).cpu().numpy()
    assert componentwise(C, C64, Cabs) <= 4e-6, (schedule, N)
    ones = torch.ones((M, N), device="cuda", dtype=torch.bfloat16)
    deg = np.diff(rowptr).astype(np.float32)
    for prec in ("bf16x6", "bf16"):
        C1 = A.spmm(ones, schedule=schedule, precision=prec, use_values=False).cpu().numpy()
        assert np.array_equal(C1, np.repeat(deg[:, None], N, 1)), (schedule, N, prec)


def _dense_reference(a, values, X, W, dH):
    """float64 CPU reference of H = A (X W) and its gradients (torch autograd)."""
    M = a.num_rows
    rows = torch.from_numpy(np.repeat(np.arange(M), np.diff(a.rowptr)))
    cols = torch.from_numpy(a.col.astype(np.int64))
    v = torch.from_numpy(values.astype(np.float64)).requires_grad_(True)
    Xd = torch.from_numpy(X.astype(np.float64)).requires_grad_(True)
    Wd = torch.from_numpy(W.astype(np.float64)).requires_grad_(True)
    Ad = torch.zeros((M, a.num_cols), dtype=torch.float64).index_put((rows, cols), v)
    H = Ad @ (Xd @ Wd)
    (H * torch.from_numpy(dH.astype(np.float64))).sum().backward()
    return H.detach().numpy(), v.grad.numpy(), Xd.grad.numpy(), Wd.grad.numpy()


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_gcn_layer_rmat_2048(ext, schedule):
    """R-MAT scale 11 (2,048 nodes), symmetric normalisation, X 2048x256 and
    W 256x256 in bf16 (Xavier), gradients vs float64 autograd and float64
    finite differences."""
    from dtc.gcn import gcn_layer
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix
    a = rmat(11, 16, seed=1)
    M, F = a.num_rows, 256
    rng = np.random.default_rng(7)
    X = torch.from_numpy(rng.standard_normal((M, F)).astype(np.float32)).to(torch.bfloat16)
    lim = np.sqrt(6.0 / (F + F))
    W = torch.from_numpy(rng.uniform(-lim, lim, (F, F)).astype(np.float32)).to(torch.bfloat16)
    dH = rng.standard_normal((M, F)).astype(np.float32)
    say("gcn layer", schedule)
    A = DTCMatrix(a.rowptr, a.col, None, num_rows=M, num_cols=M)
    v = torch.from_numpy(a.values).cuda().requires_grad_(True)
    Xc = X.cuda().requires_grad_(True)
    Wc = W.cuda().requires_grad_(True)
    H = gcn_layer(v, Xc, Wc, A, schedule=schedule)
    H.backward(torch.from_numpy(dH).cuda())
    H64, gv64, gX64, gW64 = _dense_reference(a, a.values, X.float().numpy(), W.float().numpy(), dH)

    def rel(x, ref):
        return float(np.abs(np.asarray(x, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))
    assert H.dtype == torch.float32
    assert rel(H.detach().cpu().numpy(), H64) < 2e-2
    assert rel(v.grad.cpu().numpy(), gv64) < 2e-2
    assert rel(Xc.grad.float().cpu().numpy(), gX64) < 2e-2
    assert rel(Wc.grad.float().cpu().numpy(), gW64) < 2e-2

    # float64 central finite differences of L = <H, dH> along random
    # directions of values, X and W, against the GPU gradients' projections
    # (the loss is bilinear per argument, so the difference quotient is exact
    # up to float64 rounding; the bf16 forward sets the 2e-2 tolerance).
    rows = np.repeat(np.arange(M), np.diff(a.rowptr))
    Xd, Wd, dHd = X.double().numpy(), W.double().numpy(), dH.astype(np.float64)

    def loss(vals, Xm, Wm):
        XW = Xm @ Wm
        Hm = np.zeros((M, F))
        np.add.at(Hm, rows, vals[:, None] * XW[a.col])
        return float((Hm * dHd).sum())

    rng2 = np.random.default_rng(11)
    base = (a.values.astype(np.float64), Xd, Wd)
    grads = (v.grad.double().cpu().numpy(), Xc.grad.double().cpu().numpy(), Wc.grad.double().cpu().numpy())
    for i in range(3):
        d = rng2.standard_normal(base[i].shape)
        eps = 1e-3 * np.abs(base[i]).max() / np.abs(d).max()
        plus = list(base)
        minus = list(base)
        plus[i] = base[i] + eps * d
        minus[i] = base[i] - eps * d
        fd = (loss(*plus) - loss(*minus)) / (2 * eps)
        terms = grads[i] * d  # the projection sums these; its rounding grows like their 2-norm
        assert abs(fd - float(terms.sum())) <= 2e-2 * (abs(fd) + float(np.sqrt((terms ** 2).sum()))), i


@pytest.mark.parametrize("x_dtype", ["float32", "bfloat16"])
@pytest.mark.parametrize("relu", [False, True])
def test_gcn_transform_vs_float64(ext, x_dtype, relu):
    """The GCN transform B = bf16(act(X) W) against the float64 product of the
    same bf16 operands (act(X) rounded to bf16 as the kernel's prologue does):
    |B - P| <= 2^-8 |P| (bf16 rounding of the output) + 2^-20 (|act(X)| |W|)
    (float32 accumulation).  Shapes cover partial 64-row tiles, a K chunk of
    32 after a full 128 (K = 288), F = 64 (three of a workgroup's four waves
    idle) and F = 320 (a second workgroup column)."""
    rng = np.random.default_rng(3 + int(relu) + (0 if x_dtype == "float32" else 10))
    for M in (1, 100, 1000):
        for K in (32, 96, 256, 288):
            for F in (64, 192, 320):
                X = torch.from_numpy(rng.standard_normal((M, K)).astype(np.float32)).cuda()
                if x_dtype == "bfloat16":
                    X = X.to(torch.bfloat16)
                W = torch.from_numpy(rng.standard_normal((K, F)).astype(np.float32)).cuda().to(torch.bfloat16)
                say("gcn transform", x_dtype, relu, M, K, F)
                B = ext.gcn_transform(X, ext.gcn_pack_weights(W), F, relu)
                assert B.dtype == torch.bfloat16 and B.shape == (M, F)
                Xa = (torch.relu(X) if relu else X).to(torch.bfloat16).double().cpu().numpy()
                Wd = W.double().cpu().numpy()
                P = Xa @ Wd
                bound = 2.0 ** -8 * np.abs(P) + 2.0 ** -20 * (np.abs(Xa) @ np.abs(Wd))
                err = np.abs(B.double().cpu().numpy() - P)
                assert (err <= bound).all(), (M, K, F, float((err - bound).max()))


def test_gcn_transform_rejects_shapes(ext):
    say("gcn transform shape checks")
    W = torch.zeros((48, 64), device="cuda", dtype=torch.bfloat16)
    with pytest.raises(RuntimeError, match="multiple of 32"):
        ext.gcn_pack_weights(W)
    W = torch.zeros((64, 100), device="cuda", dtype=torch.bfloat16)
    with pytest.raises(RuntimeError, match="multiple of 64"):
        ext.gcn_pack_weights(W)
    Wp = ext.gcn_pack_weights(torch.zeros((64, 64), device="cuda", dtype=torch.bfloat16))
    with pytest.raises(ValueError, match="K \\* F"):
        ext.gcn_transform(torch.zeros((8, 32), device="cuda"), Wp, 64, False)


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_gcn_two_layer_rmat_2048(ext, schedule):
    """Two fused layers H2 = A (relu(A (X W1)) W2), R-MAT scale 11 (2,048
    nodes), X 2048x256, W1 256x128, W2 128x64 in bf16: the second layer's
    transform takes the first layer's float32 output and applies the ReLU and
    the bf16 cast itself.  Output and the gradients of the values (used by
    both layers), X, W1 and W2 against float64 autograd with a dense A; the
    bf16 roundings of X W1, relu(H1) and relu(H1) W2 set the 3e-2 normwise
    tolerance.  The float64 network takes its ReLU pattern from the GPU's
    first layer (entries of H1 within a bf16 rounding of zero can fall on
    either side, and each one that does changes a gradient term outright);
    the patterns agree on all but a small fraction of entries."""
    from dtc.gcn import fused_transform_applies, gcn_two_layer
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix
    a = rmat(11, 16, seed=2)
    M = a.num_rows
    rng = np.random.default_rng(17)
    X = torch.from_numpy(rng.standard_normal((M, 256)).astype(np.float32)).to(torch.bfloat16)
    W1 = torch.from_numpy(rng.uniform(-0.1, 0.1, (256, 128)).astype(np.float32)).to(torch.bfloat16)
    W2 = torch.from_numpy(rng.uniform(-0.15, 0.15, (128, 64)).astype(np.float32)).to(torch.bfloat16)
    dH = rng.standard_normal((M, 64)).astype(np.float32)
    # layer 1: bf16 X, nothing to fuse (library GEMM); layer 2: the fused ReLU + cast
    assert not fused_transform_applies(X, W1, relu=False)
    assert fused_transform_applies(torch.zeros(1, 128), W2, relu=True)
    say("gcn two layers", schedule)
    A = DTCMatrix(a.rowptr, a.col, None, num_rows=M, num_cols=M)
    v = torch.from_numpy(a.values).cuda().requires_grad_(True)
    Xc, W1c, W2c = (t.cuda().requires_grad_(True) for t in (X, W1, W2))
    H = gcn_two_layer(v, Xc, W1c, W2c, A, schedule=schedule)
    H.backward(torch.from_numpy(dH).cuda())

    rows = torch.from_numpy(np.repeat(np.arange(M), np.diff(a.rowptr)))
    cols = torch.from_numpy(a.col.astype(np.int64))
    vd = torch.from_numpy(a.values.astype(np.float64)).requires_grad_(True)
    Xd, W1d, W2d = (t.double().requires_grad_(True) for t in (X, W1, W2))
    Ad = torch.zeros((M, M), dtype=torch.float64).index_put((rows, cols), vd)
    from dtc.gcn import gcn_layer
    with torch.no_grad():
        H1_gpu = gcn_layer(v, Xc, W1c, A, schedule=schedule).cpu()
    H1d = Ad @ (Xd @ W1d)
    mask = (H1_gpu > 0).double()
    assert float(((H1d.detach() > 0).double() != mask).double().mean()) < 2e-3
    Hd = Ad @ ((H1d * mask) @ W2d)
    (Hd * torch.from_numpy(dH.astype(np.float64))).sum().backward()

    def rel(x, ref):
        ref = ref.detach().numpy()
        return float(np.abs(np.asarray(x, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))
    assert H.dtype == torch.float32
    assert rel(H.detach().cpu().numpy(), Hd) < 3e-2
    assert rel(v.grad.cpu().numpy(), vd.grad) < 3e-2
    assert rel(Xc.grad.float().cpu().numpy(), Xd.grad) < 3e-2
    assert rel(W1c.grad.float().cpu().numpy(), W1d.grad) < 3e-2
    assert rel(W2c.grad.float().cpu().numpy(), W2d.grad) < 3e-2
