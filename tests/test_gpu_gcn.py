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
