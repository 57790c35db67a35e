"""GCN-style layer H = A * (act(X) W) over ME-TCF, differentiable in A's
values, X and W.

With bf16 W (K a multiple of 32, F of 64) and an activation or cast to fuse,
act(X) W is libdtc's GCN transform (csrc/gcn.hip): one MFMA kernel that
applies act -- the previous layer's ReLU (``input_relu``) -- and the cast of
a float32 X to bf16 while it loads X, accumulates in float32 and writes the
bf16 B that the SpMM gathers as is (b_dtype bfloat16: half the bytes of a
float32 B), with A's values split into three bf16 limbs and float32
accumulation.  A plain bf16 x bf16 product, other shapes and float32 W use a
torch GEMM.  The backward pass:
    dXW     = A^T dH            (ME-TCF of A^T; dH gathered in bf16 for a bf16 layer)
    dX      = (dXW W^T) * [X > 0 if input_relu],  dW = act(X)^T dXW
                                (dense GEMMs; dW split over K)
    dvalues = SDDMM(dH, XW)     on A's pattern, XW rows gathered as bf16.
The reference's GCN entry (`DTCSpMM_gcn`, DTCSpMM_kernel.cu:1033-1116) is a
forward only and its "backward" names return A * dY (DTCSpMM.cpp:384-399).
"""
from __future__ import annotations

import torch

from . import load_extension
from .matrix import DTCMatrix


def _xt_dy(X: torch.Tensor, dY: torch.Tensor, parts: int = 64) -> torch.Tensor:
    """X^T dY for tall X [K x F] and dY [K x G]: the product has only F x G
    outputs (one or two GEMM tiles at 256 x 256), so K is split into `parts`
    batches that run side by side and their partial products are summed in
    float32.  Both bmm operands are made contiguous first."""
    K = X.shape[0]
    if K % parts or K // parts < 512:
        return torch.matmul(X.t().contiguous(), dY.contiguous())
    Xs = X.reshape(parts, K // parts, X.shape[1]).transpose(1, 2).contiguous()
    Ys = dY.contiguous().reshape(parts, K // parts, dY.shape[1])
    return torch.bmm(Xs, Ys).float().sum(0).to(X.dtype)


def fused_transform_applies(X: torch.Tensor, W: torch.Tensor, relu: bool = True) -> bool:
    """Whether act(X) W runs on libdtc's GCN transform: bf16 W of a shape it
    takes, and something to fuse -- the activation, or the cast of a float32
    X.  A plain bf16 x bf16 product is a library GEMM (hipBLASLt, 64 vs 70 us
    for X 262,144 x 256 at F = 256; the fused ReLU + cast of a float32 X takes
    94 us against 227 us for torch's ReLU, cast and GEMM)."""
    return (W.dtype == torch.bfloat16 and X.dtype in (torch.float32, torch.bfloat16)
            and W.shape[0] % 32 == 0 and W.shape[1] % 64 == 0
            and (relu or X.dtype == torch.float32))


def _act(X: torch.Tensor, relu: bool, dtype) -> torch.Tensor:
    return (torch.relu(X) if relu else X).to(dtype)


def transform(X: torch.Tensor, W: torch.Tensor, relu: bool = False) -> torch.Tensor:
    """act(X) W, act = ReLU when relu: bf16 output from the GCN transform
    kernel for bf16 W, else a torch GEMM in W's dtype."""
    if fused_transform_applies(X, W, relu):
        ext = load_extension()
        return ext.gcn_transform(X.contiguous(), ext.gcn_pack_weights(W.contiguous()), W.shape[1], relu)
    return torch.matmul(_act(X, relu, W.dtype), W).contiguous()


class _GCNFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, X, W, A: DTCMatrix, schedule, input_relu):
        ext = load_extension()
        XW = transform(X, W, input_relu)
        vals_tc = ext.gather(values.contiguous(), A.tc_edge)
        H = A.spmm(XW, schedule=schedule, precision="bf16x6", values_tc=vals_tc)
        ctx.save_for_backward(values, X, W, XW)
        ctx.A = A
        ctx.schedule = schedule
        ctx.input_relu = input_relu
        return H

    @staticmethod
    def backward(ctx, dH):
        ext = load_extension()
        values, X, W, XW = ctx.saved_tensors
        A: DTCMatrix = ctx.A
        dH = dH.contiguous().float()
        d_values = d_X = d_W = None
        if ctx.needs_input_grad[0]:
            # XW is gathered as stored (bf16 rows: half the bytes), exactly
            d_values = ext.sddmm_csr(A.rowptr, A.col, dH, XW, A.edge_to_row)
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            At = A.transpose()
            vals_t_tc = ext.gather(ext.gather(values.contiguous(), A.perm_t), At.tc_edge)
            # A^T dH feeds GEMMs in W's precision: gather dH in it (bf16 B
            # kernel, float32 accumulation), then round the result once
            dHx = dH.to(W.dtype) if W.dtype == torch.bfloat16 else dH
            dXW = At.spmm(dHx, schedule=ctx.schedule, values_tc=vals_t_tc).to(W.dtype)
            if ctx.needs_input_grad[1]:
                d_X = torch.matmul(dXW, W.t().contiguous())
                if ctx.input_relu:
                    d_X = d_X * (X > 0)
                d_X = d_X.to(X.dtype)
            if ctx.needs_input_grad[2]:
                d_W = _xt_dy(_act(X, ctx.input_relu, W.dtype), dXW)
        return d_values, d_X, d_W, None, None, None


def gcn_layer(values: torch.Tensor, X: torch.Tensor, W: torch.Tensor, A: DTCMatrix,
              schedule: str = "balanced", input_relu: bool = False) -> torch.Tensor:
    """H = A(values) * (act(X) W), act = ReLU when input_relu (the previous
    layer's activation, applied inside the transform); X bf16 or float32, W
    bf16 (or float32), H float32."""
    return _GCNFunction.apply(values, X, W, A, schedule, input_relu)


def gcn_two_layer(values: torch.Tensor, X: torch.Tensor, W1: torch.Tensor, W2: torch.Tensor,
                  A: DTCMatrix, schedule: str = "balanced") -> torch.Tensor:
    """H2 = A (relu(A (X W1)) W2): the first layer's float32 output goes into
    the second layer's transform, which applies the ReLU and the bf16 cast as
    it loads it."""
    H1 = gcn_layer(values, X, W1, A, schedule=schedule)
    return gcn_layer(values, H1, W2, A, schedule=schedule, input_relu=True)
