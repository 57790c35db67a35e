"""GCN-style layer H = A * (X W) over ME-TCF, differentiable in A's values,
X and W.

X W is a plain dense GEMM (torch / hipBLASLt, bf16 operands, float32
accumulation, bf16 result); its bf16 output is gathered by the SpMM as is
(b_dtype bfloat16: half the bytes of a float32 B), with A's values split into
three bf16 limbs and float32 accumulation.  The backward pass:
    dXW     = A^T dH            (ME-TCF of A^T; dH gathered in bf16 for a bf16 layer)
    dX      = dXW W^T,  dW = X^T dXW   (dense GEMMs; dW split over K)
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


class _GCNFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, X, W, A: DTCMatrix, schedule):
        ext = load_extension()
        XW = torch.matmul(X, W).contiguous()
        vals_tc = ext.gather(values.contiguous(), A.tc_edge)
        H = A.spmm(XW, schedule=schedule, precision="bf16x6", values_tc=vals_tc)
        ctx.save_for_backward(values, X, W, XW)
        ctx.A = A
        ctx.schedule = schedule
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
            # A^T dH feeds bf16 GEMMs: gather dH in X's precision (bf16 B
            # kernel, float32 accumulation), then round the result once
            dHx = dH.to(X.dtype) if X.dtype == torch.bfloat16 else dH
            dXW = At.spmm(dHx, schedule=ctx.schedule, values_tc=vals_t_tc)
            dXW = dXW.to(X.dtype)
            if ctx.needs_input_grad[1]:
                d_X = torch.matmul(dXW, W.t().contiguous())
            if ctx.needs_input_grad[2]:
                d_W = _xt_dy(X, dXW)
        return d_values, d_X, d_W, None, None


def gcn_layer(values: torch.Tensor, X: torch.Tensor, W: torch.Tensor, A: DTCMatrix,
              schedule: str = "balanced") -> torch.Tensor:
    """H = A(values) * (X W); X, W bf16 (or float32), H float32."""
    return _GCNFunction.apply(values, X, W, A, schedule)
