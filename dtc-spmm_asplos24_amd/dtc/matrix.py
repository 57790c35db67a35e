"""A sparse matrix held in ME-TCF on the GPU, with SpMM and its gradients.

``DTCMatrix`` builds the ME-TCF arrays once with ``DTCSpMM.preprocess_gpu``
(the same call sequence as scripts/DTCSpMM/run_DTC_SpMM.py:45-63) and keeps
the map from TC-block positions back to CSR edges, so CSR-ordered values can
be moved into TC-block order by a gather.  ``spmm`` runs either schedule.

``spmm_autograd(values, B, A)`` is differentiable in both ``values`` (CSR
order) and ``B``: the backward pass computes dB = A^T * dC with the ME-TCF of
the transpose (built on first use, on the GPU) and dvalues = SDDMM(dC, B) on
A's pattern.  The reference has no such backward: its backward names are
bound to the forwards (DTC-SpMM/DTCSpMM.cpp:384-399).
"""
from __future__ import annotations

import torch

from . import load_extension

BLK_H = 16
BLK_W = 8


class DTCMatrix:
    def __init__(self, rowptr, col, values=None, num_rows: int | None = None,
                 num_cols: int | None = None, device="cuda"):
        ext = load_extension()
        self.device = torch.device(device)
        self.rowptr = torch.as_tensor(rowptr, dtype=torch.int32).to(self.device).contiguous()
        self.col = torch.as_tensor(col, dtype=torch.int32).to(self.device).contiguous()
        self.num_rows = int(self.rowptr.numel() - 1 if num_rows is None else num_rows)
        if num_cols is None:
            num_cols = int(self.col.max().item()) + 1 if self.col.numel() else self.num_rows
        self.num_cols = int(num_cols)
        self.nnz = int(self.col.numel())
        self.values = None
        if values is not None:
            self.values = torch.as_tensor(values, dtype=torch.float32).to(self.device).contiguous()
        W = (self.num_rows + BLK_H - 1) // BLK_H
        self.num_windows = W
        self.block_partition = torch.zeros(W, dtype=torch.int32, device=self.device)
        self.edge_to_column = torch.zeros(self.nnz, dtype=torch.int32, device=self.device)
        self.edge_to_row = torch.zeros(self.nnz, dtype=torch.int32, device=self.device)
        # the values are reordered into TC-block order by the build itself (the
        # tile-id pass), not by a separate gather over tc_edge
        out = ext.preprocess_gpu(self.col, self.rowptr, self.num_rows, BLK_H, BLK_W,
                                 self.block_partition, self.edge_to_column, self.edge_to_row,
                                 values=self.values, return_edge_order=True)
        (self.row_window_offset, self.tcblock_rowid, self.tcblocktile_id, self.tcblock_offset,
         self.sparse_atox, self.block_count) = out[:6]
        self.values_tc = out[6] if self.values is not None else None
        self.tc_edge = out[-1]
        self._reset_caches()

    def _reset_caches(self) -> None:
        """Clears what is derived lazily from the ME-TCF arrays (also used
        when a matrix is rebuilt from a cache file without ``__init__``)."""
        self._transpose = None
        self._perm_t = None
        self._max_window_blocks = None

    # ------------------------------------------------------------------ info
    def sum_unique_columns(self) -> int:
        """sum_w U_w, from the block partition (exact: the last block of a
        window holds U_w mod 8 real columns, found from SparseAToX padding)."""
        atox = self.sparse_atox.view(-1, BLK_W)
        rwo = self.row_window_offset
        nb = (rwo[1:] - rwo[:-1])
        last = rwo[1:][nb > 0] - 1
        last_rows = atox[last.long()]
        # a padding slot holds 0 and is never slot 0; a real column 0 can only
        # sit in slot 0 of a window's first block.
        real_last = 1 + (last_rows[:, 1:] != 0).sum(dim=1)
        full = (nb[nb > 0] - 1) * BLK_W
        return int((full + real_last).sum().item())

    # ----------------------------------------------------------------- SpMM
    @property
    def max_window_blocks(self) -> int:
        """TC blocks of the longest row window (computed once)."""
        if self._max_window_blocks is None:
            rwo = self.row_window_offset
            self._max_window_blocks = int((rwo[1:] - rwo[:-1]).max().item()) if rwo.numel() > 1 else 0
        return self._max_window_blocks

    def spmm(self, B: torch.Tensor, schedule: str = "auto", precision: str = "fp32",
             values_tc: torch.Tensor | None = None, use_values: bool = True, vec: int = 0,
             subtiles: int = 0, chunk_blocks: int = 0, persistent: bool = False,
             out: torch.Tensor | None = None, deterministic: bool = False,
             cut_span_blocks: int = 0) -> torch.Tensor:
        """C = A * B.  schedule "auto" runs the unbalanced schedule when no row
        window is longer than ``dtc.selector.AUTO_WINDOW_BLOCKS`` TC blocks
        and the balanced one otherwise; "windows" / "balanced" force one.
        ``deterministic`` (balanced schedule) merges windows cut across chunks
        in a fixed order, so repeated calls are bitwise identical; the
        unbalanced schedule is deterministic as it is.  ``cut_span_blocks``
        (balanced schedule, a multiple of the chunk; 0 = the chunk): windows
        of at most this many TC blocks are never cut, longer ones are cut at
        its multiples."""
        ext = load_extension()
        if schedule == "auto":
            from .selector import default_schedule
            schedule = default_schedule(self.max_window_blocks)
        vals = values_tc if values_tc is not None else (self.values_tc if use_values else None)
        idx = self.row_window_offset if schedule in ("windows", "unbalanced") else self.tcblock_rowid
        return ext.spmm(B.contiguous(), idx, self.tcblocktile_id, self.tcblock_offset,
                        self.sparse_atox, self.num_rows, schedule=schedule, values=vals,
                        precision=precision, vec=vec, subtiles=subtiles,
                        chunk_blocks=chunk_blocks, persistent=persistent, out=out,
                        deterministic=deterministic, cut_span_blocks=cut_span_blocks)

    # ------------------------------------------------------------- transpose
    def transpose(self) -> "DTCMatrix":
        """ME-TCF of A^T (built on the GPU once and cached).  Its values are
        set per call from A's CSR values through ``perm_t``."""
        if self._transpose is None:
            ext = load_extension()
            rpt, ct, perm = ext.csr_transpose(self.rowptr, self.col, self.num_rows, self.num_cols)
            self._perm_t = perm
            self._transpose = DTCMatrix(rpt, ct, None, num_rows=self.num_cols,
                                        num_cols=self.num_rows, device=self.device)
        return self._transpose

    @property
    def perm_t(self) -> torch.Tensor:
        self.transpose()
        return self._perm_t

    def transposed_operands(self, schedule: str = "windows", values: torch.Tensor | None = None) -> tuple:
        """A^T's ME-TCF operands as the reference's backward names take them in
        ``transposed`` (``backward_ptx_uint8_improved``, ``..._gcn``: schedule
        "windows"; ``balance_backward_ptx_uint8_prefetch``: "balanced"), so that
        they return dB = A^T dY: (Rowwindow_offset or TCblock_rowid,
        TCblocktile_id, TCblock_offset, sparse_AToX_idx, A^T's values in its
        TC-block order or None for all ones, rows of A^T).  ``values``: A's CSR
        values (default: this matrix's)."""
        ext = load_extension()
        At = self.transpose()
        vals = values if values is not None else self.values
        vals_t = None if vals is None else ext.gather(ext.gather(vals.contiguous(), self.perm_t), At.tc_edge)
        index = At.row_window_offset if schedule in ("windows", "unbalanced") else At.tcblock_rowid
        return (index, At.tcblocktile_id, At.tcblock_offset, At.sparse_atox, vals_t, At.num_rows)


class _SpMMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, B, A: DTCMatrix, schedule, precision):
        ext = load_extension()
        vals_tc = ext.gather(values.contiguous(), A.tc_edge)
        C = A.spmm(B, schedule=schedule, precision=precision, values_tc=vals_tc)
        ctx.save_for_backward(values, B)
        ctx.A = A
        ctx.schedule = schedule
        ctx.precision = precision
        return C

    @staticmethod
    def backward(ctx, dC):
        ext = load_extension()
        values, B = ctx.saved_tensors
        A: DTCMatrix = ctx.A
        dC = dC.contiguous()
        d_values = d_B = None
        if ctx.needs_input_grad[0]:
            d_values = ext.sddmm_csr(A.rowptr, A.col, dC, B.contiguous(), A.edge_to_row)
        if ctx.needs_input_grad[1]:
            At = A.transpose()
            vals_t_csr = ext.gather(values.contiguous(), A.perm_t)
            vals_t_tc = ext.gather(vals_t_csr, At.tc_edge)
            d_B = At.spmm(dC, schedule=ctx.schedule, precision=ctx.precision, values_tc=vals_t_tc)
            # A^T has num_cols rows; B may have more (rows no column of A
            # reaches get zero gradient) or fewer (columns of A past B's rows
            # read as zero in the forward, so their rows of A^T dC are dropped).
            kb = B.shape[0]
            if d_B.shape[0] < kb:
                d_B = torch.cat([d_B, d_B.new_zeros((kb - d_B.shape[0], d_B.shape[1]))])
            elif d_B.shape[0] > kb:
                d_B = d_B[:kb].contiguous()
        return d_values, d_B, None, None, None


def spmm_autograd(values: torch.Tensor, B: torch.Tensor, A: DTCMatrix, schedule: str = "auto",
                  precision: str = "fp32") -> torch.Tensor:
    """C = A(values) * B, differentiable in values (CSR order) and B."""
    return _SpMMFunction.apply(values, B, A, schedule, precision)
