"""Thin torch.autograd.Function over the C ABI (test harness for the adjoints).

C = A·B with A a value-carrying bitmap (trainable values in the bitmap's
(row, column) order) and B dense.  Backward uses the hand-written adjoints:

* dB      = Aᵀ·dC      -> bws_transpose_values + bws_spmm_tA (a bitmap SpMM on Aᵀ)
* dvalues = (dC·Bᵀ)⊙A  -> bws_sddmm (only the set bits are computed and stored)

Every call runs libbws.so kernels on the current stream; there is no PyTorch
fallback.
"""
from __future__ import annotations

import torch

from . import Bitmap, sddmm, spmm, spmm_tA, transpose


class BitmapSpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, B, words, rows, cols, row_ptr, precision="fp32"):
        A = Bitmap(words, rows, cols, row_ptr, values)
        C = spmm(A, B, precision)
        ctx.save_for_backward(values, B, words, row_ptr)
        ctx.shape = (rows, cols)
        ctx.precision = precision
        return C

    @staticmethod
    def backward(ctx, dC):
        values, B, words, row_ptr = ctx.saved_tensors
        rows, cols = ctx.shape
        prec = ctx.precision
        dC = dC.contiguous()
        A = Bitmap(words, rows, cols, row_ptr, values)
        d_values = d_B = None
        if ctx.needs_input_grad[0]:
            Bx = B if B.dtype == dC.dtype else B.to(dC.dtype)
            d_values = sddmm(A, dC, Bx, prec)
        if ctx.needs_input_grad[1]:
            AT = transpose(A)
            d_B = spmm_tA(AT, dC, prec)
            if d_B.dtype != B.dtype:
                d_B = d_B.to(B.dtype)
        return d_values, d_B, None, None, None, None, None


def bitmap_spmm(A: Bitmap, values: torch.Tensor, B: torch.Tensor, precision: str = "fp32"):
    """Differentiable C = A(values)·B; A supplies the pattern and row_ptr."""
    rp = A.row_ptr
    if rp is None:
        from . import row_ptr as _row_ptr
        rp = _row_ptr(A)
    return BitmapSpMM.apply(values, B, A.words, A.rows, A.cols, rp, precision)
