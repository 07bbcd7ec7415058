"""Differentiable GEMM: a torch.autograd.Function over the HIP kernels.

Forward Y = X @ W.  Backward uses the same gemm_kernel with its
transposed-operand tile loaders, so no transposed copy is ever materialised:

    dX = dY @ W^T   -> gemm(dY, W, trans_b=True)   (W read as its N x K transpose)
    dW = X^T @ dY   -> gemm(X, dY, trans_a=True)   (X read as its K x M transpose)

The reference is row-major, forward-only (thesis p45); this wrapper has no
counterpart there.  FP32 storage with mode "bf16x3" (default) or "fp32"
compute; float64 tensors use mode "fp64".
"""
from __future__ import annotations

import torch

from .gemm import gemm


def _mode_for(t: torch.Tensor, mode: str | None) -> str:
    if t.dtype == torch.float64:
        return "fp64"
    if t.dtype != torch.float32:
        raise TypeError(f"CosmaMatmul takes float32 or float64 tensors, got {t.dtype}")
    return mode or "bf16x3"


class CosmaMatmul(torch.autograd.Function):
    """Y = X @ W with hand-launched backward GEMMs."""

    @staticmethod
    def forward(ctx, X: torch.Tensor, W: torch.Tensor, mode: str | None = None):
        m = _mode_for(X, mode)
        if W.dtype != X.dtype:
            raise TypeError("X and W must share a dtype")
        Xc, Wc = X.contiguous(), W.contiguous()
        ctx.save_for_backward(Xc, Wc)
        ctx.mode = m
        return gemm(Xc, Wc, mode=m)

    @staticmethod
    def backward(ctx, dY: torch.Tensor):
        X, W = ctx.saved_tensors
        dY = dY.contiguous()
        dX = dW = None
        if ctx.needs_input_grad[0]:
            dX = gemm(dY, W, mode=ctx.mode, trans_b=True)
        if ctx.needs_input_grad[1]:
            dW = gemm(X, dY, mode=ctx.mode, trans_a=True)
        return dX, dW, None


def matmul(X: torch.Tensor, W: torch.Tensor, mode: str | None = None) -> torch.Tensor:
    """Differentiable X @ W on the cosma kernels (2-D operands)."""
    if X.dim() != 2 or W.dim() != 2:
        raise ValueError("matmul takes 2-D operands")
    return CosmaMatmul.apply(X, W, mode)
