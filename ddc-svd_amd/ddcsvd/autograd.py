"""torch.autograd harness around svd_gpu_dev / svd_gpu_backward.

    sigma, U, V = ddcsvd.autograd.svd(A)              # A: CUDA float64 tensor
    loss(sigma, U, V).backward()                      # A.grad from the GPU adjoint

The forward is the library's FP64 SVD (singular values ascending, thin U and
V with A = U diag(sigma) V^T); the backward is the library's adjoint kernel.
recompute=True keeps only A between forward and backward and runs the SVD
again in the backward (the library's SVD is deterministic, so the
recomputed factors are the stored ones bit for bit); otherwise U, sigma and
V are kept.  Either way no Householder reflectors outlive the forward call.
"""
from __future__ import annotations

import torch

from . import svd_backward_torch, svd_torch


class SvdGpuFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, recompute: bool = False, gap_guard: float = 1e-13):
        if A.dtype != torch.float64 or not A.is_cuda:
            raise ValueError("svd_gpu needs a float64 CUDA tensor")
        sigma, U, V = svd_torch(A.detach())
        ctx.recompute = bool(recompute)
        ctx.gap_guard = float(gap_guard)
        ctx.set_materialize_grads(False)
        if ctx.recompute:
            ctx.save_for_backward(A)
        else:
            ctx.save_for_backward(sigma, U, V)
        return sigma, U, V

    @staticmethod
    def backward(ctx, gsigma, gU, gV):
        if gsigma is None and gU is None and gV is None:
            return None, None, None
        if ctx.recompute:
            (A,) = ctx.saved_tensors
            sigma, U, V = svd_torch(A.detach())
        else:
            sigma, U, V = ctx.saved_tensors
        gA = svd_backward_torch(U, sigma, V, gU, gsigma, gV, ctx.gap_guard)
        return gA, None, None


def svd(A, recompute: bool = False, gap_guard: float = 1e-13):
    """Differentiable thin SVD on the GPU: returns (sigma ascending, U, V)."""
    return SvdGpuFunction.apply(A, recompute, gap_guard)
