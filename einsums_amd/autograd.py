"""torch.autograd for the einsum path.

``einsum(C_indices, A_indices, A, B_indices, B, alpha=1.0)`` returns a new tensor
``C = alpha * contraction(A, B)`` and records a ``torch.autograd.Function`` whose backward
calls the backend's adjoint kernels (ein_einsum_adjoint_*): the cotangent of A is the
contraction of dC with B, labelled like A (a GEMM, GEMV, ... or generic kernel chosen by the
same dispatch); repeated labels of A give a diagonal scatter without atomics. The
reference EinsumsPy has no autograd; this layer is built on its compile_plan surface.
"""
from __future__ import annotations

import ctypes

import torch

from . import core
from ._lib import check, lib, make_desc


def _out_shape(C_indices: str, A_indices: str, A: torch.Tensor, B_indices: str, B: torch.Tensor):
    ext = {}
    for idx, T in ((A_indices, A), (B_indices, B)):
        for ch, n in zip(idx, T.shape):
            if ext.setdefault(ch, n) != n:
                raise core.DimensionError(f"index {ch!r} has extents {ext[ch]} and {n}")
    missing = [ch for ch in C_indices if ch not in ext]
    if missing:
        raise ValueError(f"output labels {missing} appear in neither input")
    return [ext[ch] for ch in C_indices]


def _adjoint(wrt: int, alpha: float, C_indices, dC, A_indices, A, B_indices, B, grad_like) -> torch.Tensor:
    G = torch.empty_like(grad_like, memory_format=torch.contiguous_format)
    X_indices = A_indices if wrt == 0 else B_indices
    cd, ad, bd = make_desc(C_indices, dC.shape, dC.stride()), make_desc(A_indices, A.shape, A.stride()), \
        make_desc(B_indices, B.shape, B.stride())
    gd = make_desc(X_indices, G.shape, G.stride())
    used = ctypes.c_int32(-1)
    fn = lib.ein_einsum_adjoint_f64 if G.dtype == torch.float64 else lib.ein_einsum_adjoint_f32
    check(fn(wrt, float(alpha), ctypes.byref(cd), ctypes.c_void_p(dC.data_ptr()), ctypes.byref(ad),
             ctypes.c_void_p(A.data_ptr()), ctypes.byref(bd), ctypes.c_void_p(B.data_ptr()), 0.0, ctypes.byref(gd),
             ctypes.c_void_p(G.data_ptr()), ctypes.byref(used), core.stream_handle()), "einsum adjoint")
    return G


class EinsumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C_indices, A_indices, B_indices, alpha, A, B):
        C = torch.empty(_out_shape(C_indices, A_indices, A, B_indices, B), dtype=A.dtype, device=A.device)
        core.einsum(C_indices, C, A_indices, A, B_indices, B, C_prefactor=0.0, AB_prefactor=alpha)
        ctx.save_for_backward(A, B)
        ctx.idx = (C_indices, A_indices, B_indices)
        ctx.alpha = alpha
        return C

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        C_indices, A_indices, B_indices = ctx.idx
        dC = dC.contiguous() if dC.dim() > 0 else dC
        gA = gB = None
        if ctx.needs_input_grad[4]:
            gA = _adjoint(0, ctx.alpha, C_indices, dC, A_indices, A, B_indices, B, A)
        if ctx.needs_input_grad[5]:
            gB = _adjoint(1, ctx.alpha, C_indices, dC, A_indices, A, B_indices, B, B)
        return None, None, None, None, gA, gB


def einsum(C_indices: str, A_indices: str, A: torch.Tensor, B_indices: str, B: torch.Tensor,
           alpha: float = 1.0) -> torch.Tensor:
    """Differentiable ``C[C_indices] = alpha * sum A[A_indices] B[B_indices]`` on the GPU."""
    if A.dtype != B.dtype or A.dtype not in (torch.float64, torch.float32):
        raise TypeError("einsum: A and B must share float64 or float32")
    return EinsumFunction.apply(C_indices, A_indices, B_indices, float(alpha), A, B)
