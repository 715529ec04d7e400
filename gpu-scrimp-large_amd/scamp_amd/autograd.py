"""Differentiable batched self-join (torch).

``MatrixProfile.apply(T, m)`` takes a batch of series ``T`` (torch tensor
[B, n] on the GPU, float32 or float64) and returns the 1NN_INDEX matrix
profile as Pearson correlations ``P`` [B, n-m+1] (same dtype as T; NaN where
a window has no match) and the indices ``I`` [B, n-m+1] (int32, -1 = none).

Forward: the plain FP64 self-join of every series (statistics, exact seeds,
diagonal sweep) in one batched launch through the C ABI, whose persistent
sweep takes (series, band) work items, so a batch of short series fills the
GPU as one long series does.  Backward: the HIP
kernel ``scamp_hip_pearson_backward``, which reuses the forward's indices
(no join is recomputed) and scatters dP_i/dT into windows i and I_i:
    dr/dx = (y^ - r x^) / |x~|,   dr/dy = (x^ - r y^) / |y~|.
The indices are piecewise constant in T and carry no gradient.

Both run on torch's current stream, through an engine context of their own
bound to it (scamp_hip_ctx_set_stream): the engine's kernels are queued
behind torch's and torch's behind them, with no host synchronisation.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C

_INIT = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))

# device -> (engine context, the torch stream it is bound to)
_BOUND = {}


def _engine(device: int) -> C.Context:
    """This module's engine context on `device`, bound to torch's current
    stream there (rebound when the current stream changes)."""
    ctx, bound = _BOUND.get(device, (None, None))
    if ctx is None:
        ctx = C.Context(device)
    stream = torch.cuda.current_stream(device).cuda_stream
    if stream != bound:
        ctx.set_stream(stream)
    _BOUND[device] = (ctx, stream)
    return ctx


def _forward(t64: torch.Tensor, m: int, device: int):
    B, n = t64.shape
    ns = n - m + 1
    prof = torch.full((B, ns), _INIT, dtype=torch.int64, device=t64.device)
    C.selfjoin_batch_device(t64.data_ptr(), n, B, C.make_params(m, precision="double"),
                            prof.data_ptr(), device=device, ctx=_engine(device))
    # mp_entry: the float32 correlation in the low word, the int32 index in
    # the high word (views, no copies)
    corr = prof.view(torch.float32).view(B, ns, 2)[..., 0]
    idx = prof.view(torch.int32).view(B, ns, 2)[..., 1]
    return corr, idx


class MatrixProfile(torch.autograd.Function):
    @staticmethod
    def forward(ctx, T: torch.Tensor, m: int):
        if not T.is_cuda:
            raise ValueError("MatrixProfile needs a GPU tensor")
        ctx.in_shape = T.shape
        if T.dim() == 1:
            T = T.unsqueeze(0)
        device = T.device.index or 0
        t64 = T.detach().to(torch.float64).contiguous()
        P, I = _forward(t64, m, device)
        empty = P < -1
        I = torch.where(empty, -1, I)
        P = torch.where(empty, float("nan"), P)
        ctx.save_for_backward(t64, I)
        ctx.m = m
        ctx.device = device
        ctx.in_dtype = T.dtype
        ctx.mark_non_differentiable(I)
        return P.to(T.dtype), I

    @staticmethod
    def backward(ctx, grad_p, grad_i=None):
        t64, I = ctx.saved_tensors
        B, n = t64.shape
        g = torch.nan_to_num(grad_p.to(torch.float64), nan=0.0).contiguous()
        grad_t = torch.empty_like(t64)
        lib = C.load()
        cx = _engine(ctx.device)
        C.check(lib.scamp_hip_pearson_backward(
            cx.handle, ctypes.c_void_p(t64.data_ptr()), n, ctx.m, B,
            ctypes.c_void_p(I.contiguous().data_ptr()), ctypes.c_void_p(g.data_ptr()),
            ctypes.c_void_p(grad_t.data_ptr())))
        return grad_t.to(ctx.in_dtype).reshape(ctx.in_shape), None


def matrix_profile(T: torch.Tensor, m: int):
    """(P, I) of each row of T ([B, n], or [n] -> [1, n-m+1] outputs);
    differentiable in T through P."""
    return MatrixProfile.apply(T, m)
