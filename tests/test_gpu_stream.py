"""Contexts bound to a caller's stream (scamp_hip_ctx_set_stream): the
device-buffer calls queue their work behind the caller's and return at once,
and must give exactly what the blocking calls give.  The inputs below are
produced by torch work still in flight on the same stream when the engine's
call is made (no host synchronisation in between), and the outputs are read
by torch work queued after it."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import random_walk
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu


def _series(B, n, seed):
    return np.stack([random_walk(n, seed + b, nan_runs=b % 2) for b in range(B)])


def _blocking_batch(X, m):
    B, n = X.shape
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    init = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))
    out = torch.full((B, n - m + 1), init, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    C.selfjoin_batch_device(T.data_ptr(), n, B, C.make_params(m), out.data_ptr())
    return out.cpu().numpy()


def test_bound_context_batch_and_join_match_blocking():
    B, n, m = 6, 3000, 64
    X = _series(B, n, 300)
    ref = _blocking_batch(X, m)
    ctx = C.Context(0)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    init = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))
    for rep in range(3):
        # the series is still being computed on the stream when the join is
        # queued; the profile is read by a copy queued after it
        T = torch.tensor(X * 0.5, dtype=torch.float64, device="cuda") * 2.0
        out = torch.full((B, n - m + 1), init, dtype=torch.int64, device="cuda")
        C.selfjoin_batch_device(T.data_ptr(), n, B, C.make_params(m), out.data_ptr(), ctx=ctx)
        got = out.clone().cpu().numpy()
        assert np.array_equal(got, ref), rep
        t = ctx.timing()
        assert t["total_ms"] > 0 and 0 < t["sweep_ms"] <= t["total_ms"]
    # a single join on the bound context against the blocking one
    T0 = torch.tensor(X[0], dtype=torch.float64, device="cuda")
    one = torch.full((n - m + 1,), init, dtype=torch.int64, device="cuda")
    C.join_device((T0 * 1.0).data_ptr(), n, 0, 0, C.make_params(m), one.data_ptr(), ctx=ctx)
    got_one = one.cpu().numpy()
    blk = torch.full((n - m + 1,), init, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    C.join_device(T0.data_ptr(), n, 0, 0, C.make_params(m), blk.data_ptr())
    assert np.array_equal(got_one, blk.cpu().numpy())
    # back to the context's own stream: blocking again, same results
    ctx.set_stream(None)
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    out = torch.full((B, n - m + 1), init, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    C.selfjoin_batch_device(T.data_ptr(), n, B, C.make_params(m), out.data_ptr(), ctx=ctx)
    assert np.array_equal(out.cpu().numpy(), ref)
    ctx.close()


def test_bound_context_backward_matches_blocking():
    B, n, m = 4, 2500, 50
    X = _series(B, n, 320)
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    init = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))
    prof = torch.full((B, n - m + 1), init, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    C.selfjoin_batch_device(T.data_ptr(), n, B, C.make_params(m), prof.data_ptr())
    idx = prof.view(torch.int32).view(B, n - m + 1, 2)[..., 1].contiguous()
    g = torch.tensor(np.random.default_rng(5).standard_normal((B, n - m + 1)),
                     device="cuda")
    lib = C.load()

    def backward(ctx, t, gg):
        out = torch.empty_like(t)
        C.check(lib.scamp_hip_pearson_backward(
            ctx.handle, ctypes.c_void_p(t.data_ptr()), n, m, B,
            ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(gg.data_ptr()),
            ctypes.c_void_p(out.data_ptr())))
        return out

    torch.cuda.synchronize()
    ref = backward(C.context(0), T, g).cpu().numpy()
    ctx = C.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        got = backward(ctx, T * 1.0, g * 1.0).clone().cpu().numpy()
        assert np.array_equal(got, ref), rep
    ctx.close()


def test_autograd_on_a_side_stream_matches_default_stream():
    """The differentiable join follows torch's current stream: on a side
    stream (the module rebinds its context) it gives the same bits."""
    from scamp_amd.autograd import matrix_profile
    B, n, m = 4, 3000, 64
    X = _series(B, n, 340)

    def step():
        T = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
        P, I = matrix_profile(T, m)
        (1 - torch.nan_to_num(P, nan=1.0)).sum().backward()
        return P.detach().cpu().numpy(), I.cpu().numpy(), T.grad.detach().cpu().numpy()

    ref = step()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        got = step()
    torch.cuda.synchronize()
    again = step()
    for a, b, c in zip(ref, got, again):
        assert np.array_equal(a, b, equal_nan=True)
        assert np.array_equal(a, c, equal_nan=True)
