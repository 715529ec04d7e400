"""The differentiable batched join (scamp_amd.autograd): forward against the
oracle, backward (the HIP kernel reusing the forward's indices) against
central finite differences and against FP64 autograd through a brute-force
correlation matrix (CPU torch), with the index held fixed in both.

Bars: <= 1e-6 relative to both references.  The finite differences use
h = 1e-6 on the loss terms that depend on the perturbed sample (the others
have zero derivative and would only add rounding noise), evaluated in
extended precision (numpy longdouble, 64-bit significand), so their own
error stays near 1e-10 and the check is |g - fd| <= 1e-6 |fd| + 1e-9.  The
brute-force check is rtol 1e-6 with atol 1e-9 * max|ref| for entries near
zero."""
import numpy as np
import pytest
import torch

from helpers import check_1nn_index, random_walk
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FD_H = 1e-6


def fd_close(g, fd):
    return abs(g - fd) <= 1e-6 * abs(fd) + 1e-9


def _local_loss(T, I, m, rows):
    """sum over `rows` of (1 - corr(T[r:r+m], T[I_r:I_r+m])) in longdouble."""
    from numpy.lib.stride_tricks import sliding_window_view as sw
    W = sw(T, m)
    A = W[rows].astype(np.longdouble)
    B = W[I[rows]].astype(np.longdouble)
    A = A - A.mean(1, keepdims=True)
    B = B - B.mean(1, keepdims=True)
    r = (A * B).sum(1) / np.sqrt((A * A).sum(1) * (B * B).sum(1))
    return np.sum(1 - r)


def fd_grad(X, I, m, t, h=FD_H):
    """Central difference of L = sum_i (1 - P_i) (indices fixed) w.r.t. X[t]."""
    ok = I >= 0
    r = np.arange(I.size)
    rows = np.nonzero(ok & (((r <= t) & (t < r + m)) | ((I <= t) & (t < I + m))))[0]
    Xl = X.astype(np.longdouble)
    Xp, Xm = Xl.copy(), Xl.copy()
    Xp[t] += np.longdouble(h)
    Xm[t] -= np.longdouble(h)
    return float((_local_loss(Xp, I, m, rows) - _local_loss(Xm, I, m, rows))
                 / (2 * np.longdouble(h)))


def brute_force_grad(T, I, m):
    """FP64 autograd through the full window-correlation matrix (CPU)."""
    t = torch.tensor(T, dtype=torch.float64, requires_grad=True)
    W = t.unfold(0, m, 1)
    Wc = W - W.mean(1, keepdim=True)
    Wn = Wc / Wc.norm(dim=1, keepdim=True)
    D = Wn @ Wn.T
    idx = torch.tensor(I, dtype=torch.int64)
    ok = idx >= 0
    rows = torch.nonzero(ok).squeeze(1)
    P = D[rows, idx[rows]]
    (1 - P).sum().backward()
    return t.grad.numpy()


def test_forward_matches_oracle_and_batch_rows():
    from scamp_amd.autograd import matrix_profile
    B, n, m = 4, 2048, 128
    X = np.stack([random_walk(n, 100 + b) for b in range(B)])
    T = torch.tensor(X, dtype=torch.float32, device="cuda")
    P, I = matrix_profile(T, m)
    X32 = T.double().cpu().numpy()
    for b in range(B):
        oc, oi = O.join(X32[b], None, m)
        res = check_1nn_index(X32[b], None, m, P[b].cpu().numpy(), I[b].cpu().numpy(),
                              oc, oi)
        assert res == {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


def test_backward_matches_finite_differences():
    from scamp_amd.autograd import matrix_profile
    n, m = 1024, 64
    X = random_walk(n, 7)
    T = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
    P, I = matrix_profile(T, m)
    loss = (1 - torch.nan_to_num(P, nan=1.0)).sum()
    loss.backward()
    g = T.grad.detach().cpu().numpy()
    Ih = I.cpu().numpy()[0]
    rng = np.random.default_rng(0)
    for t in rng.choice(n, 40, replace=False):
        fd = fd_grad(X, Ih, m, t)
        assert fd_close(g[t], fd), (t, g[t], fd)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_backward_matches_bruteforce_autograd_config4_shape(dtype):
    """64 float32 random walks of n = 8192 (seeds 100..163), m = 128: forward
    and backward of L = sum_i (1 - P_i); gradients of 4 of the series checked
    against FP64 autograd through the brute-force correlation matrix at rtol
    1e-6.  The gradient is computed in FP64 whatever the input dtype; a
    float32 input gets it rounded to float32 (2^-24 relative), which the bar
    covers."""
    from scamp_amd.autograd import matrix_profile
    B, n, m = 64, 8192, 128
    X = np.stack([np.cumsum(np.random.Generator(np.random.PCG64(100 + b))
                            .standard_normal(n)).astype(np.float32) for b in range(B)])
    T = torch.tensor(X, device="cuda", dtype=dtype, requires_grad=True)
    P, I = matrix_profile(T, m)
    (1 - torch.nan_to_num(P, nan=1.0)).sum().backward()
    G = T.grad.detach().cpu().numpy().astype(np.float64)
    assert G.shape == (B, n) and np.isfinite(G).all()
    worst = 0.0
    for b in (0, 17, 40, 63):
        ref = brute_force_grad(X[b].astype(np.float64), I[b].cpu().numpy(), m)
        atol = 1e-9 * np.abs(ref).max()
        np.testing.assert_allclose(G[b], ref, rtol=1e-6, atol=atol)
        big = np.abs(ref) > 1e3 * atol
        worst = max(worst, float(np.max(np.abs(G[b][big] - ref[big]) / np.abs(ref[big]))))
    print(f"backward {dtype}: max relative error vs brute-force FP64 autograd {worst:.3g}",
          flush=True)


def test_backward_is_deterministic_and_matches_per_series():
    """Every sample's gradient is gathered by one thread in a fixed order (no
    atomics): two backward passes give the same bits, and a batch gives each
    series exactly what that series alone gives."""
    from scamp_amd.autograd import matrix_profile
    B, n, m = 6, 3000, 64
    X = np.stack([random_walk(n, 40 + b, nan_runs=b % 2) for b in range(B)])
    grads = []
    for _ in range(2):
        T = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
        P, _ = matrix_profile(T, m)
        (1 - torch.nan_to_num(P, nan=1.0)).sum().backward()
        grads.append(T.grad.detach().cpu().numpy())
    assert np.array_equal(grads[0], grads[1])
    for b in (0, 3, 5):
        T1 = torch.tensor(X[b:b + 1], dtype=torch.float64, device="cuda", requires_grad=True)
        P1, _ = matrix_profile(T1, m)
        (1 - torch.nan_to_num(P1, nan=1.0)).sum().backward()
        assert np.array_equal(T1.grad.detach().cpu().numpy()[0], grads[0][b])


@pytest.mark.parametrize("n,m", [
    (8255, 64),    # 8192 windows: the largest per-series sort in one workgroup
    (10000, 64),   # 9937 windows: the device-wide sort of all entries
])
def test_backward_both_sort_paths(n, m):
    """Series up to 8192 windows sort their entries per series in LDS, longer
    ones through one device-wide sort; both give finite-difference gradients,
    the same bits on every call, and per series what the series alone gives."""
    from scamp_amd.autograd import matrix_profile
    B = 2
    X = np.stack([random_walk(n, 70 + b) for b in range(B)])
    grads = []
    for _ in range(2):
        T = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
        P, I = matrix_profile(T, m)
        (1 - torch.nan_to_num(P, nan=1.0)).sum().backward()
        grads.append(T.grad.detach().cpu().numpy())
    assert np.array_equal(grads[0], grads[1])
    T1 = torch.tensor(X[1:2], dtype=torch.float64, device="cuda", requires_grad=True)
    P1, _ = matrix_profile(T1, m)
    (1 - torch.nan_to_num(P1, nan=1.0)).sum().backward()
    assert np.array_equal(T1.grad.detach().cpu().numpy()[0], grads[0][1])
    Ih = I.cpu().numpy()[0]
    rng = np.random.default_rng(1)
    for t in rng.choice(n, 12, replace=False):
        fd = fd_grad(X[0], Ih, m, t)
        assert fd_close(grads[0][0, t], fd), (t, grads[0][0, t], fd)
