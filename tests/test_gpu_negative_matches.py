"""Windows whose every correlation is negative, in the long-chunk FP64 sweep.

A = 2^21-sample random walk (PCG64 seed 41), B = a ramp t + N(0, 1) noise of
the same length (seed 42), m = 64, AB-join with keep_rows_separate (both
profiles, so each pass updates rows and columns), 1NN_INDEX profile of A.  A
strongly falling window of A correlates at about -0.9 with every (rising)
window of B, so its best match is negative; where that match lies at a B
index at or below the A index, it reaches the profile through the sweep's
column side (the upper triangle: columns = A, rows = B), whose threshold test
starts from "no match yet" (the AB-join has no warm start).  At this size the
passes run in long row chunks, i.e. in the two-stage filter kernel with the
integer image of the column threshold.

Checked at 24 seeded falling windows of A (correlation with a pure ramp below
-0.8): each window's exact FP64 correlation row over all of B (FFT sliding
dot products of the centred window divided by directly computed norms) has
a negative maximum, the stored correlation equals it to float32 rounding,
and the stored index is its argmax (or a partner within 1e-9 of it).
"""
import numpy as np
import pytest

from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

N, M, WINDOWS = 1 << 21, 64, 24


def _series():
    a = np.cumsum(np.random.Generator(np.random.PCG64(41)).standard_normal(N))
    b = np.arange(N, dtype=np.float64) + np.random.Generator(
        np.random.PCG64(42)).standard_normal(N)
    return a, b


def _norms(y, m):
    from numpy.lib.stride_tricks import sliding_window_view as sw
    W = sw(y, m)
    out = np.empty(W.shape[0])
    for s in range(0, W.shape[0], 1 << 16):
        w = W[s:s + (1 << 16)]
        c = w - w.mean(axis=1, keepdims=True)
        out[s:s + (1 << 16)] = np.sqrt((c * c).sum(axis=1))
    return out


def _corr_row(q, y, norms):
    m = q.size
    qc = q - q.mean()
    yc = y - y.mean()
    nfft = 1 << int(np.ceil(np.log2(y.size + m)))
    F = np.fft.rfft(yc, nfft) * np.conj(np.fft.rfft(qc, nfft))
    return np.fft.irfft(F, nfft)[: y.size - m + 1] / (norms * np.linalg.norm(qc))


def test_negative_best_matches_reach_the_column_side():
    import torch
    a, b = _series()
    ns = N - M + 1
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(a).to(dev)
    b_d = torch.from_numpy(b).to(dev)
    init = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))
    prof = torch.full((ns,), init, dtype=torch.int64, device=dev)
    prof_b = torch.full((ns,), init, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    print(f"[negative matches] AB-join n={N} m={M} FP64", flush=True)
    C.join_device(a_d.data_ptr(), N, b_d.data_ptr(), N,
                  C.make_params(M, "1NN_INDEX", precision="double", has_b=True,
                                keep_rows_separate=True),
                  prof.data_ptr(), prof_b.data_ptr())
    corr, idx = C.split_mp_entry(prof.cpu().numpy().view(np.uint64))

    # falling windows of A: correlation with a pure ramp below -0.8
    t = np.arange(M, dtype=np.float64) - (M - 1) / 2
    an = _norms(a, M)
    ramp_corr = np.convolve(a, t[::-1], mode="valid") / (an * np.linalg.norm(t))
    falling = np.flatnonzero(ramp_corr < -0.8)
    assert falling.size > 1000
    pick = np.sort(np.random.Generator(np.random.PCG64(43)).choice(falling, WINDOWS,
                                                                  replace=False))
    bn = _norms(b, M)
    column_side = 0
    for i in pick:
        row = _corr_row(a[i:i + M], b, bn)
        j = int(np.argmax(row))
        best = row[j]
        assert best < 0, (i, best)
        column_side += j <= i
        assert abs(float(corr[i]) - best) <= 2.0 ** -23 * abs(best) + 1e-12, (i, corr[i], best)
        assert idx[i] == j or row[idx[i]] >= best - 1e-9, (i, idx[i], j)
    # the case the test exists for: best partners at B indices <= the A index
    assert column_side >= WINDOWS // 4, column_side
