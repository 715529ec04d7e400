"""Self-join at one GPU's real scale: n = 2^25 float32 random walk (PCG64 seed
5; FP64 statistics from the float32 values), m = 4096, the mixed form's
1NN_INDEX profile with the autotuned re-anchor interval and a 1024 x 1024
MATRIX_SUMMARY.

Checks (the oracle is the C restatement of SCAMP's statistics plus textbook
FP64 dot products; nothing here uses the engine's own FP64 form):
  * 10^5 seeded sampled rows: the exact FP64 correlation of each reported
    pair is within 1e-5 of the stored value;
  * every reported match lies outside the exclusion zone |i - j| >= m/4 and
    no window reports a NaN (the input has no non-finite samples);
  * 16 sampled rows against the whole FP64 correlation row (overlap-add FFT
    of the centred window against the series, the best candidates recomputed
    directly): the stored value is within 1e-5 of the row's true maximum, and
    an index that differs from the true argmax is a near-tie within 1e-5;
  * for every band of subsequences, the larger of the summary's row and
    column maxima for that band (a self-join summary holds each pair i < j
    once, in cell (band of i, band of j); src/common.cpp:98-106 cell
    geometry: rows per cell = ceil(n_sub / 1024)) is the profile's maximum
    over the band, up to the mixed form's rounding.
The run takes a few minutes; progress is printed as it goes.
"""
import time

import numpy as np
import pytest

from helpers import exact_corr
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

N, M, H = 1 << 25, 4096, 1024
SAMPLED_ROWS = 100_000
FFT_ROWS = 16
BAR = 1e-5


def series():
    steps = np.random.Generator(np.random.PCG64(5)).standard_normal(N).astype(np.float32)
    return np.cumsum(steps, dtype=np.float32).astype(np.float64)


_CAPSYS = None


def log(msg, t0):
    """Progress to the terminal even under pytest's output capture (each
    phase of this test runs for a minute or more)."""
    line = f"[n=2^25 m=4096] {msg} ({time.perf_counter() - t0:.1f} s)"
    if _CAPSYS is not None:
        with _CAPSYS.disabled():
            print(line, flush=True)
    else:
        print(line, flush=True)


def fft_row(a, i, inorm, e):
    """(true max, its index, exact FP64 correlations of the best candidates)
    of row i over all windows outside the exclusion zone."""
    from scipy.signal import oaconvolve
    x = a[i:i + M] - a[i:i + M].mean()
    qt = oaconvolve(a, x[::-1], mode="valid")  # sum_t a[j + t] * x[t] = cov(i, j)
    corr = qt * inorm / np.sqrt(x @ x)
    lo, hi = max(0, i - e + 1), min(corr.size, i + e)
    corr[lo:hi] = -np.inf
    corr[~np.isfinite(corr)] = -np.inf
    cand = np.argpartition(corr, -8)[-8:]
    ex = exact_corr(a, a, M, np.full(cand.size, i), cand)
    k = int(np.argmax(ex))
    return float(ex[k]), int(cand[k])


@pytest.mark.timeout(1500)
def test_large_selfjoin_mixed_profile_and_summary(capsys):
    global _CAPSYS
    _CAPSYS = capsys
    import torch
    from scamp_amd import tuning
    t0 = time.perf_counter()
    a = series()
    ns = N - M + 1
    e = M // 4
    dev = torch.device("cuda:0")
    a_d = torch.from_numpy(a).to(dev)
    anchor = tuning.autotune_anchor(a, M)
    log(f"autotuned anchor interval {anchor}", t0)

    init = int(C.init_profile("1NN_INDEX", 1)[0].view(np.int64))
    prof = torch.full((ns,), init, dtype=torch.int64, device=dev)
    C.join_device(a_d.data_ptr(), N, 0, 0,
                  C.make_params(M, "1NN_INDEX", precision="mixed", anchor_interval=anchor),
                  prof.data_ptr())
    c, i = C.split_mp_entry(prof.cpu().numpy().view(np.uint64))
    ran = C.context(0).last_precision()
    log(f"mixed 1NN_INDEX done (ran in {ran[0]}, audit failures {ran[1]})", t0)
    assert ran == ("mixed", 0)  # the FP32 sweep stayed within its bound

    init_m = int(C.init_profile("MATRIX_SUMMARY", 1, 1, 1)[0].view(np.int32))
    mat = torch.full((H * H,), init_m, dtype=torch.int32, device=dev)
    C.join_device(a_d.data_ptr(), N, 0, 0,
                  C.make_params(M, "MATRIX_SUMMARY", precision="mixed",
                                anchor_interval=anchor, matrix_height=H, matrix_width=H),
                  mat.data_ptr())
    summary = mat.cpu().numpy().view(np.float32).reshape(H, H)
    assert C.context(0).last_precision()[0] == "mixed"
    del a_d, prof, mat
    log("mixed MATRIX_SUMMARY done", t0)

    # every window valid, every match outside the exclusion zone
    assert np.all(c > -2) and np.all(i >= 0) and np.all(i < ns)
    assert np.all(np.abs(np.arange(ns) - i.astype(np.int64)) >= e)

    # 10^5 sampled rows against exact FP64 dot products
    rng = np.random.default_rng(5)
    rows = np.sort(rng.choice(ns, SAMPLED_ROWS, replace=False))
    worst = 0.0
    for k in range(0, rows.size, 2000):
        r = rows[k:k + 2000]
        ex = exact_corr(a, a, M, r, i[r].astype(np.int64))
        worst = max(worst, float(np.max(np.abs(ex - c[r]))))
    log(f"{SAMPLED_ROWS} sampled rows: max |stored - exact FP64| {worst:.3g}", t0)
    assert worst <= BAR

    # whole correlation rows
    _, _, inorm, _, _ = O.stats(a, M)
    worst_gap = 0.0
    for r in rng.choice(ns, FFT_ROWS, replace=False):
        best, j = fft_row(a, int(r), inorm, e)
        gap = best - float(c[r])
        worst_gap = max(worst_gap, abs(gap))
        assert abs(gap) <= BAR, (int(r), best, float(c[r]))
        if j != int(i[r]):
            assert exact_corr(a, a, M, np.array([r]), np.array([i[r]]))[0] >= best - BAR
    log(f"{FFT_ROWS} full FP64 rows: max |stored - true row max| {worst_gap:.3g}", t0)

    # summary bands against the profile's band maxima: the pairs of band r
    # sit in cell row r (as the earlier subsequence) and cell column r (as
    # the later one)
    rpc = -(-ns // H)
    band_max = np.array([c[r * rpc:(r + 1) * rpc].max() for r in range(H)], dtype=np.float64)
    filled = np.where(np.isnan(summary), -2.0, summary)
    row_max = np.maximum(filled.max(axis=1), filled.max(axis=0)).astype(np.float64)
    diff = float(np.max(np.abs(row_max - np.where(band_max > 0, band_max, -2.0))))
    log(f"summary row maxima vs profile band maxima: max |diff| {diff:.3g}", t0)
    assert diff <= 2e-6
