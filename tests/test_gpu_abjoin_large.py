"""AB-join at full size: A = 2^21-sample random walk, B = 2^23 samples of three
sinusoids plus noise with 200 planted copies of a 256-sample motif
(bench.make_abjoin_series), window m = 256; 1NN_INDEX with both profiles
(A's columns, B's rows kept separate) and SUM_THRESH at 0.9, each in the
plain FP64 form and the mixed form.

The FP64 results are checked at sampled windows against exact correlation
rows: FFT sliding dot products of the centred query window with the whole
other series (numpy FP64) divided by directly computed window norms.  The
mixed results are checked against the FP64 ones.
"""
import numpy as np
import pytest

import bench
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

M = 256
THR = 0.9


@pytest.fixture(scope="module")
def series():
    a, b, _ = bench.make_abjoin_series()
    return a, b


def window_norms(y, m, direct, chunk=1 << 16):
    """FP64 norms of the centred windows of y: direct per-window sums, or
    (for series of small amplitude, where they lose nothing that matters)
    running sums of the globally centred series."""
    if not direct:
        yc = y - y.mean()
        c1 = np.concatenate(([0.0], np.cumsum(yc)))
        c2 = np.concatenate(([0.0], np.cumsum(yc * yc)))
        s1 = c1[m:] - c1[:-m]
        return np.sqrt(np.maximum((c2[m:] - c2[:-m]) - s1 * s1 / m, 0.0))
    from numpy.lib.stride_tricks import sliding_window_view as sw
    W = sw(y, m)
    out = np.empty(W.shape[0])
    for s in range(0, W.shape[0], chunk):
        w = W[s:s + chunk]
        c = w - w.mean(axis=1, keepdims=True)
        out[s:s + chunk] = np.sqrt((c * c).sum(axis=1))
    return out


def corr_row(q, y, norms):
    """Exact correlation of the window q with every window of y."""
    m = q.size
    qc = q - q.mean()
    yc = y - y.mean()  # a constant shift leaves centred dot products unchanged
    nfft = 1 << int(np.ceil(np.log2(y.size + m)))
    F = np.fft.rfft(yc, nfft) * np.conj(np.fft.rfft(qc, nfft))
    dots = np.fft.irfft(F, nfft)[: y.size - m + 1]
    return dots / (norms * np.linalg.norm(qc))


def direct_corr(q, y, j):
    x = q - q.mean()
    w = y[j:j + q.size] - y[j:j + q.size].mean()
    return float(x @ w / np.sqrt((x @ x) * (w @ w)))


def joins(a, b, precision):
    nn_a, nn_b = C.join(a, b, C.make_params(M, "1NN_INDEX", precision=precision, has_b=True,
                                            computing_rows=True, keep_rows_separate=True))
    ran = C.context(0).last_precision()
    print(f"1NN_INDEX {precision}: ran in {ran[0]}, audit failures {ran[1]}", flush=True)
    # the mixed results below must come from the FP32 sweep itself
    assert ran == (precision, 0), ran
    s_a, s_b = C.join(a, b, C.make_params(M, "SUM_THRESH", precision=precision, has_b=True,
                                          computing_rows=True, keep_rows_separate=True,
                                          threshold=THR))
    return C.split_mp_entry(nn_a), C.split_mp_entry(nn_b), s_a, s_b


@pytest.fixture(scope="module")
def results(series):
    a, b = series
    return {"double": joins(a, b, "double"), "mixed": joins(a, b, "mixed")}


def test_fp64_against_exact_rows(series, results):
    a, b = series
    (ca, ia), (cb, ib), sa, sb = results["double"]
    assert ca.size == a.size - M + 1 and cb.size == b.size - M + 1
    # A is a random walk (|values| up to ~2e3): direct sums; B stays within ~20
    na, nb = window_norms(a, M, direct=True), window_norms(b, M, direct=False)
    rng = np.random.default_rng(7)
    _, _, offsets = bench.make_abjoin_series()
    cases = [("a", int(r)) for r in rng.integers(0, ca.size, 3)]
    cases += [("b", int(offsets[0])), ("b", int(rng.integers(0, cb.size)))]
    for side, r in cases:
        if side == "a":
            row = corr_row(a[r:r + M], b, nb)
            c, i, s = ca[r], ia[r], sa[r]
        else:
            row = corr_row(b[r:r + M], a, na)
            c, i, s = cb[r], ib[r], sb[r]
        q, y = (a, b) if side == "a" else (b, a)
        q = q[r:r + M]
        # the FFT row locates the candidates; direct dot products decide
        cand = np.nonzero(row >= row.max() - 1e-6)[0]
        best = max(direct_corr(q, y, int(j)) for j in cand)
        assert abs(float(c) - best) <= 1.2e-7, (side, r, c, best)
        assert best - direct_corr(q, y, int(i)) <= 1e-9, (side, r, i, cand)
        exact_sum = row[row > THR].sum()
        border = np.count_nonzero(np.abs(row - THR) <= 1e-8)
        assert abs(s - exact_sum) <= 1e-8 * max(1.0, exact_sum) + border, (side, r)


def test_mixed_against_fp64(series, results):
    a, b = series
    d, x = results["double"], results["mixed"]
    from helpers import exact_corr
    for k, (q, y) in enumerate(((a, b), (b, a))):
        (dc, di), (mc, mi) = d[k], x[k]
        assert np.array_equal(dc < -1, mc < -1)
        ok = dc > -1
        dmax = float(np.max(np.abs(mc[ok].astype(np.float64) - dc[ok])))
        assert dmax <= 1e-5, dmax
        diff = np.nonzero(ok & (mi != di))[0]
        agree = float(np.mean(mi[ok] == di[ok]))
        # B is periodic: many of its windows are near-identical, so raw index
        # agreement is data-bound; what must hold is that every mismatch is a
        # near-tie in exact FP64 correlation
        rows = diff[np.random.default_rng(k).permutation(diff.size)[:4000]]
        gap = 0.0
        if rows.size:
            gap = float(np.max(exact_corr(q, y, M, rows, di[rows])
                               - exact_corr(q, y, M, rows, mi[rows])))
        print(f"profile {'ab'[k]}: max|dcorr| {dmax:.3g} agreement {agree:.6f} "
              f"mismatches {diff.size} max gap {gap:.3g}")
        # the mixed choice can be worse than the FP64 one by at most the
        # error of both correlations
        assert gap <= 2 * dmax + 1e-7, (gap, dmax)
        assert agree >= 0.99, agree
    for k in (2, 3):
        s64, s32 = d[k], x[k]
        rel = np.abs(s32 - s64) / np.maximum(1.0, np.abs(s64))
        # a correlation within ~1e-6 of the threshold may fall on either side
        print(f"sums {'ab'[k - 2]}: rows within 1e-5 {np.mean(rel <= 1e-5):.6f} max rel {rel.max():.3g}")
        assert np.mean(rel <= 1e-5) >= 0.999, np.mean(rel <= 1e-5)
