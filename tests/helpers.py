"""Shared checks for the parity tests.

The brute-force references here are textbook FP64 Pearson correlations of
the windows (numpy), reduced the way the reference's test oracle reduces its
distance matrix (test/distance_matrix_fast.py:222-262); the index comparison
follows the reference's tie rule (test/test_common.py:3-17): an index that
differs from the expected one is accepted when the two candidates'
correlations agree (FP64 within 1e-9, or equal after FP32 rounding).
"""
import os

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "sample_io.npz")


def golden():
    return np.load(GOLDEN)


def random_walk(n, seed, nan_runs=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    a = np.cumsum(rng.standard_normal(n))
    for _ in range(nan_runs):
        L = int(rng.integers(1, 6))
        s = int(rng.integers(0, n - L))
        a[s:s + L] = np.nan
    return a


def centred_windows(a, m):
    W = sliding_window_view(np.nan_to_num(a, nan=0.0), m).astype(np.float64)
    W = W - W.mean(axis=1, keepdims=True)
    bad = ~np.isfinite(sliding_window_view(a, m)).all(axis=1)
    ss = (W * W).sum(axis=1)
    bad |= ss <= 1e-13
    return W, ss, bad


def exact_corr(a, b, m, i_idx, j_idx):
    """FP64 Pearson correlation of a[i:i+m] and b[j:j+m] for index arrays."""
    A = sliding_window_view(a, m)[i_idx].astype(np.float64)
    B = sliding_window_view(b, m)[j_idx].astype(np.float64)
    A = A - A.mean(1, keepdims=True)
    B = B - B.mean(1, keepdims=True)
    return (A * B).sum(1) / np.sqrt((A * A).sum(1) * (B * B).sum(1))


def corr_matrix(a, b, m, self_join):
    """Full FP64 distance matrix D[row in b, col in a] with -2 for NaN windows
    and (self-join) the exclusion zone |i - j| < m // 4."""
    bb = a if b is None else b
    A, ssa, bada = centred_windows(a, m)
    B, ssb, badb = centred_windows(bb, m)
    with np.errstate(invalid="ignore", divide="ignore"):
        D = (B @ A.T) / np.sqrt(ssb[:, None] * ssa[None, :])
    D[badb, :] = -2
    D[:, bada] = -2
    if self_join:
        e = m // 4
        ii, jj = np.indices(D.shape)
        D[np.abs(ii - jj) < e] = -2
    D[~np.isfinite(D)] = -2
    return D


def reduce_1nn_index(D):
    corr = D.max(axis=0)
    idx = D.argmax(axis=0)
    idx[corr <= -2] = -1
    return corr, idx


def reduce_sum(D, thr):
    return np.where(D > thr, D, 0.0).sum(axis=0)


def upper_pairs(D):
    """A self-join distance matrix with only the pairs (row i, column j > i):
    what a self-join MATRIX_SUMMARY sees (the reference's upper-triangular
    tiles with rows switched off, src/SCAMP.cpp:34-43, src/common.cpp:98-106)."""
    ii, jj = np.indices(D.shape)
    return np.where(jj > ii, D, -2.0)


def reduce_matrix(D, H, W, thr):
    rpc = -(-D.shape[0] // H)
    cpc = -(-D.shape[1] // W)
    out = np.full((H, W), -2.0)
    for r in range(H):
        for c in range(W):
            blk = D[r * rpc:(r + 1) * rpc, c * cpc:(c + 1) * cpc]
            if blk.size and blk.max() > thr:
                out[r, c] = blk.max()
    return out


def check_1nn_index(a, b, m, corr, idx, ref_corr, ref_idx, corr_tol=None):
    """Compare a 1NN_INDEX profile with an expected one (tie rule).  Returns a
    dict of failure counts (all zero when the profiles agree)."""
    bb = a if b is None else b
    corr = np.asarray(corr, dtype=np.float64)
    ref_corr = np.asarray(ref_corr, dtype=np.float64)
    idx = np.where(corr < -1, -1, np.asarray(idx, dtype=np.int64))
    ref_idx = np.where(ref_corr < -1, -1, np.asarray(ref_idx, dtype=np.int64))
    out = {"nan_mismatch": int(np.count_nonzero((corr < -1) != (ref_corr < -1)))}
    diff = np.nonzero(idx != ref_idx)[0]
    diff = diff[(idx[diff] >= 0) & (ref_idx[diff] >= 0)]
    non_tie = 0
    if diff.size:
        cg = exact_corr(a, bb, m, diff, idx[diff])
        cr = exact_corr(a, bb, m, diff, ref_idx[diff])
        tie = (np.abs(cg - cr) <= 1e-9) | (
            corr[diff].astype(np.float32) == ref_corr[diff].astype(np.float32))
        non_tie = int(np.count_nonzero(~tie))
    out["non_tie_index"] = non_tie
    both = (corr > -1) & (ref_corr > -1)
    tol = corr_tol if corr_tol is not None else 2 ** -23
    out["corr_off"] = int(np.count_nonzero(np.abs(corr[both] - ref_corr[both]) > tol))
    return out


def euclidean(corr, m):
    c = np.asarray(corr, dtype=np.float64)
    d = np.sqrt(np.maximum(2 * m * (1 - c), 0))
    d[c < -1] = np.nan
    return d


def reduced_precision_errors(a, b, m, corr, idx, ref_corr, ref_idx):
    """Errors of a reduced-precision 1NN_INDEX profile (a's windows matched in
    b; b None: a itself) against FP64 reference values: max |dcorr| over the
    valid entries, the raw index agreement, the agreement when ties under the
    reference's rule (test/test_common.py:3-17: the two candidates' FP64
    correlations within 1e-9 or equal after FP32 rounding) count as agreeing,
    the number of non-tie disagreements, and the largest exact FP64 gap between
    the reference's match and the reduced form's at a disagreement."""
    bb = a if b is None else b
    corr = np.asarray(corr, dtype=np.float64)
    ref_corr = np.asarray(ref_corr, dtype=np.float64)
    ok = ref_corr > -2
    n = int(ok.sum())
    dmax = float(np.max(np.abs(corr[ok] - ref_corr[ok]), initial=0.0))
    dis = np.nonzero(ok & (np.asarray(idx) != np.asarray(ref_idx)))[0]
    gap, non_tie = 0.0, 0
    if dis.size:
        er = exact_corr(a, bb, m, dis, np.asarray(ref_idx)[dis])
        eg = exact_corr(a, bb, m, dis, np.asarray(idx)[dis])
        gap = float(np.max(er - eg))
        tie = (np.abs(er - eg) <= 1e-9) | (er.astype(np.float32) == eg.astype(np.float32))
        non_tie = int(np.count_nonzero(~tie))
    return {"entries": n, "max_dcorr": dmax, "raw_agreement": 1 - dis.size / max(n, 1),
            "agreement": 1 - non_tie / max(n, 1), "non_tie": non_tie,
            "disagreements": int(dis.size), "max_gap": gap}
