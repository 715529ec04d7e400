"""Autotuning of the mixed form's FP64 re-anchoring interval R.

The FP32 recurrence (compensated: the leading part of each covariance is
rounded once per block of 4 rows, a small partial sum every row) drifts from
the FP64 covariance by FP32 roundings that accumulate along the diagonal;
re-anchoring to an exact FP64 value every R rows bounds that drift, at a
cost of about m/R FP64 FMAs per pair.  ``autotune_anchor`` joins a prefix of
the caller's own series (self-join, or AB-join with a prefix of ``b``) with
the plain FP64 form and with the mixed form at several R, and keeps the
largest R whose error stays a safety factor below the target on that sample
and that passes the engine's own FP64 audit of the FP32 sweep (an R at which
the engine had to recompute the join in FP64 is never chosen).
"""
from __future__ import annotations

import numpy as np

from . import _capi as C

CANDIDATES = (16384, 8192, 4096, 2048, 1024, 512, 256)


def _exact_corr(a, b, m, rows, cols):
    """FP64 Pearson correlation of windows a[r:r+m] and b[c:c+m]."""
    from numpy.lib.stride_tricks import sliding_window_view as sw
    x = sw(a, m)[rows]
    y = sw(b, m)[cols]
    x = x - x.mean(axis=1, keepdims=True)
    y = y - y.mean(axis=1, keepdims=True)
    return (x * y).sum(1) / np.sqrt((x * x).sum(1) * (y * y).sum(1))


def mixed_error(a, m, anchor, n_sample=1 << 17, device=0, b=None):
    """(max |dcorr|, index agreement, near-tie gap) of the mixed form against
    the FP64 form on a prefix of ``a`` (and of ``b`` for an AB-join, A's
    profile).  The gap is the largest exact FP64 correlation difference
    between the FP64 form's match and the mixed form's match over the rows
    where the two indices differ (0 if none): a small gap means every
    disagreement is a near-tie.  An R at which the engine's audit sent the
    join back to FP64 reports an infinite error."""
    s = np.ascontiguousarray(a[: min(len(a), n_sample)], dtype=np.float64)
    t = None if b is None else np.ascontiguousarray(b[: min(len(b), n_sample)],
                                                    dtype=np.float64)
    kw = dict(has_b=t is not None, computing_rows=t is None)
    p64, _ = C.join(s, t, C.make_params(m, precision="double", **kw), device=device)
    p32, _ = C.join(s, t, C.make_params(m, precision="mixed", anchor_interval=anchor, **kw),
                    device=device)
    ran, drift = C.context(device).last_precision()
    if ran != "mixed" or drift:
        return float("inf"), 0.0, float("inf")
    c64, i64 = C.split_mp_entry(p64)
    c32, i32 = C.split_mp_entry(p32)
    valid = c64 > -1
    if not valid.any():
        return 0.0, 1.0, 0.0
    d = float(np.max(np.abs(c32[valid].astype(np.float64) - c64[valid])))
    agree = float(np.mean(i32[valid] == i64[valid]))
    rows = np.nonzero(valid & (i32 != i64) & (c32 > -1))[0]
    gap = 0.0
    if rows.size:
        other = s if t is None else t
        gap = float(np.max(np.abs(_exact_corr(s, other, m, rows, i64[rows])
                                  - _exact_corr(s, other, m, rows, i32[rows]))))
    return d, agree, gap


def autotune_anchor(a, m, target=1e-5, safety=4.0, n_sample=1 << 17, device=0, b=None):
    """Largest R in CANDIDATES whose sample error and near-tie gap are both
    <= target / safety (the smallest R otherwise)."""
    for r in CANDIDATES:
        d, _, gap = mixed_error(a, m, r, n_sample, device, b=b)
        if d <= target / safety and gap <= target / safety:
            return r
    return CANDIDATES[-1]
