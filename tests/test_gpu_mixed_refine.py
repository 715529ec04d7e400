"""The mixed form's FP64 refinement of 1NN_INDEX profiles (refine_kernel.hip).

After the FP32 sweep each improved entry's match and the lags within two of
it are re-evaluated in FP64, so the mixed form picks what the FP64 form picks
wherever the FP64 match lies within two lags of the FP32 one, and stores
that match's FP64-accurate correlation.  Checked against the FP64 GPU form
(itself checked against the oracle and SCAMP's CPU path by the parity tests):
index agreement, the stored correlations where the indices agree, every
disagreement a near-tie of the exact FP64 correlations, and a prior that
beats the join kept in place."""
import numpy as np
import pytest

from helpers import exact_corr, random_walk
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu


def _both_forms(a, b, m, **kw):
    out = {}
    for prec in ("double", "mixed"):
        p = C.make_params(m, "1NN_INDEX", precision=prec, has_b=b is not None, **kw)
        pa, pb = C.join(a, b, p)
        out[prec] = (pa, pb)
    return out


def _compare(a, b, m, p64, p32, agree_min, mask=None):
    """Profiles of a's windows matched in b (b None: a self-join)."""
    c64, i64 = C.split_mp_entry(p64)
    c32, i32 = C.split_mp_entry(p32)
    valid = c64 > -1
    assert np.array_equal(valid, c32 > -1)
    if mask is not None:
        valid &= mask
    same = valid & (i32 == i64)
    agree = same.sum() / valid.sum()
    assert agree >= agree_min, agree
    # where the indices agree the stored values are FP64-accurate: equal after
    # FP32 rounding up to a rounding-boundary case of one unit
    d = np.abs(c32[same].astype(np.float64) - c64[same])
    assert d.max() <= 1.2e-7, d.max()
    assert np.mean(d == 0) >= 0.999
    # a disagreement is a near-tie of the exact correlations
    rows = np.nonzero(valid & ~same)[0]
    if rows.size:
        bb = a if b is None else b
        g = np.abs(exact_corr(a, bb, m, rows, i64[rows]) -
                   exact_corr(a, bb, m, rows, i32[rows]))
        assert g.max() <= 1e-6, g.max()
    return agree


@pytest.mark.parametrize("n,m,agree_min", [
    (1 << 18, 256, 0.9999),
    # short series: many matches lie within two lags of either end, where the
    # candidates are evaluated one at a time
    (1500, 24, 0.995),
    (700, 100, 0.995),
])
def test_selfjoin_refined_to_the_fp64_choice(n, m, agree_min):
    a = random_walk(n, 41, nan_runs=8 if n > 10000 else 1)
    r = _both_forms(a, None, m)
    _compare(a, None, m, r["double"][0], r["mixed"][0], agree_min)


def test_abjoin_short_b_edges():
    """B barely longer than the window: most of A's matches sit within two
    lags of B's ends."""
    a = random_walk(5000, 45)
    b = random_walk(70, 46)
    m = 64
    r = _both_forms(a, b, m, computing_rows=True, keep_rows_separate=True)
    _compare(a, b, m, r["double"][0], r["mixed"][0], 0.995)
    _compare(b, a, m, r["double"][1], r["mixed"][1], 0.995)


def test_abjoin_refined_both_profiles():
    a = random_walk(150000, 42, nan_runs=3)
    b = random_walk(120000, 43)
    m = 128
    r = _both_forms(a, b, m, computing_rows=True, keep_rows_separate=True)
    _compare(a, b, m, r["double"][0], r["mixed"][0], 0.9999)   # A's windows in B
    _compare(b, a, m, r["double"][1], r["mixed"][1], 0.9999)   # B's windows in A


def test_refinement_keeps_a_better_prior():
    """Entries whose prior beats everything in the join keep the prior's value
    and index untouched (the refinement only re-evaluates entries the join
    improved)."""
    a = random_walk(60000, 44)
    m = 100
    ns = a.size - m + 1
    prior = C.init_profile("1NN_INDEX", ns).copy()
    rng = np.random.default_rng(3)
    cols = rng.choice(ns, 50, replace=False)
    e = np.zeros(1, np.uint64)
    e.view(np.float32)[0] = 1.0
    e.view(np.uint32)[1] = 7
    prior[cols] = e[0]
    p = C.make_params(m, "1NN_INDEX", precision="mixed")
    out, _ = C.join(a, None, p, profile_a=prior)
    assert np.array_equal(out[cols], prior[cols])
    ref, _ = C.join(a, None, C.make_params(m, "1NN_INDEX", precision="double"))
    rest = np.ones(ns, bool)
    rest[cols] = False
    _compare(a, None, m, ref, out, 0.9999, mask=rest)


@pytest.mark.parametrize("grow,gcol", [(0, 0), (0, 131072), (131072, 0)])
def test_aligned_abjoin_refinement_respects_exclusion(grow, gcol):
    """An aligned AB-join of a series with itself excludes |global col -
    global row| < m/4; the refined matches (lags j - 2 .. j + 2 of the sweep's)
    must stay outside that zone, in both profiles, and agree with FP64."""
    a = random_walk(6000, 62)
    b = a.copy()
    m = 64
    e = m // 4
    kw = dict(is_aligned=True, distributed_start_row=grow, distributed_start_col=gcol,
              computing_rows=True, keep_rows_separate=True)
    r = _both_forms(a, b, m, **kw)
    delta = gcol - grow
    for k, (src, dst) in enumerate(((a, b), (b, a))):
        c32, i32 = C.split_mp_entry(r["mixed"][k])
        ok = c32 > -1
        rows = np.nonzero(ok)[0].astype(np.int64)
        j = i32[ok].astype(np.int64)
        # profile A (k = 0): A index = row, B index = match; profile B the reverse
        d = (rows - j + delta) if k == 0 else (j - rows + delta)
        assert np.all(np.abs(d) >= e)
        _compare(src, dst, m, r["double"][k], r["mixed"][k], 0.999)
