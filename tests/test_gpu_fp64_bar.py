"""The plain FP64 form at its stated bar: every correlation within 1e-9
absolute of SCAMP's CPU double path, indices exact except at ties.

The engine stores 1NN_INDEX correlations in FP32 (SCAMP's mp_entry layout,
src/common.h:70-74), so the FP64 value behind each stored one is checked
through its rounding: a stored value f is consistent with an FP64 correlation
within 1e-9 of the oracle's FP64 value r exactly when
fl32(r - 1e-9) <= f <= fl32(r + 1e-9) (rounding is monotone).  SUM_THRESH
sums are stored in FP64, so at a high threshold (few terms per entry) each sum
is compared with the oracle's FP64-mode sum directly: |difference| <= 1e-9
per term.  The oracle's FP64 mode (oracle.join(..., fp64=True)) is SCAMP's CPU
double path without the FP32 rounding of stored values
(src/cpu_kernels.cpp:146-149); indices are compared with the oracle's default
mode, which rounds as SCAMP does, under the reference's tie rule
(test/test_common.py:3-17).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import check_1nn_index, golden, random_walk
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

TOL = 1e-9
CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


def k1_series():
    """The n = 131072 random walk (cumulative N(0,1) steps, PCG64 seed 1)."""
    return np.cumsum(np.random.Generator(np.random.PCG64(1)).standard_normal(131072))


def consistent_fp32(stored, ref64, tol=TOL):
    """Entries whose stored FP32 value could be the rounding of an FP64 value
    within `tol` of ref64."""
    lo = (ref64 - tol).astype(np.float32)
    hi = (ref64 + tol).astype(np.float32)
    f = np.asarray(stored, dtype=np.float32)
    return (f >= lo) & (f <= hi)


def fp64_bar_1nn_index(a, b, m, ref32, ref64, **kw):
    """GPU FP64 1NN_INDEX against the oracle in both modes."""
    p, pb = C.join(a, b, C.make_params(m, has_b=b is not None, **kw))
    c, i = C.split_mp_entry(p)
    oc, oi = ref32[0], ref32[1]
    assert check_1nn_index(a, b, m, c, i, oc, oi, corr_tol=2 ** -24) == CLEAN
    rc = ref64[0]
    both = (c > -2) & (rc > -2)
    assert np.array_equal(c > -2, rc > -2)
    ok = consistent_fp32(c[both], rc[both])
    assert ok.all(), (np.count_nonzero(~ok), rc[both][~ok][:4], c[both][~ok][:4])
    return int(both.sum())


def test_k1_selfjoin_corr_within_1e9_of_scamp_double():
    """n = 131072, m = 100: SCAMP's CPU path (the compiled reference when
    built, else the oracle) for the indices, the oracle's FP64 mode for the
    values."""
    a = k1_series()
    m = 100
    impl = "reference" if O.ref_available() else "port"
    with ThreadPoolExecutor(2) as ex:
        f32 = ex.submit(O.join, a, None, m, impl=impl, num_workers=8)
        f64 = ex.submit(O.join, a, None, m, fp64=True)
        ref32, ref64 = f32.result(), f64.result()
    assert fp64_bar_1nn_index(a, None, m, ref32, ref64) > 130000


@pytest.mark.parametrize("na,nb,m,nan_runs", [(20000, None, 256, 4), (9000, None, 33, 2),
                                              (12000, 7000, 64, 2), (3000, 15000, 500, 0)])
def test_small_joins_corr_within_1e9(na, nb, m, nan_runs):
    a = random_walk(na, na + m, nan_runs)
    b = None if nb is None else random_walk(nb, nb + m + 1, 1)
    kw = dict(computing_rows=b is None)
    ref32 = O.join(a, b, m)
    ref64 = O.join(a, b, m, fp64=True)
    fp64_bar_1nn_index(a, b, m, ref32, ref64, **kw)


@pytest.mark.parametrize("name", ["randomwalk64K_nan", "poorly_conditioned_test",
                                  "randomwalk1K_nan"])
def test_reference_inputs_corr_within_1e9(name):
    a = golden()["in_" + name]
    for m in (5, 25, 100):
        fp64_bar_1nn_index(a, None, m, O.join(a, None, m), O.join(a, None, m, fp64=True))


def sum_bar(a, b, m, thr, oracle_sum):
    s, _ = C.join(a, b, C.make_params(m, "SUM_THRESH", threshold=thr, has_b=b is not None,
                                      computing_rows=b is None))
    o = oracle_sum
    # each term is above thr, so an entry holds at most |sum| / thr of them
    terms = np.maximum(1.0, np.ceil(np.abs(o) / thr))
    d = np.abs(s - o)
    off = d > TOL * terms
    # a correlation within rounding of the threshold may be counted on one
    # side only (the reference's straddle rule, test/test_common.py:27-51)
    q = d[off] / thr
    straddle = np.abs(q - np.round(q)) <= TOL * (terms[off] + 1)
    assert straddle.all(), (np.nonzero(off)[0][~straddle][:5], d[off][~straddle][:5])
    assert np.count_nonzero(off) <= 2, np.count_nonzero(off)
    return int(np.count_nonzero(o > 0))


def test_k1_sum_thresh_within_1e9_per_term():
    """n = 131072, m = 100 self-join at threshold 0.95: most entries sum a few
    FP64 correlations, each of which must match the oracle's to 1e-9."""
    a = k1_series()
    thr = 0.95
    o, _ = O.join(a, None, 100, "SUM_THRESH", threshold=thr, fp64=True)
    assert sum_bar(a, None, 100, thr, o) > 10000


@pytest.mark.parametrize("na,nb,m,thr", [(30000, None, 64, 0.9), (20000, 25000, 128, 0.8)])
def test_sum_thresh_within_1e9_per_term(na, nb, m, thr):
    a = random_walk(na, na + 3, 2)
    b = None if nb is None else random_walk(nb, nb + 5, 1)
    o, _ = O.join(a, b, m, "SUM_THRESH", threshold=thr, fp64=True,
                  computing_rows=b is None)
    assert sum_bar(a, b, m, thr, o) > 100
