"""The reference's own edge-case grid, run on the HIP engine.

Restates test/run_tests.py:12-17,63-92,263-270: inputs of 20, 100, 250, 500,
1500 and 8000 samples drawn with numpy's legacy generator seeded 13
(np.random.seed(13); np.random.randn per size, in that order), plus the two
static inputs randomwalk1K_nan and poorly_conditioned_test; every ordered pair
(A, B) of them -- the self-join when A is B -- at every window of 5, 7, 13,
25, 33, 78, 92, 100, 189, 339 and 500 that generate_tests admits (window <=
len - len // 4 on a self-join, <= both lengths otherwise); profile types
1NN_INDEX, 1NN and SUM_THRESH at thresholds 0, 0.125 and 0.5, all in the
plain FP64 form.

Expected results come from the brute-force FP64 correlation matrix
(helpers.corr_matrix, the reference's distance_matrix_np,
test/distance_matrix_fast.py:141-196) reduced as the reference reduces it
(:222-262).  The comparisons are the reference's (test/test_common.py:3-51:
index equal unless the FP32 correlations agree, vectors allclose, SUM_THRESH
allclose or off by whole threshold multiples) with tighter tolerances on top:
1NN correlations within one FP32 rounding step of the FP64 value, 1NN_INDEX
through helpers.check_1nn_index, and SUM_THRESH within 1e-9 relative.
"""
import numpy as np
import pytest

from helpers import check_1nn_index, corr_matrix, golden, reduce_1nn_index, reduce_sum
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

SIZES = [20, 100, 250, 500, 1500, 8000]
WINDOWS = [5, 7, 13, 25, 33, 78, 92, 100, 189, 339, 500]
THRESHOLDS = [0.0, 0.125, 0.5]
STATIC = ["randomwalk1K_nan", "poorly_conditioned_test"]
CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}

_inputs = None


def inputs():
    """generate_input_arrays + read_file_inputs (test/run_tests.py:75-86)."""
    global _inputs
    if _inputs is None:
        rs = np.random.RandomState(13)
        arrs = [rs.randn(s) for s in SIZES]
        g = golden()
        arrs += [g["in_" + name] for name in STATIC]
        _inputs = arrs
    return _inputs


def windows_for(a, b):
    """generate_tests (test/run_tests.py:88-98)."""
    la, lb = len(inputs()[a]), len(inputs()[b])
    excl = la // 4 if a == b else 0
    return [w for w in WINDOWS if w <= la - excl and w <= lb - excl]


def compare_vectors_sum(valid, check, thresh):
    """test/test_common.py:27-51: equal within allclose, or every difference a
    whole number of thresholds (a correlation straddling the threshold)."""
    if np.allclose(valid, check, equal_nan=False):
        return True
    diff = np.round(np.abs(valid - check) / thresh, decimals=6)
    diff = diff - np.floor(diff)
    diff[diff < 1e-6] = 0
    return np.count_nonzero(diff) == 0


N_INPUTS = len(SIZES) + len(STATIC)
PAIRS = [(a, b) for a in range(N_INPUTS) for b in range(N_INPUTS)]


@pytest.mark.parametrize("a,b", PAIRS)
def test_reference_grid(a, b):
    A = inputs()[a]
    B = None if a == b else inputs()[b]
    for m in windows_for(a, b):
        tag = (len(A), None if B is None else len(B), m)
        kw = dict(has_b=B is not None, computing_rows=B is None)
        D = corr_matrix(A, B, m, B is None)  # rows: B windows, columns: A windows
        # 1NN_INDEX
        p, _ = C.join(A, B, C.make_params(m, "1NN_INDEX", **kw))
        c, i = C.split_mp_entry(p)
        rc, ri = reduce_1nn_index(D)
        assert check_1nn_index(A, B, m, c, i, rc, ri) == CLEAN, tag
        has = rc > -2
        # the reference's compare_index / compare_vectors
        no_agree = has & (np.where(c < -1, -1, i) != ri)
        assert np.all(c[no_agree].astype(np.float32) == rc[no_agree].astype(np.float32)), tag
        assert np.allclose(np.where(has, rc, np.nan), np.where(c > -2, c, np.nan),
                           equal_nan=True), tag
        # 1NN
        p1, _ = C.join(A, B, C.make_params(m, "1NN", **kw))
        assert np.array_equal(p1 > -2, has), tag
        assert np.max(np.abs(p1[has].astype(np.float64) - rc[has]), initial=0) <= 2 ** -23, tag
        # SUM_THRESH
        for thr in THRESHOLDS:
            s, _ = C.join(A, B, C.make_params(m, "SUM_THRESH", threshold=thr, **kw))
            ref = reduce_sum(D, thr)
            assert compare_vectors_sum(ref, s, thr) if thr > 0 else \
                np.allclose(ref, s, equal_nan=False), (tag, thr)
            # every pair above the threshold summed in FP64; an entry may
            # differ by whole thresholds only where a correlation lies within
            # rounding of the threshold (the reference's straddle rule)
            off = np.abs(s - ref) > 1e-9 * np.maximum(1.0, np.abs(ref))
            if np.any(off):
                assert thr > 0, (tag, thr, np.nonzero(off)[0][:5])
                q = np.abs(s - ref)[off] / thr
                assert np.all(np.abs(q - np.round(q)) <= 1e-9 * (1 + np.abs(ref[off]) / thr)), \
                    (tag, thr)
