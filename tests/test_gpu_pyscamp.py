"""pyscamp on the GPU: the reference's module test (test/test_pyscamp.py:1-80)
re-run against this engine -- selfjoin / abjoin / *_sum / *_matrix on
np.random.random(8000) with m = 1024 -- checked with the reference's own
comparison rules (test/test_common.py) against a brute-force FP64 distance
matrix, plus the Euclidean/Pearson output conversions
(src/scamp_utils.cpp:127-141) and the path pyscamp -> SCAMP::do_SCAMP."""
import numpy as np
import pytest

from helpers import corr_matrix, reduce_1nn_index, reduce_matrix, reduce_sum, upper_pairs

pytestmark = pytest.mark.gpu


def _compare_index(valid, valid_corr, check, check_corr):
    # test/test_common.py:3-17
    no_agree = valid != check
    bad = valid_corr[no_agree].astype(np.float32) != check_corr[no_agree].astype(np.float32)
    return np.count_nonzero(bad) == 0


def _compare_sum(valid, check, thresh):
    # test/test_common.py:27-51
    if np.allclose(valid, check):
        return True
    diff = np.round(np.abs(valid - check) / thresh, 6)
    diff = diff - np.floor(diff)
    diff[diff < 1e-6] = 0
    return np.count_nonzero(diff) == 0


@pytest.fixture(scope="module")
def data():
    rng = np.random.RandomState(2024)
    arr = rng.random_sample(8000)
    arr2 = rng.random_sample(8000)
    return arr, arr2, corr_matrix(arr, None, 1024, True), corr_matrix(arr, arr2, 1024, False)


def test_selfjoin_pearson(data):
    import pyscamp as mp
    arr, _, dm_self, _ = data
    dist, index = mp.selfjoin(arr, 1024, pearson=True)
    vdist, vindex = reduce_1nn_index(dm_self)
    vdist = np.where(vdist <= -2, np.nan, vdist)
    assert dist.dtype == np.float32 and index.dtype == np.int32
    assert np.allclose(vdist, dist, equal_nan=True)
    assert _compare_index(vindex, vdist, index, dist)


def test_abjoin_pearson(data):
    import pyscamp as mp
    arr, arr2, _, dm_ab = data
    dist, index = mp.abjoin(arr, arr2, 1024, pearson=True)
    vdist, vindex = reduce_1nn_index(dm_ab)
    assert dist.shape == (arr.size - 1023,)
    assert np.allclose(vdist, dist, equal_nan=True)
    assert _compare_index(vindex, vdist, index, dist)


def test_selfjoin_euclidean_default(data):
    import pyscamp as mp
    arr, _, dm_self, _ = data
    dist, index = mp.selfjoin(arr, 1024)
    vdist, _ = reduce_1nn_index(dm_self)
    ed = np.sqrt(np.maximum(2 * 1024 * (1 - vdist), 0))
    assert np.allclose(ed, dist, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("which", ["self", "ab"])
def test_sum_thresh(data, which):
    import pyscamp as mp
    arr, arr2, dm_self, dm_ab = data
    if which == "self":
        dist = mp.selfjoin_sum(arr, 1024, threshold=0.90, pearson=True)
        vdist = reduce_sum(dm_self, 0.90)
    else:
        dist = mp.abjoin_sum(arr, arr2, 1024, threshold=0.90, pearson=True)
        vdist = reduce_sum(dm_ab, 0.90)
    assert dist.dtype == np.float64
    assert _compare_sum(vdist, dist, 0.90)
    # a threshold low enough that many correlations count
    rw = np.cumsum(np.random.RandomState(3).standard_normal(3000))
    s = mp.selfjoin_sum(rw, 128, threshold=0.5)
    np.testing.assert_allclose(s, reduce_sum(corr_matrix(rw, None, 128, True), 0.5),
                               rtol=1e-9, atol=1e-9)


def test_matrix_summary(data):
    import pyscamp as mp
    arr, arr2, _, dm_ab = data
    m = mp.abjoin_matrix(arr, arr2, 1024, threshold=0.125, mwidth=10, mheight=5,
                         pearson=True)
    assert m.shape == (5, 10) and m.dtype == np.float32
    ref = reduce_matrix(dm_ab, 5, 10, 0.125)
    ref = np.where(ref <= -2, np.nan, ref)
    assert np.allclose(ref, m, equal_nan=True, atol=2 ** -23)
    rw = np.cumsum(np.random.RandomState(4).standard_normal(4000))
    ms = mp.selfjoin_matrix(rw, 100, pearson=True)  # default 50 x 50
    ref = reduce_matrix(upper_pairs(corr_matrix(rw, None, 100, True)), 50, 50, 0.0)
    ref = np.where(ref <= -2, np.nan, ref)
    assert np.allclose(ref, ms, equal_nan=True, atol=2 ** -23)


def test_threads_gpus_and_precisions(data):
    import pyscamp as mp
    arr = data[0]
    d1, i1 = mp.selfjoin(arr, 1024, threads=1)
    d2, i2 = mp.selfjoin(arr, 1024, threads=2, gpus=[0])
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2, equal_nan=True)
    dm, im = mp.selfjoin(arr, 1024, precision="mixed", pearson=True)
    dd, _ = mp.selfjoin(arr, 1024, precision="double", pearson=True)
    assert np.nanmax(np.abs(dm - dd)) <= 1e-4
    du, iu = mp.selfjoin(arr, 1024, precision="ultra")
    assert np.array_equal(iu, i1)
