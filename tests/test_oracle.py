"""Pinning the oracle (CPU only).

The oracle (oracle/scamp_oracle.c) restates SCAMP's CPU double path.  It is
checked against
  * the reference's own golden files (test/SampleOutput, with the reference's
    tolerances: distance within 0.01, at most 1 % differing indices --
    test/run_tests.sh:34-44, test/difference.py:23-26);
  * SCAMP's CPU sources compiled unchanged into oracle/_ref (bit for bit),
    when that library was built;
  * textbook FP64 computations (brute-force window statistics, direct dot
    products, the full Pearson matrix of tiny inputs) and invariants (planted
    copies, the exclusion zone, NaN-window masking, shift invariance).
"""
import numpy as np
import pytest

from helpers import (check_1nn_index, corr_matrix, euclidean, golden, random_walk,
                     reduce_1nn_index, reduce_sum)
from oracle import oracle as O

# ------------------------------------------------------------ statistics


def test_stats_match_brute_force():
    a = random_walk(3000, 11)
    m = 64
    clean, mu, inv, df, dg = O.stats(a, m)
    W = np.lib.stride_tricks.sliding_window_view(a, m)
    np.testing.assert_allclose(mu, W.mean(1), rtol=1e-12, atol=1e-12)
    ssd = ((W - W.mean(1, keepdims=True)) ** 2).sum(1)
    np.testing.assert_allclose(inv, 1 / np.sqrt(ssd), rtol=1e-9)
    n = a.size - m + 1
    np.testing.assert_array_equal(df[:-1], (a[m:] - a[:n - 1]) / 2)
    np.testing.assert_allclose(dg[:-1], (a[m:] - mu[1:]) + (a[:n - 1] - mu[:-1]), rtol=1e-12,
                               atol=1e-9)
    assert df[-1] == 0 and dg[-1] == 0


def test_stats_nan_windows_and_flat_windows():
    m = 10
    a = random_walk(200, 3)
    a[50] = np.nan
    a[120:140] = 7.0  # a flat stretch: windows fully inside it are flat
    _, mu, inv, _, _ = O.stats(a, m)
    bad = np.isnan(inv)
    # exactly the windows containing sample 50 are NaN around it
    assert bad[41:51].all() and not bad[40] and not bad[51]
    # windows entirely inside the flat stretch have SSD 0 -> NaN
    assert bad[120:131].all()


def test_stats_integer_windows_exact_means():
    a = np.arange(100, dtype=np.float64) % 7
    _, mu, _, _, _ = O.stats(a, 7)
    assert np.all(mu == 3.0)  # each window of 7 consecutive residues sums to 21


# -------------------------------------------------------------- QT seed


def test_qt_seed_brute_force_and_shift_invariance():
    rng = np.random.default_rng(5)
    T = np.cumsum(rng.standard_normal(500))
    Q = np.cumsum(rng.standard_normal(80))
    m = 40
    qt = O.qt_seed(T, Q, m)
    W = np.lib.stride_tricks.sliding_window_view(T, m)
    ref = ((W - W.mean(1, keepdims=True)) * (Q[:m] - Q[:m].mean())).sum(1)
    np.testing.assert_allclose(qt, ref, rtol=1e-10, atol=1e-9)
    qt2 = O.qt_seed(T + 123.0, Q - 5.0, m)
    np.testing.assert_allclose(qt2, qt, rtol=1e-8, atol=1e-7)


# ----------------------------------------------------------------- joins


@pytest.mark.parametrize("n,m,nan_runs", [(400, 20, 0), (650, 37, 3), (300, 5, 1)])
def test_selfjoin_1nn_index_matches_brute_force(n, m, nan_runs):
    a = random_walk(n, n + m, nan_runs)
    D = corr_matrix(a, None, m, True)
    rc, ri = reduce_1nn_index(D)
    for fp64 in (False, True):
        c, i = O.join(a, None, m, max_tile_size=1024, fp64=fp64)
        bad = check_1nn_index(a, None, m, c, i, rc, ri, corr_tol=1e-9 if fp64 else 2 ** -23)
        assert bad == {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


@pytest.mark.parametrize("na,nb,m", [(500, 700, 30), (900, 350, 64)])
def test_abjoin_matches_brute_force(na, nb, m):
    a = random_walk(na, 21)
    b = random_walk(nb, 22)
    D = corr_matrix(a, b, m, False)
    rc, ri = reduce_1nn_index(D)
    c, i = O.join(a, b, m, max_tile_size=1024)
    assert check_1nn_index(a, b, m, c, i, rc, ri) == {"nan_mismatch": 0, "non_tie_index": 0,
                                                      "corr_off": 0}
    # rows kept separately = the column profile of the transposed join
    _, _, cb, ib = O.join(a, b, m, max_tile_size=1024, computing_rows=True,
                          keep_rows_separate=True)
    rcb, rib = reduce_1nn_index(D.T)
    assert check_1nn_index(b, a, m, cb, ib, rcb, rib)["non_tie_index"] == 0


def test_1nn_and_sum_thresh_match_brute_force():
    a = random_walk(600, 31)
    b = random_walk(450, 32)
    m = 25
    for bb, self_join in ((None, True), (b, False)):
        D = corr_matrix(a, bb, m, self_join)
        c, _ = O.join(a, bb, m, "1NN", max_tile_size=1024)
        ref = D.max(axis=0)
        ok = ref > -2
        np.testing.assert_allclose(c[ok], ref[ok], atol=2 ** -23)
        assert np.all(c[~ok] == -2)
        for thr in (0.0, 0.5, 0.9):
            s, _ = O.join(a, bb, m, "SUM_THRESH", max_tile_size=1024, threshold=thr, fp64=True)
            np.testing.assert_allclose(s, reduce_sum(D, thr), rtol=1e-9, atol=1e-9)


def test_planted_copy_and_exclusion_zone():
    rng = np.random.default_rng(9)
    a = rng.standard_normal(2000)
    m = 50
    a[1500:1550] = a[200:250] * 3.0 + 1.0  # affine copy: correlation exactly 1
    c, i = O.join(a, None, m, max_tile_size=1024, fp64=True)
    assert abs(c[200] - 1) < 1e-12 and i[200] == 1500
    assert abs(c[1500] - 1) < 1e-12 and i[1500] == 200
    valid = i >= 0
    assert np.all(np.abs(np.arange(c.size)[valid] - i[valid]) >= m // 4)


def test_tiling_does_not_change_the_profile():
    a = random_walk(5000, 41)
    m = 100
    c1, i1 = O.join(a, None, m, max_tile_size=1024)
    c2, i2 = O.join(a, None, m, max_tile_size=1 << 17)
    assert check_1nn_index(a, None, m, c1, i1, c2, i2) == {"nan_mismatch": 0,
                                                          "non_tie_index": 0, "corr_off": 0}


# --------------------------------------------------------- golden files


@pytest.mark.parametrize("name", ["randomwalk8K", "randomwalk16K", "randomwalk64K_nan"])
def test_oracle_against_reference_goldens(name):
    g = golden()
    a = g["in_" + name]
    c, i = O.join(a, None, 100, max_tile_size=1 << 17)
    d = euclidean(c, 100)
    mp, mpi = g["mp_" + name], g["mpi_" + name]
    assert np.array_equal(np.isnan(d), np.isnan(mp))
    assert np.nanmax(np.abs(d - mp)) <= 0.01
    i = np.where(c < -1, -1, i)
    assert np.count_nonzero(i != mpi) <= mpi.size // 100


@pytest.mark.parametrize("a_name,b_name", [("randomwalk16K", "randomwalk32K"),
                                           ("randomwalk32K", "randomwalk16K")])
def test_oracle_against_reference_ab_goldens(a_name, b_name):
    g = golden()
    c, i = O.join(g["in_" + a_name], g["in_" + b_name], 100)
    d = euclidean(c, 100)
    mp, mpi = g[f"mp_{a_name}_{b_name}"], g[f"mpi_{a_name}_{b_name}"]
    assert np.nanmax(np.abs(d - mp)) <= 0.01
    assert np.count_nonzero(i != mpi) <= mpi.size // 100


# ------------------------------------------- SCAMP's own compiled CPU path

needs_ref = pytest.mark.skipif(not O.ref_available(),
                               reason="oracle/_ref was not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("ptype", ["1NN_INDEX", "1NN"])
@pytest.mark.parametrize("ab", [False, True])
def test_oracle_bit_exact_vs_compiled_reference(ptype, ab):
    a = random_walk(3000, 51, nan_runs=2)
    b = random_walk(2500, 52) if ab else None
    r1 = O.join(a, b, 64, ptype, max_tile_size=1024, impl="port")
    r2 = O.join(a, b, 64, ptype, max_tile_size=1024, impl="reference")
    np.testing.assert_array_equal(r1[0], r2[0])
    if ptype == "1NN_INDEX":
        np.testing.assert_array_equal(r1[1], r2[1])


@needs_ref
def test_oracle_keep_rows_and_aligned_vs_compiled_reference():
    a = random_walk(2600, 61)
    b = random_walk(2600, 61)  # identical series: an aligned AB-join
    for kw in (dict(computing_rows=True, keep_rows_separate=True),
               dict(is_aligned=True),
               dict(is_aligned=True, global_start_row=3000, global_start_col=2990)):
        r1 = O.join(a, b, 50, max_tile_size=1024, impl="port", **kw)
        r2 = O.join(a, b, 50, max_tile_size=1024, impl="reference", **kw)
        for x, y in zip(r1, r2):
            np.testing.assert_array_equal(x, y)


@needs_ref
def test_oracle_sum_thresh_vs_compiled_reference():
    a = random_walk(2000, 71)
    s1, _ = O.join(a, None, 40, "SUM_THRESH", max_tile_size=1024, threshold=0.5)
    s2, _ = O.join(a, None, 40, "SUM_THRESH", max_tile_size=1024, threshold=0.5,
                   impl="reference")
    # the reference build fuses some of its additions into FMAs
    np.testing.assert_allclose(s1, s2, rtol=1e-12, atol=1e-12)


def test_oracle_rejects_invalid_arguments():
    a = random_walk(100, 1)
    with pytest.raises(RuntimeError):
        O.join(a, None, 2)  # window < 3
    with pytest.raises(RuntimeError):
        O.join(a, None, 10, max_tile_size=512)  # tile < 1024
    with pytest.raises(RuntimeError):
        O.join(a[:5], None, 10)  # series shorter than the window
