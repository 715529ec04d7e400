"""Parity of the HIP engine with the oracle (SCAMP's CPU double path), the
reference's golden files and textbook FP64 computations.  All calls go
through the C ABI (libscamp_hip.so).

Bars: statistics within 1e-12 relative on valid windows (NaN masks identical);
seeds within 1e-14 of the largest seed; 1NN / 1NN_INDEX correlations equal to
the oracle's FP32 values within one FP32 rounding step (2^-23) and indices
equal except at ties (the reference's rule, test/test_common.py:3-17, with
ties also accepted when the FP64 correlations agree within 1e-9);
SUM_THRESH within 1e-12 relative of the oracle's FP64 sums; MATRIX_SUMMARY
within one FP32 rounding of the brute-force max-pool.
"""
import numpy as np
import pytest

from helpers import (check_1nn_index, corr_matrix, euclidean, exact_corr, golden,
                     random_walk, reduce_matrix, reduce_sum, upper_pairs)
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


def gpu_1nn_index(a, b, m, **kw):
    p, pb = C.join(a, b, C.make_params(m, has_b=b is not None,
                                       computing_rows=kw.pop("computing_rows", b is None),
                                       **kw))
    c, i = C.split_mp_entry(p)
    if pb is not None:
        return c, i, *C.split_mp_entry(pb)
    return c, i


# ------------------------------------------------------------ units


def test_statistics_match_oracle():
    a = random_walk(20000, 3, nan_runs=6)
    for m in (16, 100, 1024):
        mu, inv, df, dg = C.stats(a, m)
        _, omu, oinv, odf, odg = O.stats(a, m)
        assert np.array_equal(np.isnan(inv), np.isnan(oinv))
        ok = ~np.isnan(oinv)
        np.testing.assert_allclose(mu[ok], omu[ok], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(inv[ok], oinv[ok], rtol=1e-10)
        both = ok[:-1] & ok[1:]  # df/dg of steps between two valid windows
        np.testing.assert_allclose(df[:-1][both], odf[:-1][both], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(dg[:-1][both], odg[:-1][both], rtol=1e-9, atol=1e-9)


def test_exact_seeds_match_direct_dot_products():
    rng = np.random.default_rng(4)
    X = np.cumsum(rng.standard_normal(6000))
    Y = np.cumsum(rng.standard_normal(3000))
    from numpy.lib.stride_tricks import sliding_window_view as sw
    for m, row, d0, cnt in ((64, 10, 5, 700), (1024, 0, 0, 1500), (3, 7, 0, 100)):
        q = C.qt_seed(X, Y, m, row, d0, cnt)
        xw = sw(X, m)[row + d0: row + d0 + cnt]
        yw = Y[row:row + m]
        ref = ((xw - xw.mean(1, keepdims=True)) * (yw - yw.mean())).sum(1)
        assert np.max(np.abs(q - ref)) <= 1e-13 * np.max(np.abs(ref)) + 1e-12


# ------------------------------------------------------- 1NN_INDEX


@pytest.mark.parametrize("n,m,nan_runs", [(3000, 100, 2), (1000, 16, 0), (5000, 50, 4),
                                         (700, 7, 1), (600, 3, 0), (2048, 512, 0),
                                         (1500, 1, 0)])
def test_selfjoin_matches_oracle(n, m, nan_runs):
    a = random_walk(n, n * 7 + m, nan_runs)
    if m < 3:
        with pytest.raises(C.ScampError):
            gpu_1nn_index(a, None, m)
        return
    c, i = gpu_1nn_index(a, None, m)
    oc, oi = O.join(a, None, m)
    assert check_1nn_index(a, None, m, c, i, oc, oi) == CLEAN


def test_selfjoin_flat_and_tiny_inputs():
    a = random_walk(800, 5)
    a[300:420] = 2.5  # flat stretch -> windows inside it are masked
    c, i = gpu_1nn_index(a, None, 40)
    # against the brute-force matrix: SCAMP's sequential SSD update can leave a
    # rounding residue above its 1e-13 flatness bound inside a flat stretch
    # (then it divides by ~0); the direct window sums here find SSD = 0.
    from helpers import reduce_1nn_index
    rc, ri = reduce_1nn_index(corr_matrix(a, None, 40, True))
    assert check_1nn_index(a, None, 40, c, i, rc, ri) == CLEAN
    # a series exactly one window long: one subsequence, no match
    c, i = gpu_1nn_index(random_walk(40, 6), None, 40)
    assert c.size == 1 and c[0] == -2 and i[0] == -1


@pytest.mark.parametrize("na,nb,m", [(3000, 4500, 64), (5000, 1200, 100), (700, 9000, 33)])
def test_abjoin_matches_oracle(na, nb, m):
    a = random_walk(na, na, 1)
    b = random_walk(nb, nb + 1, 1)
    c, i = gpu_1nn_index(a, b, m)
    oc, oi = O.join(a, b, m)
    assert check_1nn_index(a, b, m, c, i, oc, oi) == CLEAN
    c, i, cb, ib = gpu_1nn_index(a, b, m, computing_rows=True, keep_rows_separate=True)
    oc, oi, ocb, oib = O.join(a, b, m, computing_rows=True, keep_rows_separate=True)
    assert check_1nn_index(a, b, m, c, i, oc, oi) == CLEAN
    assert check_1nn_index(b, a, m, cb, ib, ocb, oib) == CLEAN


# Offsets on SCAMP's tile grid (distributed joins place tiles on a grid):
# there its per-tile corner rule (src/tile.cpp:44-138) equals the zone
# |global col - global row| < m/4 this engine applies.
@pytest.mark.parametrize("grow,gcol", [(-1, -1), (131072, 131072), (262144, 131072),
                                       (0, 131072)])
def test_aligned_abjoin_matches_oracle(grow, gcol):
    a = random_walk(2600, 61)
    c, i = gpu_1nn_index(a, a.copy(), 50, is_aligned=True, distributed_start_row=grow,
                         distributed_start_col=gcol)
    oc, oi = O.join(a, a.copy(), 50, is_aligned=True, global_start_row=grow,
                    global_start_col=gcol)
    assert check_1nn_index(a, a, 50, c, i, oc, oi) == CLEAN


def test_prior_profile_is_read_as_best_so_far():
    a = random_walk(3000, 8)
    c, i = gpu_1nn_index(a, None, 64)
    # a prior claiming a perfect match for every entry must survive
    # (index 0: an equal correlation with a higher index cannot displace it)
    prior = np.full(c.size, np.uint64(int(np.float32(1.0).view(np.uint32))))
    p, _ = C.join(a, None, C.make_params(64), profile_a=prior.copy())
    pc, pi = C.split_mp_entry(p)
    assert np.all(pc == 1.0) and np.all(pi == 0)


def test_selfjoin_is_deterministic():
    a = random_walk(20000, 12, 3)
    p1, _ = C.join(a, None, C.make_params(256))
    p2, _ = C.join(a, None, C.make_params(256))
    assert np.array_equal(p1, p2)  # order-independent merges: bit-identical


@pytest.mark.parametrize("scale", [2.0 ** 140, 2.0 ** -10])
def test_join_is_scale_invariant(scale):
    """A power-of-two scale of the series changes no rounding in the
    statistics, seeds, recurrence or correlations (nothing overflows or
    underflows here), so the profiles are bit-identical.  At 2^140 the inverse
    norms fall below 2^-128 and the first-stage factor (a lowered reciprocal
    of in_col, carried in FP32) saturates at FLT_MAX: the bound turns loose but
    stays conservative, which this pins for the self-join and for an AB-join
    with both profiles (both run the two-stage filter)."""
    a = random_walk(20000, 12, 3)
    b = random_walk(9000, 13, 2)
    p1, _ = C.join(a, None, C.make_params(256))
    p2, _ = C.join(a * scale, None, C.make_params(256))
    assert np.array_equal(p1, p2)
    params = C.make_params(100, has_b=True, keep_rows_separate=True)
    qa, qb = C.join(a, b, params)
    ra, rb = C.join(a * scale, b * scale, params)
    assert np.array_equal(qa, ra) and np.array_equal(qb, rb)


# ---------------------------------------------------- golden files


@pytest.mark.parametrize("name", ["randomwalk8K", "randomwalk16K", "randomwalk32K",
                                  "randomwalk64K_nan"])
def test_selfjoin_reference_goldens(name):
    g = golden()
    a = g["in_" + name]
    c, i = gpu_1nn_index(a, None, 100)
    d = euclidean(c, 100)
    mp, mpi = g["mp_" + name], g["mpi_" + name]
    assert np.array_equal(np.isnan(d), np.isnan(mp))
    assert np.nanmax(np.abs(d - mp)) <= 0.01
    i = np.where(c < -1, -1, i)
    assert np.count_nonzero(i != mpi) <= mpi.size // 100
    oc, oi = O.join(a, None, 100)
    assert check_1nn_index(a, None, 100, c, i, oc, oi) == CLEAN


@pytest.mark.parametrize("a_name,b_name", [("randomwalk16K", "randomwalk32K"),
                                           ("randomwalk32K", "randomwalk16K")])
def test_abjoin_reference_goldens(a_name, b_name):
    g = golden()
    a, b = g["in_" + a_name], g["in_" + b_name]
    c, i = gpu_1nn_index(a, b, 100)
    d = euclidean(c, 100)
    mp, mpi = g[f"mp_{a_name}_{b_name}"], g[f"mpi_{a_name}_{b_name}"]
    assert np.nanmax(np.abs(d - mp)) <= 0.01
    assert np.count_nonzero(i != mpi) <= mpi.size // 100


# ------------------------------------------- other profile types


def test_1nn_matches_oracle():
    a = random_walk(4000, 13, 2)
    b = random_walk(3000, 14)
    for bb in (None, b):
        p, _ = C.join(a, bb, C.make_params(64, "1NN", has_b=bb is not None,
                                           computing_rows=bb is None))
        o, _ = O.join(a, bb, 64, "1NN")
        assert np.array_equal(p < -1, o < -1)
        assert np.max(np.abs(p.astype(np.float64) - o)) <= 2 ** -23


@pytest.mark.parametrize("thr", [0.0, 0.5, 0.9, -0.3])
def test_sum_thresh_matches_oracle_and_brute_force(thr):
    a = random_walk(4000, 15, 1)
    b = random_walk(3000, 16)
    for bb in (None, b):
        s, _ = C.join(a, bb, C.make_params(64, "SUM_THRESH", has_b=bb is not None,
                                           computing_rows=bb is None, threshold=thr))
        if thr >= 0:  # the reference's CPU path thresholds block sums (see oracle)
            o, _ = O.join(a, bb, 64, "SUM_THRESH", threshold=thr, fp64=True)
            np.testing.assert_allclose(s, o, rtol=1e-12, atol=1e-9)
        if True:
            D = corr_matrix(a[:1500], None if bb is None else bb[:1200], 64, bb is None)
            s2, _ = C.join(a[:1500], None if bb is None else bb[:1200],
                           C.make_params(64, "SUM_THRESH", has_b=bb is not None,
                                         computing_rows=bb is None, threshold=thr))
            np.testing.assert_allclose(s2, reduce_sum(D, thr), rtol=1e-9, atol=1e-8)


@pytest.mark.parametrize("na,nb,m,H,W,thr", [(1200, None, 32, 5, 10, 0.1),
                                             (1500, 1100, 50, 7, 9, 0.0),
                                             (400, None, 20, 50, 50, 0.0),
                                             (3000, 2000, 64, 10, 5, 0.125),
                                             # wide cells: whole blocks of rows
                                             # fall in one cell (self-join: the
                                             # cells on and above the diagonal;
                                             # AB: each pass on its own side)
                                             (6000, None, 64, 10, 10, 0.0),
                                             (9000, None, 100, 7, 7, 0.3),
                                             (5000, 7000, 50, 6, 4, 0.1)])
def test_matrix_summary_is_exact_max_pool(na, nb, m, H, W, thr):
    a = random_walk(na, na + 3)
    b = None if nb is None else random_walk(nb, nb + 4)
    D = corr_matrix(a, b, m, b is None)
    ref = reduce_matrix(D if b is not None else upper_pairs(D), H, W, thr)
    pm, _ = C.join(a, b, C.make_params(m, "MATRIX_SUMMARY", has_b=b is not None,
                                       threshold=thr, matrix_height=H, matrix_width=W))
    pm = pm.reshape(H, W)
    assert np.array_equal(pm < -1, ref < -1)
    ok = ref > -1
    assert np.max(np.abs(pm[ok] - ref[ok])) <= 2 ** -23


# --------------------------------------------------- the config-1 case


def test_config1_selfjoin_against_scamp_cpu_path():
    """n = 131072 random walk (PCG64 seed 1), m = 100: the plain FP64 form
    against SCAMP's CPU double path (compiled reference when built, else the
    oracle's restatement)."""
    a = np.cumsum(np.random.Generator(np.random.PCG64(1)).standard_normal(131072))
    c, i = gpu_1nn_index(a, None, 100)
    impl = "reference" if O.ref_available() else "port"
    oc, oi = O.join(a, None, 100, impl=impl, num_workers=8)
    assert check_1nn_index(a, None, 100, c, i, oc, oi) == CLEAN


# ------------------------------------ size-independent checks at n = 2^22


def test_full_size_profile_properties():
    """K2 shape (n = 2^22, m = 1024, NaN runs): every reported match is a
    valid pair outside the exclusion zone whose FP64 correlation rounds to the
    reported value, NaN windows report no match, and each sampled entry is at
    least as good as random candidates."""
    import bench
    a = bench.make_series()
    m = 1024
    p, _ = C.join(a, None, C.make_params(m))
    c, i = C.split_mp_entry(p)
    nonfinite = np.concatenate([[0], np.cumsum(~np.isfinite(a))])
    bad = (nonfinite[m:] - nonfinite[:-m]) > 0  # windows holding a NaN
    assert np.array_equal(c < -1, bad)
    rng = np.random.default_rng(0)
    rows = rng.choice(np.nonzero(~bad)[0], 4000, replace=False)
    assert np.all(np.abs(rows - i[rows]) >= m // 4)
    exact = exact_corr(a, a, m, rows, i[rows])
    assert np.max(np.abs(exact - c[rows])) <= 2 ** -24 + 1e-9
    cand = rng.integers(0, c.size, size=(rows.size, 8))
    for k in range(cand.shape[1]):
        j = cand[:, k]
        ok = (np.abs(rows - j) >= m // 4) & ~bad[j]
        cj = exact_corr(a, a, m, rows[ok], j[ok])
        assert np.all(cj <= c[rows[ok]] + 2 ** -23)


# ------------------------------------------------------- mixed form


def _mixed_vs_fp64(a, b, m, anchor):
    kw = dict(has_b=b is not None, computing_rows=b is None)
    pd, _ = C.join(a, b, C.make_params(m, **kw))
    dc, di = C.split_mp_entry(pd)
    pm, _ = C.join(a, b, C.make_params(m, precision="mixed", anchor_interval=anchor, **kw))
    mc, mi = C.split_mp_entry(pm)
    assert np.array_equal(dc < -1, mc < -1)
    ok = dc > -1
    dmax = float(np.max(np.abs(mc[ok].astype(np.float64) - dc[ok])))
    diff = np.nonzero(ok & (mi != di))[0]
    gap = 0.0
    if diff.size:
        bb = a if b is None else b
        gap = float(np.max(exact_corr(a, bb, m, diff, di[diff])
                           - exact_corr(a, bb, m, diff, mi[diff])))
    return dmax, float(np.mean(mi[ok] == di[ok])), gap


@pytest.mark.parametrize("anchor", [256, 4096, 16384])
def test_mixed_form_error_bounds(anchor):
    """Compensated FP32 recurrence: |dcorr| stays far below 1e-5 even with
    sparse anchors, and every index mismatch is a near-tie (the mixed choice
    is within 2e-6 of the FP64 best in exact FP64 correlation)."""
    dmax, agree, gap = _mixed_vs_fp64(random_walk(1 << 17, 21, 20), None, 256, anchor)
    assert dmax <= 5e-6
    assert gap <= 2e-6
    assert agree >= 0.9995


def test_mixed_form_abjoin():
    a = random_walk(40000, 31, 5)
    b = random_walk(70000, 32, 5)
    dmax, agree, gap = _mixed_vs_fp64(a, b, 300, 2048)
    assert dmax <= 5e-6 and gap <= 2e-6 and agree >= 0.999


def test_autotuned_anchor_meets_target():
    from scamp_amd import tuning
    a = random_walk(1 << 17, 22, 10)
    r = tuning.autotune_anchor(a, 512)
    assert r in tuning.CANDIDATES
    d, _, gap = tuning.mixed_error(a, 512, r)
    assert d <= 1e-5 / 4 and gap <= 1e-5 / 4


# ------------------------------------------------------- error paths


def test_invalid_arguments_return_codes():
    a = random_walk(500, 1)
    with pytest.raises(C.ScampError, match="at least 3"):
        C.join(a, None, C.make_params(2))
    with pytest.raises(C.ScampError, match="max tile size"):
        C.join(a, None, C.make_params(10, max_tile_size=512))
    with pytest.raises(C.ScampError, match="window size"):
        C.join(a[:5], None, C.make_params(10))
    with pytest.raises(C.ScampError, match="matrix height"):
        C.join(a, None, C.make_params(10, "MATRIX_SUMMARY", matrix_height=0))
    # lengths past the sweep's 32-bit offsets are refused before any buffer
    # is read (the claimed length here is far beyond the 500-sample array)
    import ctypes
    prof = C.init_profile("1NN_INDEX", a.size)
    ctx = C.context(0)
    rc = ctx._lib.scamp_hip_join(ctx.handle, C._ptr(a), ctypes.c_size_t(1 << 30), None, 0,
                                 ctypes.byref(C.make_params(10)), C._ptr(prof), None)
    assert rc == 16
    with pytest.raises(C.ScampError, match="too long"):
        C.check(rc)
