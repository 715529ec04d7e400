"""The long-chunk sweep kernels (the two-stage 1NN filter: FP64 with the
carried first-stage factor, and FP32) on short inputs.  Joins of short series
are cut into short row chunks and run the short-chunk kernel; with
SCAMP_HIP_NO_CHUNKS set the library cuts no chunks, so these joins run the
kernels the benchmark-sized joins use.  Bars: the oracle's 1NN_INDEX (SCAMP's
CPU double path) under the reference's tie rule for the FP64 form, as the
other parity tests; the mixed form against the oracle's FP64 mode with the
mixed bar of test_gpu_mixed_vs_oracle.py (max |dcorr| <= 1e-5, index
agreement >= 99.99 % with ties counting, every disagreement a near-tie).
"""
import numpy as np
import pytest

from helpers import check_1nn_index, random_walk
from oracle import oracle as O
from scamp_amd import _capi as C
from test_gpu_mixed_vs_oracle import check_against_fp64, mixed_join

pytestmark = pytest.mark.gpu

CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


@pytest.fixture
def no_chunks(monkeypatch):
    monkeypatch.setenv("SCAMP_HIP_NO_CHUNKS", "1")


def split(p):
    return C.split_mp_entry(p)


@pytest.mark.parametrize("n,m,nan_runs", [(1600, 40, 1), (5000, 100, 2), (20000, 64, 3)])
def test_long_chunk_selfjoin_matches_oracle(no_chunks, n, m, nan_runs):
    a = random_walk(n, 3 * n + m, nan_runs)
    p, _ = C.join(a, None, C.make_params(m))
    c, i = split(p)
    oc, oi = O.join(a, None, m)
    assert check_1nn_index(a, None, m, c, i, oc, oi) == CLEAN


def test_long_chunk_flat_windows_against_brute_force(no_chunks):
    # windows inside a flat stretch can never match (their columns' first
    # stage has an infinite bound); against the brute-force matrix, as SCAMP's
    # sequential SSD update can leave a residue above its flatness bound there
    from helpers import corr_matrix, reduce_1nn_index
    a = random_walk(3000, 5, 1)
    a[1000:1120] = 2.5
    p, _ = C.join(a, None, C.make_params(40))
    rc, ri = reduce_1nn_index(corr_matrix(a, None, 40, True))
    assert check_1nn_index(a, None, 40, *split(p), rc, ri) == CLEAN


@pytest.mark.parametrize("na,nb,m", [(6000, 9000, 64), (9000, 2500, 50)])
def test_long_chunk_abjoin_both_profiles_match_oracle(no_chunks, na, nb, m):
    a = random_walk(na, na + 5, 1)
    b = random_walk(nb, nb + 6, 1)
    pa, pb = C.join(a, b, C.make_params(m, has_b=True, computing_rows=True,
                                        keep_rows_separate=True))
    oc, oi, ocb, oib = O.join(a, b, m, computing_rows=True, keep_rows_separate=True)
    assert check_1nn_index(a, b, m, *split(pa), oc, oi) == CLEAN
    assert check_1nn_index(b, a, m, *split(pb), ocb, oib) == CLEAN


@pytest.mark.parametrize("na,nb,m", [(6000, 9000, 64), (9000, 2500, 50)])
def test_long_chunk_one_sided_abjoin_matches_oracle(no_chunks, na, nb, m):
    # A's profile only: the one-sided long-chunk kernel (no two-stage filter)
    a = random_walk(na, na + 7, 1)
    b = random_walk(nb, nb + 8, 1)
    pa, _ = C.join(a, b, C.make_params(m, has_b=True))
    oc, oi = O.join(a, b, m)
    assert check_1nn_index(a, b, m, *split(pa), oc, oi) == CLEAN


@pytest.mark.parametrize("grow,gcol", [(131072, 131072), (0, 131072)])
def test_long_chunk_aligned_abjoin_matches_oracle(no_chunks, grow, gcol):
    a = random_walk(6000, 61)
    pa, pb = C.join(a, a.copy(), C.make_params(50, has_b=True, computing_rows=True,
                                               keep_rows_separate=True, is_aligned=True,
                                               distributed_start_row=grow,
                                               distributed_start_col=gcol))
    oc, oi, ocb, oib = O.join(a, a.copy(), 50, computing_rows=True, keep_rows_separate=True,
                              is_aligned=True, global_start_row=grow, global_start_col=gcol)
    assert check_1nn_index(a, a, 50, *split(pa), oc, oi) == CLEAN
    assert check_1nn_index(a, a, 50, *split(pb), ocb, oib) == CLEAN


@pytest.mark.parametrize("scale", [2.0 ** 140, 2.0 ** -10])
def test_long_chunk_join_is_scale_invariant(no_chunks, scale):
    """Power-of-two scales change no rounding: bit-identical profiles, also
    where the FP64 first-stage factor saturates at FLT_MAX (2^140)."""
    a = random_walk(12000, 9, 2)
    p1, _ = C.join(a, None, C.make_params(128))
    p2, _ = C.join(a * scale, None, C.make_params(128))
    assert np.array_equal(p1, p2)


def test_long_chunk_and_short_chunk_kernels_agree(monkeypatch):
    a = random_walk(30000, 17, 4)
    p1, _ = C.join(a, None, C.make_params(128))  # short-chunk kernel
    monkeypatch.setenv("SCAMP_HIP_NO_CHUNKS", "1")
    p2, _ = C.join(a, None, C.make_params(128))
    p3, _ = C.join(a, None, C.make_params(128))
    assert np.array_equal(p2, p3)  # deterministic
    assert check_1nn_index(a, None, 128, *split(p2), *split(p1)) == CLEAN


def test_long_chunk_mixed_form_against_fp64_oracle(no_chunks):
    a = random_walk(1 << 16, 23, 8)
    ref = O.join(a, None, 256, fp64=True)
    check_against_fp64(a, None, 256, mixed_join(a, None, 256)[0], ref,
                       "long-chunk mixed walk_2^16_m256")
