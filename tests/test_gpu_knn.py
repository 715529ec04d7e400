"""ALL_NEIGHBORS (the k best matches of every subsequence) against brute
force: the exact FP64 distance matrix (NaN windows and the self-join
exclusion zone removed), each column's correlations above the threshold
ranked best first (FP32-rounded values; ties to the lower row).  The engine
keeps FP32 correlations, so ranks are compared by value, and a different row
at a rank must be a near-tie of the expected one.  Through the C ABI,
pyscamp (selfjoin_knn / abjoin_knn: tuples, worst first per column, as the
reference's SplitProfileKNN pops them) and the command-line tool."""
import os
import subprocess

import numpy as np
import pytest

from helpers import corr_matrix, exact_corr, random_walk
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute_topk(D, k, thr):
    """Per column j of D[row, col]: (values float32[k], rows[k]) best first,
    NaN / -1 padded."""
    n = D.shape[1]
    vals = np.full((n, k), np.nan, np.float32)
    rows = np.full((n, k), -1, np.int64)
    for j in range(n):
        col = D[:, j].astype(np.float32)
        ok = np.nonzero(col > np.float32(thr))[0]
        order = ok[np.lexsort((ok, -col[ok]))][:k]
        vals[j, :order.size] = col[order]
        rows[j, :order.size] = order
    return vals, rows


def check_lists(a, b, m, got_v, got_r, ref_v, ref_r):
    bb = a if b is None else b
    assert np.array_equal(np.isnan(got_v), np.isnan(ref_v))
    ok = ~np.isnan(ref_v)
    assert np.max(np.abs(got_v[ok] - ref_v[ok])) <= 2.5e-7
    diff = ok & (got_r != ref_r)
    if diff.any():
        cols = np.nonzero(diff)[0]
        e_got = exact_corr(bb, a, m, got_r[diff], cols)
        e_ref = exact_corr(bb, a, m, ref_r[diff], cols)
        assert np.max(np.abs(e_got - e_ref)) <= 1e-6


@pytest.mark.parametrize("n,m,k,thr,nan_runs", [(1500, 40, 5, 0.0, 0), (2200, 64, 3, 0.5, 2),
                                                (900, 16, 12, -0.2, 1), (3000, 100, 1, 0.0, 0)])
def test_selfjoin_topk_matches_brute_force(n, m, k, thr, nan_runs):
    a = random_walk(n, 300 + n, nan_runs)
    D = corr_matrix(a, None, m, True)
    prof, _ = C.join(a, None, C.make_params(m, "ALL_NEIGHBORS", threshold=thr, max_matches=k))
    gv, gr = C.knn_lists(prof, k)
    rv, rr = brute_topk(D, k, thr)
    check_lists(a, None, m, gv, gr, rv, rr)


def test_k1_equals_the_matrix_profile():
    a = random_walk(4000, 77, 3)
    m = 50
    prof, _ = C.join(a, None, C.make_params(m, "ALL_NEIGHBORS", threshold=-1.0, max_matches=1))
    gv, gr = C.knn_lists(prof, 1)
    p1, _ = C.join(a, None, C.make_params(m))
    c1, i1 = C.split_mp_entry(p1)
    ok = c1 > -1
    assert np.array_equal(np.isnan(gv[:, 0]), ~ok)
    assert np.array_equal(gv[ok, 0], c1[ok])
    assert np.array_equal(gr[ok, 0], i1[ok].astype(np.int64))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_abjoin_topk_matches_brute_force(precision):
    a = random_walk(1700, 5, 1)
    b = random_walk(2300, 6, 1)
    m, k = 48, 4
    D = corr_matrix(a, b, m, False)
    prof, _ = C.join(a, b, C.make_params(m, "ALL_NEIGHBORS", precision=precision, has_b=True,
                                         computing_rows=False, threshold=0.2, max_matches=k))
    gv, gr = C.knn_lists(prof, k)
    rv, rr = brute_topk(D, k, 0.2)
    if precision == "double":
        check_lists(a, b, m, gv, gr, rv, rr)
    else:  # FP32 recurrence: values within 1e-5, ranks may swap at near-ties
        ok = ~np.isnan(rv) & ~np.isnan(gv)
        assert np.max(np.abs(gv[ok] - rv[ok])) <= 1e-5
        assert np.mean(np.isnan(gv) == np.isnan(rv)) >= 0.999


def test_prior_lists_are_kept_and_merged():
    """A second join with the first join's lists as the prior keeps them and
    adds nothing new (every pair was already considered): identical lists."""
    a = random_walk(1200, 9)
    m, k = 32, 4
    params = C.make_params(m, "ALL_NEIGHBORS", max_matches=k)
    first, _ = C.join(a, None, params)
    again, _ = C.join(a, None, params, profile_a=first.copy())
    assert np.array_equal(first, again)


def test_pyscamp_knn_tuples():
    import pyscamp as mp
    a = random_walk(1300, 12)
    m, k = 40, 3
    out = mp.selfjoin_knn(a, m, k, threshold=0.3, pearson=True)
    assert all(len(t) == 3 for t in out)
    D = corr_matrix(a, None, m, True)
    rv, rr = brute_topk(D, k, 0.3)
    by_col = {}
    for col, row, d in out:
        by_col.setdefault(col, []).append((row, d))
    for j in range(D.shape[1]):
        got = by_col.get(j, [])
        want = [(int(rr[j, t]), float(rv[j, t])) for t in range(k) if rr[j, t] >= 0]
        assert len(got) == len(want)
        # worst first, as popped from the reference's min-heap
        assert [d for _, d in got] == sorted(d for _, d in got)
        np.testing.assert_allclose(sorted(d for _, d in got), sorted(d for _, d in want),
                                   atol=2.5e-7)
    # Euclidean output and the AB form
    b = random_walk(1600, 13)
    out_ab = mp.abjoin_knn(a, b, m, 2)
    assert {c for c, _, _ in out_ab} <= set(range(a.size - m + 1))
    assert all(d >= 0 for _, _, d in out_ab)


def test_cli_all_neighbors_file(tmp_path):
    a = random_walk(1000, 21)
    m, k = 30, 3
    with open(tmp_path / "in.txt", "w") as f:
        for v in a:
            f.write(f"{float(v)!r}\n")
    cli = os.path.join(ROOT, "gpu-scrimp-large_amd", "SCAMP")
    p = subprocess.run([cli, f"--window={m}", "--input_a_file_name=in.txt",
                        "--profile_type=ALL_NEIGHBORS", f"--max_matches_per_column={k}",
                        "--threshold=0.5", "--output_pearson"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    rows = np.loadtxt(tmp_path / "mp_columns_out", ndmin=2)
    D = corr_matrix(a, None, m, True)
    rv, rr = brute_topk(D, k, 0.5)
    assert rows.shape[0] == int(np.count_nonzero(rr >= 0))
    # per column, best first
    for j in np.unique(rows[:, 0]).astype(int):
        v = rows[rows[:, 0] == j, 2]
        assert np.all(np.diff(v) <= 0)
        np.testing.assert_allclose(v, rv[j, :v.size], atol=1e-6)
