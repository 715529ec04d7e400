"""Precision modes on the reference's ill-conditioned inputs.

* earthquake_precision_test100K (test/SampleInput; the input
  docs/source/cli.rst:127 names as one on which SCAMP's mixed and single
  modes fail): 95001 samples whose window standard deviation spans 44 to
  7e6.  SCAMP's own DOUBLE path is itself inexact here -- its sliding
  sum-of-squares recurrence (src/cpu_stats.cpp:103-120) loses up to 1.1e-4
  of correlation at m = 5 against textbook FP64 -- while its ULTRA path
  (brute-force statistics, src/cpu_stats.cpp:31-41,85-97) is exact to the
  FP32 rounding of its stored values.  This engine's statistics are direct
  compensated window sums, so its FP64 form is held to the textbook FP64
  correlations (1e-9) and to SCAMP's ULTRA indices (tie rule).
* ULTRA maps to the FP64 form (the statistics are already direct sums): it is
  compared with SCAMP's compiled ULTRA path on the reference's static inputs
  and goldens, index-exact under the tie rule.
* the mixed form on the same inputs against textbook FP64 at the mixed
  form's bar (max |dcorr| <= 1e-5, every index disagreement a near-tie within
  1e-5).
"""
import numpy as np
import pytest

from helpers import check_1nn_index, exact_corr, golden, reduced_precision_errors
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}
needs_ref = pytest.mark.skipif(not O.ref_available(),
                               reason="SCAMP's compiled CPU path (oracle/_ref) is not built")


def gpu_selfjoin(a, m, precision):
    p, _ = C.join(a, None, C.make_params(m, precision=precision))
    return C.split_mp_entry(p)


def textbook_check(a, m, c, i, tol):
    """Every reported pair's stored correlation against its textbook FP64
    correlation (fl32 bounds of +-tol); returns (max |stored - exact|, rows)."""
    rows = np.nonzero(c > -2)[0]
    ex = exact_corr(a, a, m, rows, i[rows])
    lo = (ex - tol).astype(np.float32)
    hi = (ex + tol).astype(np.float32)
    ok = (c[rows] >= lo) & (c[rows] <= hi)
    return float(np.max(np.abs(c[rows] - ex))), rows, ok


@needs_ref
@pytest.mark.parametrize("m", [5, 25, 100, 1000])
def test_earthquake_fp64_and_ultra(m):
    a = golden()["in_earthquake_precision_test100K"]
    cu, iu = O.join(a, None, m, impl="reference", precision="ultra", num_workers=8)
    for prec in ("double", "ultra"):
        c, i = gpu_selfjoin(a, m, prec)
        # indices: SCAMP's ULTRA path, tie rule
        assert check_1nn_index(a, None, m, c, i, cu, iu) == CLEAN, (m, prec)
        # values: FP64 correlations within 1e-9 of textbook FP64
        err, rows, ok = textbook_check(a, m, c, i, 1e-9)
        print(f"earthquake m={m} {prec}: max |stored - exact| {err:.3g}", flush=True)
        assert ok.all(), (m, prec, np.count_nonzero(~ok))


@needs_ref
@pytest.mark.parametrize("name", ["poorly_conditioned_test", "randomwalk1K_nan",
                                  "randomwalk8K", "randomwalk64K_nan"])
def test_ultra_matches_scamp_ultra(name):
    a = golden()["in_" + name]
    for m in (5, 33, 100):
        cu, iu = O.join(a, None, m, impl="reference", precision="ultra", num_workers=8)
        c, i = gpu_selfjoin(a, m, "ultra")
        assert check_1nn_index(a, None, m, c, i, cu, iu) == CLEAN, (name, m)
        # ULTRA and DOUBLE name the same form here
        cd, idd = gpu_selfjoin(a, m, "double")
        assert np.array_equal(c, cd) and np.array_equal(i, idd)


@pytest.mark.parametrize("m", [5, 25, 100, 1000])
def test_earthquake_mixed_form(m):
    """At the mixed form's bar against the FP64 form (itself held to textbook
    FP64 above).  Where the FP32 sweep drifts past its bound on this input,
    the FP64 audit recomputes the join in FP64, which the printout shows."""
    a = golden()["in_earthquake_precision_test100K"]
    c64, i64 = gpu_selfjoin(a, m, "double")
    c, i = gpu_selfjoin(a, m, "mixed")
    ran = C.context(0).last_precision()
    assert np.array_equal(c > -2, c64 > -2)
    r = reduced_precision_errors(a, None, m, c, i, c64, i64)
    print(f"earthquake m={m} mixed: max |dcorr| {r['max_dcorr']:.3g}, agreement "
          f"{r['raw_agreement']:.6f} raw / {r['agreement']:.6f} with ties, largest gap "
          f"{r['max_gap']:.3g}; ran in {ran[0]} (audit failures {ran[1]})", flush=True)
    assert r["max_dcorr"] <= 1e-5 and r["max_gap"] <= 1e-5
    assert r["agreement"] >= 0.9999
    if ran[1]:
        assert ran[0] == "double"
