"""The mixed form (compensated FP32 recurrence, FP64 statistics, seeds,
re-anchoring and refinement) against the oracle's FP64 mode -- SCAMP's CPU
double path with its stored correlations kept in FP64
(oracle.join(..., fp64=True); src/cpu_kernels.cpp:146-149).

Bar, for every 1NN_INDEX result:
  * the NaN-window masks are identical;
  * max |corr_mixed - corr_fp64| <= 1e-5 over all valid entries;
  * index agreement >= 99.99 %, a disagreement that is a tie under the
    reference's rule (test/test_common.py:3-17: the two candidates' exact FP64
    correlations equal after FP32 rounding, or within 1e-9) counting as
    agreement (inputs of fewer than 10^4 entries: at most one non-tie
    disagreement);
  * every disagreement is a near-tie: the exact FP64 correlation of the
    mixed form's match is within 1e-5 of the FP64 best.
The raw and tie-rule agreement, max |dcorr| and the largest gap are printed
for every case.  These random-walk inputs must not trip the mixed form's
FP64 audit (the join stays in the mixed form).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import golden, random_walk, reduced_precision_errors
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

CORR_BAR = 1e-5
AGREE_BAR = 0.9999


def mixed_join(a, b, m, anchor=0, precision="mixed", **kw):
    p, pb = C.join(a, b, C.make_params(m, "1NN_INDEX", precision=precision,
                                       anchor_interval=anchor, has_b=b is not None, **kw))
    out = [C.split_mp_entry(p)]
    if pb is not None:
        out.append(C.split_mp_entry(pb))
    return out


def check_against_fp64(a, b, m, mixed, ref64, label, audit_clean=True):
    """mixed = (corr, idx) of a's windows matched in b (b None: a itself);
    ref64 = the oracle's FP64-mode (corr, idx) of the same profile."""
    c, i = mixed
    rc, ri = ref64
    assert np.array_equal(c > -2, rc > -2), label
    r = reduced_precision_errors(a, b, m, c, i, rc, ri)
    ran = C.context(0).last_precision()
    print(f"{label}: entries {r['entries']}, max |dcorr| {r['max_dcorr']:.3g}, index agreement "
          f"{r['raw_agreement']:.6f} raw / {r['agreement']:.6f} with ties "
          f"({r['disagreements']} disagreements, {r['non_tie']} not ties, largest exact gap "
          f"{r['max_gap']:.3g}); ran in {ran[0]}, audit failures {ran[1]}", flush=True)
    if audit_clean:
        assert ran == ("mixed", 0) or ran[0] == "single" and ran[1] == 0, label
    assert r["max_dcorr"] <= CORR_BAR, label
    assert r["max_gap"] <= CORR_BAR, label
    if r["entries"] >= 10000:
        assert r["agreement"] >= AGREE_BAR, label
    else:
        assert r["non_tie"] <= 1, label
    return r


def k2_prefix():
    """The first 2^17 samples of the n = 2^22 benchmark series (random walk,
    PCG64 seed 2, 0.05 % NaN in runs of 1-50)."""
    import bench
    return bench.make_series()[: 1 << 17]


CASES = {
    "walk_2^17_m256": lambda: (random_walk(1 << 17, 21, 20), None, 256),
    "k2_prefix_2^17_m1024": lambda: (k2_prefix(), None, 1024),
    "ab_40000x70000_m300": lambda: (random_walk(40000, 31, 5), random_walk(70000, 32, 5), 300),
}

_ref = {}


def oracle_fp64(name):
    """The oracle's FP64-mode profiles of every case, computed once, the cases
    in parallel threads (the oracle is single-threaded C and releases the GIL)."""
    if not _ref:
        with ThreadPoolExecutor(len(CASES)) as ex:
            futs = {}
            for k, make in CASES.items():
                a, b, m = make()
                futs[k] = (a, b, m, ex.submit(O.join, a, b, m, fp64=True,
                                              computing_rows=True,
                                              keep_rows_separate=b is not None))
            for k, (a, b, m, f) in futs.items():
                _ref[k] = (a, b, m, f.result())
    return _ref[name]


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_against_fp64_oracle(name):
    a, b, m, ref = oracle_fp64(name)
    if b is None:
        check_against_fp64(a, None, m, mixed_join(a, None, m)[0], ref, name)
    else:
        ma, mb = mixed_join(a, b, m, computing_rows=True, keep_rows_separate=True)
        check_against_fp64(a, b, m, ma, (ref[0], ref[1]), name + " (A in B)")
        check_against_fp64(b, a, m, mb, (ref[2], ref[3]), name + " (B in A)")


@pytest.mark.parametrize("anchor", [256, 16384])
def test_mixed_anchor_intervals_against_fp64_oracle(anchor):
    a, b, m, ref = oracle_fp64("walk_2^17_m256")
    check_against_fp64(a, None, m, mixed_join(a, None, m, anchor=anchor)[0], ref,
                       f"walk_2^17_m256 R={anchor}")


@pytest.mark.parametrize("name,windows", [("randomwalk1K_nan", (5, 32, 100)),
                                          ("poorly_conditioned_test", (5, 25, 100, 500))])
@pytest.mark.parametrize("precision", ["mixed", "single"])
def test_reference_static_inputs_reduced_precision(name, windows, precision):
    """The reference's static test inputs (test/run_tests.py:12) in the mixed
    form and under SINGLE, which names the same form.  These are short
    windows of quantized or NaN-broken series; where the FP32 sweep drifts
    past its bound the engine's audit recomputes the join in FP64 (printed),
    and the result must meet the bar either way."""
    a = golden()["in_" + name]
    for m in windows:
        ref = O.join(a, None, m, fp64=True)
        check_against_fp64(a, None, m, mixed_join(a, None, m, precision=precision)[0], ref,
                           f"{name} m={m} {precision}", audit_clean=False)


def test_single_is_the_mixed_form():
    """SINGLE names the mixed form: the same profile bit for bit."""
    a = random_walk(20000, 77, 3)
    ps, _ = C.join(a, None, C.make_params(128, precision="single"))
    pm, _ = C.join(a, None, C.make_params(128, precision="mixed"))
    assert np.array_equal(ps, pm)
