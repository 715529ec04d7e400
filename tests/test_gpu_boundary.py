"""Boundary lifetime and threading of SCAMP::do_SCAMP (through pyscamp, which
releases the GIL around it), and the mixed form's refinement on one-sided
self-join profiles.

* do_SCAMP owns its device state for the length of one call (the reference's
  per-call workers and tiles, src/SCAMP.cpp:240-256, src/common.cpp:109-124):
  the device's free memory returns to its level before the call.
* Two host threads calling do_SCAMP at once each get the oracle's result.
* A self-join profile of columns only (partners below each window), rows only
  (partners above) or both kept apart holds the pairs of its side only; the
  mixed form's FP64 refinement of lags j - 2 .. j + 2 must keep to that side,
  which the exclusion zone alone does not enforce for m < 8 (zone <= 1).
"""
import threading

import numpy as np
import pytest

from helpers import check_1nn_index, random_walk, reduced_precision_errors
from oracle import oracle as O
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

CLEAN = {"nan_mismatch": 0, "non_tie_index": 0, "corr_off": 0}


def test_do_scamp_returns_device_memory():
    import pyscamp
    import torch
    a = random_walk(200000, 91, 3)
    pyscamp.selfjoin(a[:5000], 100)  # code objects load on first use
    pyscamp.selfjoin_sum(a[:5000], 100, threshold=0.5)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for m in (64, 256):
        pyscamp.selfjoin(a, m)
        pyscamp.selfjoin_sum(a, m, threshold=0.5)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    # the joins above use ~30 MB of device blocks each; none may stay behind
    assert free0 - free1 <= 2 << 20, (free0, free1)


def test_concurrent_do_scamp_calls():
    import pyscamp
    series = [random_walk(60000, 95 + k, 2) for k in range(2)]
    m = 128
    out = [None, None]
    errors = []
    barrier = threading.Barrier(2)

    def work(k):
        try:
            barrier.wait()
            for _ in range(3):
                out[k] = pyscamp.selfjoin(series[k], m, pearson=True)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(2):
        c, i = out[k]
        c = np.where(np.isnan(c), -2.0, c)
        oc, oi = O.join(series[k], None, m)
        assert check_1nn_index(series[k], None, m, c, i, oc, oi) == CLEAN


@pytest.mark.parametrize("m", [4, 5, 7, 12])
@pytest.mark.parametrize("mode", ["columns", "rows", "separate"])
def test_mixed_one_sided_selfjoin_profiles(m, mode):
    a = random_walk(6000, 97 + m, 1)
    kw = {"columns": dict(computing_rows=False),
          "rows": dict(computing_columns=False),
          "separate": dict(keep_rows_separate=True)}[mode]
    p, pb = C.join(a, None, C.make_params(m, precision="mixed", **kw))
    okw = {"columns": dict(computing_rows=False),
           "rows": dict(computing_cols=False),
           "separate": dict(keep_rows_separate=True)}[mode]
    ref = O.join(a, None, m, fp64=True, **okw)
    profiles = [(p, ref[0], ref[1])]
    if mode == "separate":
        profiles.append((pb, ref[2], ref[3]))
    for k, (prof, rc, ri) in enumerate(profiles):
        c, i = C.split_mp_entry(prof)
        assert np.array_equal(c > -2, rc > -2), (mode, k)
        ok = c > -2
        rows = np.nonzero(ok)[0]
        # the side each profile holds: columns see partners i <= j - m/4,
        # rows partners >= i + m/4
        below = mode == "columns" or (mode == "separate" and k == 0)
        e = m // 4
        if below:
            assert np.all(i[rows] <= rows - e), (mode, k)
        else:
            assert np.all(i[rows] >= rows + e), (mode, k)
        r = reduced_precision_errors(a, None, m, c, i, rc, ri)
        assert r["max_dcorr"] <= 1e-5 and r["max_gap"] <= 1e-5, (mode, k, r)
        assert r["agreement"] >= 0.9995, (mode, k, r)
