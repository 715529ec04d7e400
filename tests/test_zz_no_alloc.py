"""The no-allocation join (scamp_hip_join_device_noalloc) against the
allocating one (scamp_hip_join_device), through the C ABI on device buffers.

Each case runs two calls in a row through both entry points on the same
seeded inputs, the second on a shorter series (so the scratch array holds
the first call's larger buffers when the second starts), then the first size
once more.  Bars: 1NN_INDEX / 1NN / MATRIX_SUMMARY profiles bit for bit (both
entry points run the same kernels on the same work plan; keys merge with
atomic max, independent of order); SUM_THRESH sums within 1e-9 relative
(floating-point atomic adds in a scheduling-dependent order);
ALL_NEIGHBORS lists equal as (correlation, row) sets per subsequence.  A join
whose buffers exceed the scratch array is refused with INVALID_ARGS and
leaves the profile untouched.
"""
import numpy as np
import pytest

from helpers import random_walk
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu

# (name, profile type, precision, AB-join, keep_rows_separate, window, sizes)
CASES = [
    ("fp64-self-1nn_index", "1NN_INDEX", "double", False, False, 256, (200_000, 40_000)),
    ("mixed-ab-1nn_index-rows", "1NN_INDEX", "mixed", True, True, 128, (120_000, 30_000)),
    ("ultra-self-1nn", "1NN", "ultra", False, False, 64, (30_000, 6_000)),
    ("fp64-self-sum_thresh", "SUM_THRESH", "double", False, False, 64, (20_000, 5_000)),
    ("fp64-ab-matrix", "MATRIX_SUMMARY", "double", True, False, 64, (20_000, 5_000)),
    ("fp64-self-knn", "ALL_NEIGHBORS", "double", False, False, 64, (8_000, 2_000)),
]


def _device(arr, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).to(dev)


def _run(entry, a_d, na, b_d, nb, params, prof0, prof0_b, dev):
    import torch
    pa = _device(prof0, dev)
    pb = _device(prof0_b, dev) if prof0_b is not None else None
    torch.cuda.synchronize()
    entry(a_d.data_ptr(), na, b_d.data_ptr() if b_d is not None else 0, nb, params,
          pa.data_ptr(), pb.data_ptr() if pb is not None else None)
    out_a = pa.cpu().numpy().view(prof0.dtype)
    out_b = pb.cpu().numpy().view(prof0.dtype) if pb is not None else None
    return out_a, out_b


def _same(pt, x, y, k):
    if pt == "SUM_THRESH":
        np.testing.assert_allclose(x, y, rtol=1e-9, atol=1e-9)
    elif pt == "ALL_NEIGHBORS":
        xs = np.sort(x.reshape(-1, k), axis=1, order=("corr", "row"))
        ys = np.sort(y.reshape(-1, k), axis=1, order=("corr", "row"))
        assert np.array_equal(xs["corr"], ys["corr"])
        assert np.array_equal(xs["row"], ys["row"])
    else:
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


@pytest.mark.parametrize("name,pt,prec,ab,sep,m,sizes", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_join_device(name, pt, prec, ab, sep, m, sizes):
    import torch
    dev = torch.device("cuda:0")
    k, H, W = 5, 40, 30
    for call, n in enumerate((sizes[0], sizes[1], sizes[0])):
        print(f"[noalloc {name}] call {call}: n={n} m={m}", flush=True)
        a = random_walk(n, 11 + call, nan_runs=2)
        b = random_walk(n - n // 7, 23 + call, nan_runs=1) if ab else None
        na, nb = a.size, (b.size if ab else 0)
        a_d = _device(a, dev)
        b_d = _device(b, dev) if ab else None
        params = C.make_params(m, pt, precision=prec, has_b=ab, keep_rows_separate=sep,
                               threshold=0.5, matrix_height=H, matrix_width=W,
                               max_matches=k)
        ns_a = na - m + 1
        ns_b = (nb if ab else na) - m + 1
        if pt == "MATRIX_SUMMARY":
            prof0, prof0_b = C.init_profile(pt, 0, H, W), None
        else:
            prof0 = C.init_profile(pt, ns_a, max_matches=k)
            prof0_b = C.init_profile(pt, ns_b, max_matches=k) if sep else None
        ref_a, ref_b = _run(C.join_device, a_d, na, b_d, nb, params, prof0, prof0_b, dev)
        got_a, got_b = _run(C.join_device_noalloc, a_d, na, b_d, nb, params, prof0,
                            prof0_b, dev)
        _same(pt, got_a, ref_a, k)
        if sep:
            _same(pt, got_b, ref_b, k)
        # the join did something: some entry moved off its initial value
        assert not np.array_equal(got_a.view(np.uint8), prof0.view(np.uint8))


def test_noalloc_refuses_a_join_beyond_its_scratch():
    import torch
    dev = torch.device("cuda:0")
    n, m = 3_000_000, 256  # about 156 MB of FP64 buffers: above 128 MiB
    a_d = torch.zeros(n, dtype=torch.float64, device=dev)
    prof0 = C.init_profile("1NN_INDEX", n - m + 1)
    pa = _device(prof0, dev)
    torch.cuda.synchronize()
    with pytest.raises(C.ScampError, match="scratch array holds") as err:
        C.join_device_noalloc(a_d.data_ptr(), n, 0, 0, C.make_params(m), pa.data_ptr())
    assert err.value.code == 16  # INVALID_ARGS
    assert np.array_equal(pa.cpu().numpy().view(np.uint64), prof0)
