"""The batched self-join (scamp_hip_selfjoin_batch): `batch` series in one
launch must give exactly what one scamp_hip_join_device self-join per series
gives (batches this small cut their bands into the same row chunks as a single
series, so the statistics, seeds, rows and merges are the same and the
profiles bit-identical), for FP64 and the mixed form, 1NN_INDEX and 1NN, with
NaN runs, with prior profiles that must be kept where they are better, and for
series shorter than one 512-diagonal band as well as several bands long.  A
batch with enough bands to fill the GPU is not cut into chunks; it is checked
against the per-series joins to FP64 rounding.  The per-series join is itself
checked against the oracle by the parity tests."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import random_walk
from scamp_amd import _capi as C

pytestmark = pytest.mark.gpu


def _per_series(X, m, profile_type, precision, prior):
    B, n = X.shape
    out = prior.clone()
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    params = C.make_params(m, profile_type, precision=precision)
    for b in range(B):
        C.join_device(T[b].data_ptr(), n, 0, 0, params, out[b].data_ptr())
    return out


def _batched(X, m, profile_type, precision, prior):
    B, n = X.shape
    out = prior.clone()
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    C.selfjoin_batch_device(T.data_ptr(), n, B,
                            C.make_params(m, profile_type, precision=precision),
                            out.data_ptr())
    return out


def _prior(profile_type, B, ns, X, m, seed):
    """Empty profiles, with a few seeded entries better than any real match
    (corr 1, a marker index) that the join must keep."""
    init = C.init_profile(profile_type, ns)
    P = np.tile(init, (B, 1))
    rng = np.random.default_rng(seed)
    for b in range(B):
        cols = rng.choice(ns, 5, replace=False)
        if profile_type == "1NN_INDEX":
            e = np.zeros(1, np.uint64)
            e.view(np.float32)[0] = 1.0
            e.view(np.uint32)[1] = 7
            P[b, cols] = e[0]
        else:
            P[b, cols] = 1.0
    t = torch.tensor(P.view(np.int64) if P.dtype == np.uint64 else P, device="cuda")
    return t


@pytest.mark.parametrize("B,n,m,precision,profile_type", [
    (3, 700, 32, "double", "1NN_INDEX"),      # one band per series
    (17, 2500, 64, "double", "1NN_INDEX"),    # several bands
    (5, 5000, 100, "mixed", "1NN_INDEX"),
    (4, 3000, 50, "double", "1NN"),
    (6, 1800, 40, "mixed", "1NN"),
    (1, 4000, 128, "double", "1NN_INDEX"),
])
def test_batch_equals_per_series_joins(B, n, m, precision, profile_type):
    X = np.stack([random_walk(n, 500 + 13 * b, nan_runs=b % 3) for b in range(B)])
    ns = n - m + 1
    prior = _prior(profile_type, B, ns, X, m, B * n)
    want = _per_series(X, m, profile_type, precision, prior)
    got = _batched(X, m, profile_type, precision, prior)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    # the seeded prior entries survive (nothing beats corr 1 at index 7)
    empty = torch.tensor(C.init_profile(profile_type, 1)[:1].view(
        np.int64 if profile_type == "1NN_INDEX" else np.float32), device="cuda")
    seeded = prior != empty
    assert int(seeded.sum()) == 5 * B
    assert torch.equal(got[seeded], prior[seeded])


def test_batch_argument_errors():
    lib = C.load()
    ctx = C.context(0)
    T = torch.zeros(2, 1000, dtype=torch.float64, device="cuda")
    out = torch.zeros(2, 1000, dtype=torch.int64, device="cuda")
    p = C.make_params(50)
    rc = lib.scamp_hip_selfjoin_batch(ctx.handle, ctypes.c_void_p(T.data_ptr()), 1000, 0,
                                      ctypes.byref(p), ctypes.c_void_p(out.data_ptr()))
    assert rc == 16 and "batch" in C.last_error()
    p = C.make_params(50, "SUM_THRESH")
    rc = lib.scamp_hip_selfjoin_batch(ctx.handle, ctypes.c_void_p(T.data_ptr()), 1000, 2,
                                      ctypes.byref(p), ctypes.c_void_p(out.data_ptr()))
    assert rc == 1
    p = C.make_params(50, has_b=True)
    rc = lib.scamp_hip_selfjoin_batch(ctx.handle, ctypes.c_void_p(T.data_ptr()), 1000, 2,
                                      ctypes.byref(p), ctypes.c_void_p(out.data_ptr()))
    assert rc == 16 and "self-join" in C.last_error()
    p = C.make_params(2000)
    rc = lib.scamp_hip_selfjoin_batch(ctx.handle, ctypes.c_void_p(T.data_ptr()), 1000, 2,
                                      ctypes.byref(p), ctypes.c_void_p(out.data_ptr()))
    assert rc == 16
    p = C.make_params(50, keep_rows_separate=True)
    rc = lib.scamp_hip_selfjoin_batch(ctx.handle, ctypes.c_void_p(T.data_ptr()), 1000, 2,
                                      ctypes.byref(p), ctypes.c_void_p(out.data_ptr()))
    assert rc == 16 and "one profile" in C.last_error()


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("rows,cols", [(False, True), (True, False)])
def test_batch_honours_row_and_column_flags(rows, cols, precision):
    """A self-join computing only columns (or only rows) keeps, for each
    subsequence, its best match on one side: the batch must agree with the
    single joins under the same flags."""
    B, n, m = 4, 1500, 40
    X = np.stack([random_walk(n, 900 + b, nan_runs=b % 2) for b in range(B)])
    ns = n - m + 1
    prior = torch.tensor(np.tile(C.init_profile("1NN_INDEX", ns), (B, 1)).view(np.int64),
                         device="cuda")
    T = torch.tensor(X, dtype=torch.float64, device="cuda")
    params = C.make_params(m, computing_rows=rows, computing_columns=cols,
                           precision=precision)
    want = prior.clone()
    torch.cuda.synchronize()
    for b in range(B):
        C.join_device(T[b].data_ptr(), n, 0, 0, params, want[b].data_ptr())
    got = prior.clone()
    C.selfjoin_batch_device(T.data_ptr(), n, B, params, got.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(got, want)


def test_large_batch_without_chunks_matches_per_series():
    """1600 series of 2500 samples (m = 64): 5 bands each, 8000 bands in all,
    enough to fill the GPU without cutting bands into row chunks, while each
    series alone is cut.  Correlations agree to FP32 rounding of values that
    differ by FP64 rounding; an index may differ only at a tie."""
    B, n, m = 1600, 2500, 64
    rng = np.random.default_rng(5)
    X = np.cumsum(rng.standard_normal((B, n)), axis=1)
    ns = n - m + 1
    prior = torch.tensor(np.tile(C.init_profile("1NN_INDEX", ns), (B, 1)).view(np.int64),
                         device="cuda")
    got = _batched(X, m, "1NN_INDEX", "double", prior).cpu().numpy().view(np.uint64)
    pick = rng.choice(B, 12, replace=False)
    want = _per_series(X[pick], m, "1NN_INDEX", "double",
                       prior[:12].clone()).cpu().numpy().view(np.uint64)
    for q, b in enumerate(pick):
        cg, ig = C.split_mp_entry(got[b])
        cw, iw = C.split_mp_entry(want[q])
        assert np.max(np.abs(cg.astype(np.float64) - cw)) <= 1.2e-7
        diff = ig != iw
        assert np.mean(diff) < 1e-3
        if diff.any():
            assert np.max(np.abs(cg[diff].astype(np.float64) - cw[diff])) <= 1.2e-7
