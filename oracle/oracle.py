"""ctypes access to the oracle libraries (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only to check results against.  It wraps

* ``liboracle.so``         -- the C restatement of SCAMP's CPU double path
  (oracle/scamp_oracle.c, which cites the reference's files line by line);
* ``_ref/libscamp_ref.so`` -- SCAMP's own CPU sources compiled by
  oracle/Makefile with a small C driver (oracle/ref_driver.cpp).

Profile layouts returned here are "wide": float64 correlations plus int64
indices (-1 where there is no match), so that the FP64 mode of the
restatement can be compared without FP32 rounding.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PTYPES = {"1NN_INDEX": 1, "SUM_THRESH": 2, "1NN": 6}
# SCAMPPrecisionType (src/common.h:53-58) of the modes SCAMP's CPU path runs
PRECISIONS = {"double": 3, "ultra": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


class _OraArgs(ctypes.Structure):
    # mirrors ora_args in scamp_oracle.h
    _fields_ = [
        ("window", ctypes.c_int64),
        ("max_tile_size", ctypes.c_int64),
        ("profile_type", ctypes.c_int32),
        ("has_b", ctypes.c_int32),
        ("computing_rows", ctypes.c_int32),
        ("computing_cols", ctypes.c_int32),
        ("keep_rows_separate", ctypes.c_int32),
        ("is_aligned", ctypes.c_int32),
        ("global_start_row", ctypes.c_int64),
        ("global_start_col", ctypes.c_int64),
        ("threshold", ctypes.c_double),
        ("num_workers", ctypes.c_int32),
        ("fp64_dist", ctypes.c_int32),
    ]


def build():
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.ora_stats.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64] + [_dp] * 5
        lib.ora_stats.restype = ctypes.c_int
        lib.ora_qt_seed.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, _dp]
        lib.ora_qt_seed.restype = None
        lib.ora_join.argtypes = [_dp, ctypes.c_int64, _dp, ctypes.c_int64,
                                 ctypes.POINTER(_OraArgs), _dp, _ip, _dp, _ip]
        lib.ora_join.restype = ctypes.c_int
        _lib = lib
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libscamp_ref.so"))


def _load_ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(os.path.join(_HERE, "_ref", "libscamp_ref.so"))
        lib.ref_join.argtypes = [
            _dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
            ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_int, ctypes.c_int,
            _dp, _ip, _dp, _ip]
        lib.ref_join.restype = ctypes.c_int
        _ref = lib
    return _ref


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def stats(T, m):
    """SCAMP's CPU statistics: (clean, means, inverse norms, df, dg)."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    n = T.size
    ns = n - m + 1
    clean = np.empty(n)
    out = [np.empty(ns) for _ in range(4)]
    if _load().ora_stats(_p(T), n, m, _p(clean), *[_p(o) for o in out]):
        raise ValueError("invalid statistics arguments")
    return (clean, *out)


def qt_seed(T, Q, m):
    """Direct centred sliding dot products of T against Q[:m]."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    nq = T.size - m + 1
    out = np.empty(nq)
    _load().ora_qt_seed(_p(T), _p(Q), nq, m, _p(out))
    return out


def _alloc(ptype, n):
    init = 0.0 if ptype == "SUM_THRESH" else -2.0
    return np.full(n, init), np.full(n, -1, dtype=np.int64)


def join(a, b, m, profile_type="1NN_INDEX", max_tile_size=1 << 17,
         computing_rows=None, computing_cols=True, keep_rows_separate=False,
         is_aligned=False, global_start_row=-1, global_start_col=-1,
         threshold=0.0, num_workers=1, fp64=False, impl="port", precision="double"):
    """Run a join through the restatement (impl="port") or SCAMP's own
    compiled CPU path (impl="reference").

    Returns (corr_a, idx_a) -- plus (corr_b, idx_b) when keep_rows_separate.
    ``idx`` is None for 1NN and SUM_THRESH.  computing_rows defaults to the
    pyscamp choice: True for self-joins, False for AB-joins.  ``precision``
    ("double" or "ultra") selects SCAMP's CPU precision mode and applies to
    impl="reference" only; the restatement follows the DOUBLE path.
    """
    a = np.ascontiguousarray(a, dtype=np.float64)
    has_b = b is not None
    if has_b:
        b = np.ascontiguousarray(b, dtype=np.float64)
    if computing_rows is None:
        computing_rows = not has_b
    na = a.size
    nb = b.size if has_b else na
    if m < 1 or na < m or nb < m:
        raise RuntimeError("oracle join failed with code 7 (series shorter than the window)")
    ca, ia = _alloc(profile_type, na - m + 1)
    cb, ib = _alloc(profile_type, nb - m + 1) if keep_rows_separate else (None, None)
    pt = PTYPES[profile_type]
    if precision != "double" and impl != "reference":
        raise ValueError("the restatement follows SCAMP's DOUBLE path only")
    if impl == "reference":
        if fp64:
            raise ValueError("the compiled reference has no FP64-output mode")
        rc = _load_ref().ref_join(
            _p(a), na, _p(b) if has_b else None, nb, m, pt, max_tile_size,
            int(computing_rows), int(computing_cols), int(keep_rows_separate),
            int(is_aligned), global_start_row, global_start_col, threshold,
            num_workers, PRECISIONS[precision], _p(ca), _p(ia, _ip), _p(cb), _p(ib, _ip))
    else:
        args = _OraArgs(m, max_tile_size, pt, int(has_b), int(computing_rows),
                        int(computing_cols), int(keep_rows_separate), int(is_aligned),
                        global_start_row, global_start_col, threshold, num_workers,
                        int(fp64))
        rc = _load().ora_join(_p(a), na, _p(b) if has_b else None, nb,
                              ctypes.byref(args), _p(ca), _p(ia, _ip), _p(cb),
                              _p(ib, _ip))
    if rc:
        raise RuntimeError(f"oracle join failed with code {rc}")
    if profile_type != "1NN_INDEX":
        ia = ib = None
    if keep_rows_separate:
        return ca, ia, cb, ib
    return ca, ia
