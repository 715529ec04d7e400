"""ctypes binding of the engine's C ABI (include/scamp_hip.h).

The shared library ``libscamp_hip.so`` (HIP kernels for gfx950 + C ABI) is
built in-tree by ``gpu-scrimp-large_amd/Makefile``.  Loading fails loudly when
it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscamp_hip.so")

# include/scamp_hip.h enums
PROFILE_TYPES = {"1NN_INDEX": 1, "SUM_THRESH": 2, "1NN": 6, "ALL_NEIGHBORS": 7,
                 "MATRIX_SUMMARY": 8}
# scamp_hip_match
MATCH_DTYPE = np.dtype([("corr", np.float32), ("row", np.uint32), ("col", np.uint32)])
PRECISIONS = {"single": 1, "mixed": 2, "double": 3, "ultra": 4}
STATUS = {0: "OK", 1: "FUNCTIONALITY_UNIMPLEMENTED", 2: "ILLEGAL_TYPE",
          3: "DEVICE_ERROR", 6: "DIM_INCOMPATIBLE", 16: "INVALID_ARGS",
          17: "OUT_OF_MEMORY", 18: "NO_DEVICE"}

# symbols include/scamp_hip.h declares
EXPORTS = ("scamp_hip_ctx_create", "scamp_hip_ctx_destroy", "scamp_hip_device_count",
           "scamp_hip_last_error", "scamp_hip_join", "scamp_hip_join_device",
           "scamp_hip_stats", "scamp_hip_qt_seed", "scamp_hip_last_timing",
           "scamp_hip_pearson_backward", "scamp_hip_selfjoin_batch",
           "scamp_hip_last_precision", "scamp_hip_join_device_noalloc",
           "scamp_hip_ctx_set_stream")
# SCAMP_HIP_NOALLOC_SCRATCH_BYTES: the no-allocation join's working memory per GPU
NOALLOC_SCRATCH_BYTES = 128 << 20


class Params(ctypes.Structure):
    """scamp_hip_params"""
    _fields_ = [
        ("window", ctypes.c_int64),
        ("profile_type", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("has_b", ctypes.c_int32),
        ("computing_rows", ctypes.c_int32),
        ("computing_columns", ctypes.c_int32),
        ("keep_rows_separate", ctypes.c_int32),
        ("is_aligned", ctypes.c_int32),
        ("silent_mode", ctypes.c_int32),
        ("distributed_start_row", ctypes.c_int64),
        ("distributed_start_col", ctypes.c_int64),
        ("distance_threshold", ctypes.c_double),
        ("matrix_height", ctypes.c_int64),
        ("matrix_width", ctypes.c_int64),
        ("max_tile_size", ctypes.c_int64),
        ("anchor_interval", ctypes.c_int64),
        ("max_matches_per_column", ctypes.c_int64),
    ]


class ScampError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libscamp_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`make -C gpu-scrimp-large_amd` (no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, dp, sz, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.c_int64
    lib.scamp_hip_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.scamp_hip_ctx_destroy.argtypes = [vp]
    lib.scamp_hip_ctx_set_stream.argtypes = [vp, vp]
    lib.scamp_hip_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.scamp_hip_last_error.argtypes = [ctypes.c_char_p, sz]
    lib.scamp_hip_join.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(Params), vp, vp]
    lib.scamp_hip_join_device.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(Params), vp, vp]
    lib.scamp_hip_join_device_noalloc.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(Params),
                                                  vp, vp]
    lib.scamp_hip_stats.argtypes = [vp, vp, sz, i64, vp, vp, vp, vp]
    lib.scamp_hip_qt_seed.argtypes = [vp, vp, sz, vp, sz, i64, i64, i64, i64, vp]
    lib.scamp_hip_last_timing.argtypes = [vp, dp, dp, dp]
    lib.scamp_hip_pearson_backward.argtypes = [vp, vp, sz, i64, i64, vp, vp, vp]
    lib.scamp_hip_selfjoin_batch.argtypes = [vp, vp, sz, i64, ctypes.POINTER(Params), vp]
    lib.scamp_hip_last_precision.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_uint64)]
    for name in EXPORTS:
        if hasattr(lib, name):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().scamp_hip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != 0:
        raise ScampError(rc, last_error())


def device_count() -> int:
    n = ctypes.c_int(0)
    load().scamp_hip_device_count(ctypes.byref(n))
    return n.value


class Context:
    """A device + stream (scamp_hip_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        self._h = ctypes.c_void_p()
        check(self._lib.scamp_hip_ctx_create(device, ctypes.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self._lib.scamp_hip_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """Queue this context's work on `stream` (an integer hipStream_t
        handle, e.g. torch.cuda.current_stream().cuda_stream; None or 0: the
        context's own stream).  On a caller's stream the device-buffer calls
        return once their work is queued (scamp_hip_ctx_set_stream)."""
        check(self._lib.scamp_hip_ctx_set_stream(self._h, ctypes.c_void_p(stream or None)))

    def timing(self):
        t, p, s = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self._lib.scamp_hip_last_timing(self._h, ctypes.byref(t), ctypes.byref(p),
                                              ctypes.byref(s)))
        return {"total_ms": t.value, "prep_ms": p.value, "sweep_ms": s.value}

    def last_precision(self):
        """(precision name the last join ran in, entries that failed the
        mixed form's FP64 audit)."""
        prec, drift = ctypes.c_int(), ctypes.c_uint64()
        check(self._lib.scamp_hip_last_precision(self._h, ctypes.byref(prec),
                                                 ctypes.byref(drift)))
        names = {v: k for k, v in PRECISIONS.items()}
        return names.get(prec.value, prec.value), drift.value


_ctx_cache = {}


def context(device: int = 0) -> Context:
    if device not in _ctx_cache:
        _ctx_cache[device] = Context(device)
    return _ctx_cache[device]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def make_params(window, profile_type="1NN_INDEX", precision="double", has_b=False,
                computing_rows=True, computing_columns=True, keep_rows_separate=False,
                is_aligned=False, silent_mode=True, distributed_start_row=-1,
                distributed_start_col=-1, threshold=0.0, matrix_height=50,
                matrix_width=50, max_tile_size=1 << 17, anchor_interval=0,
                max_matches=5) -> Params:
    return Params(int(window), PROFILE_TYPES[profile_type], PRECISIONS[precision], int(has_b),
                  int(computing_rows), int(computing_columns), int(keep_rows_separate),
                  int(is_aligned), int(silent_mode), int(distributed_start_row),
                  int(distributed_start_col), float(threshold), int(matrix_height),
                  int(matrix_width), int(max_tile_size), int(anchor_interval),
                  int(max_matches))


def init_profile(profile_type, n, matrix_height=50, matrix_width=50, max_matches=5):
    """Profile::Alloc initial values (src/common.cpp:283-319); ALL_NEIGHBORS:
    n * max_matches empty scamp_hip_match entries."""
    if profile_type == "ALL_NEIGHBORS":
        out = np.zeros(n * max_matches, dtype=MATCH_DTYPE)
        out["corr"] = -2.0
        out["row"] = 0xFFFFFFFF
        out["col"] = np.repeat(np.arange(n, dtype=np.uint32), max_matches)
        return out
    if profile_type == "1NN_INDEX":
        e = np.uint64((0xFFFFFFFF << 32) | int(np.float32(-2.0).view(np.uint32)))
        return np.full(n, e, dtype=np.uint64)
    if profile_type == "1NN":
        return np.full(n, -2.0, dtype=np.float32)
    if profile_type == "SUM_THRESH":
        return np.zeros(n, dtype=np.float64)
    if profile_type == "MATRIX_SUMMARY":
        return np.full(matrix_height * matrix_width, -2.0, dtype=np.float32)
    raise ValueError(profile_type)


def knn_lists(matches, k):
    """scamp_hip_match[n * k] -> (corr float32[n, k], row int64[n, k]), best
    first per subsequence; empty entries are NaN / -1."""
    m = np.asarray(matches).reshape(-1, k)
    corr = m["corr"].astype(np.float32)
    row = m["row"].astype(np.int64)
    empty = corr <= -2
    return np.where(empty, np.float32(np.nan), corr), np.where(empty, -1, row)


def split_mp_entry(p):
    """uint64 mp_entry -> (float32 correlation, int32 index)."""
    p = np.ascontiguousarray(p, dtype=np.uint64)
    corr = (p & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.float32)
    idx = (p >> np.uint64(32)).astype(np.uint32).view(np.int32)
    return corr, idx


def join(a, b, params: Params, profile_a=None, profile_b=None, device=0):
    """Host-buffer join (scamp_hip_join).  Returns (profile_a, profile_b)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64) if b is not None else None
    m = params.window
    pt = {v: k for k, v in PROFILE_TYPES.items()}[params.profile_type]
    na = a.size
    nb = b.size if b is not None else na
    if pt == "MATRIX_SUMMARY":
        la, lb = params.matrix_height * params.matrix_width, 0
    else:
        la, lb = max(na - m + 1, 1), max(nb - m + 1, 1)
    k = params.max_matches_per_column
    if profile_a is None:
        profile_a = init_profile(pt, la, params.matrix_height, params.matrix_width, k)
    sep = params.computing_rows and params.keep_rows_separate and pt != "MATRIX_SUMMARY"
    if sep and profile_b is None:
        profile_b = init_profile(pt, lb, max_matches=k)
    ctx = context(device)
    check(ctx._lib.scamp_hip_join(ctx.handle, _ptr(a), na, _ptr(b), nb if b is not None else 0,
                                  ctypes.byref(params), _ptr(profile_a),
                                  _ptr(profile_b) if sep else None))
    return profile_a, (profile_b if sep else None)


def stats(T, m, device=0):
    T = np.ascontiguousarray(T, dtype=np.float64)
    ns = T.size - m + 1
    out = [np.empty(ns) for _ in range(4)]
    ctx = context(device)
    check(ctx._lib.scamp_hip_stats(ctx.handle, _ptr(T), T.size, m, *[_ptr(o) for o in out]))
    return tuple(out)


def qt_seed(X, Y, m, row, diag0, count, device=0):
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    out = np.empty(count)
    ctx = context(device)
    check(ctx._lib.scamp_hip_qt_seed(ctx.handle, _ptr(X), X.size, _ptr(Y), Y.size, m, row,
                                     diag0, count, _ptr(out)))
    return out


def join_device(ptr_a, na, ptr_b, nb, params: Params, ptr_pa, ptr_pb=None, device=0,
                ctx=None):
    """Device-buffer join (scamp_hip_join_device) on raw device pointers
    (e.g. torch tensors' data_ptr()).  Blocks until the profiles are written
    (queues the join, on a context bound to a caller's stream)."""
    ctx = ctx or context(device)
    check(ctx._lib.scamp_hip_join_device(
        ctx.handle, ctypes.c_void_p(ptr_a), na, ctypes.c_void_p(ptr_b) if ptr_b else None,
        nb, ctypes.byref(params), ctypes.c_void_p(ptr_pa),
        ctypes.c_void_p(ptr_pb) if ptr_pb else None))


def join_device_noalloc(ptr_a, na, ptr_b, nb, params: Params, ptr_pa, ptr_pb=None,
                        device=0):
    """join_device without allocating (scamp_hip_join_device_noalloc): its
    working buffers come from the device's fixed scratch array of
    NOALLOC_SCRATCH_BYTES; larger joins raise ScampError (INVALID_ARGS)."""
    ctx = context(device)
    check(ctx._lib.scamp_hip_join_device_noalloc(
        ctx.handle, ctypes.c_void_p(ptr_a), na, ctypes.c_void_p(ptr_b) if ptr_b else None,
        nb, ctypes.byref(params), ctypes.c_void_p(ptr_pa),
        ctypes.c_void_p(ptr_pb) if ptr_pb else None))


def selfjoin_batch_device(ptr_T, n, batch, params: Params, ptr_profiles, device=0,
                          ctx=None):
    """`batch` self-joins of the float64 device series at ptr_T ([batch, n],
    back to back) in one launch (scamp_hip_selfjoin_batch); the profiles at
    ptr_profiles ([batch, n - m + 1], 1NN_INDEX or 1NN) are the prior and
    are updated in place.  Blocks until done (queues the joins, on a context
    bound to a caller's stream)."""
    ctx = ctx or context(device)
    check(ctx._lib.scamp_hip_selfjoin_batch(
        ctx.handle, ctypes.c_void_p(ptr_T), n, batch, ctypes.byref(params),
        ctypes.c_void_p(ptr_profiles)))
