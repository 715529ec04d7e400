"""The drop-in boundary on the CPU: the C-ABI library loads and exports every
entry point include/scamp_hip.h declares; argument errors come back as status
codes / exceptions without touching a GPU; pyscamp's keyword checking
mirrors python/SCAMP_python.cpp:91-177; without a GPU the engine fails loudly
(it has no CPU path)."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT
from scamp_amd import _capi as C

HEADER = os.path.join(ROOT, "include", "scamp_hip.h")


def declared_symbols():
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\bint\s+(scamp_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.load()
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(C.EXPORTS)


def test_cpp_api_library_exports_do_scamp():
    lib = ctypes.CDLL(os.path.join(ROOT, "gpu-scrimp-large_amd", "scamp_amd", "libscamp.so"))
    # mangled SCAMP::do_SCAMP(SCAMPArgs*) and the three-argument overload
    assert hasattr(lib, "_ZN5SCAMP8do_SCAMPEPNS_9SCAMPArgsE")
    assert hasattr(lib, "_ZN5SCAMP8do_SCAMPEPNS_9SCAMPArgsERKSt6vectorIiSaIiEEi")


def test_null_arguments_are_rejected_with_status_codes():
    lib = C.load()
    p = C.make_params(10)
    a = np.zeros(100)
    rc = lib.scamp_hip_join(None, a.ctypes.data, 100, None, 0, ctypes.byref(p), a.ctypes.data,
                            None)
    assert rc == 16  # SCAMP_HIP_INVALID_ARGS
    assert "null" in C.last_error()


def test_context_without_gpu_fails_loudly():
    if C.device_count() > 0:
        pytest.skip("a GPU is visible")
    lib = C.load()
    h = ctypes.c_void_p()
    assert lib.scamp_hip_ctx_create(0, ctypes.byref(h)) == 18  # SCAMP_HIP_NO_DEVICE


def test_pyscamp_keyword_validation():
    import pyscamp as mp
    a = np.random.default_rng(0).standard_normal(500)
    with pytest.raises(ValueError, match="unknown argument: bogus"):
        mp.selfjoin(a, 32, bogus=1)
    with pytest.raises(ValueError, match="unknown argument: threshold"):
        mp.selfjoin(a, 32, threshold=0.5)  # 1NN_INDEX takes no threshold
    with pytest.raises(ValueError, match="between -1 and 1"):
        mp.selfjoin_sum(a, 32, threshold=1.5)
    with pytest.raises(ValueError, match="precision"):
        mp.selfjoin(a, 32, precision="quad")
    with pytest.raises(ValueError, match="matrix height"):
        mp.selfjoin_matrix(a, 32, mheight=0)
    with pytest.raises(ValueError, match="cpu worker threads"):
        mp.selfjoin(a, 32, threads=-1)
    # ALL_NEIGHBORS takes the threshold keyword but no matrix size
    with pytest.raises(ValueError, match="unknown argument: mheight"):
        mp.selfjoin_knn(a, 32, 5, mheight=4)
    assert mp.gpu_supported() is True


def test_pyscamp_scamp_validation_messages():
    import pyscamp as mp
    a = np.random.default_rng(0).standard_normal(500)
    with pytest.raises(RuntimeError, match="Subsequence length must be at least 3"):
        mp.selfjoin(a, 2)
    with pytest.raises(RuntimeError, match="at least 'subesequence window size'"):
        mp.abjoin(a, a[:10], 32)
    with pytest.raises(RuntimeError, match="2x larger than the window"):
        mp.selfjoin(np.zeros(200000), 70000)


def test_pyscamp_without_gpu_fails_loudly():
    if C.device_count() > 0:
        pytest.skip("a GPU is visible")
    import pyscamp as mp
    with pytest.raises(RuntimeError, match="no HIP device|no compute_resources"):
        mp.selfjoin(np.random.default_rng(1).standard_normal(300), 16)


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The Python binding's Params and match records have the C layouts
    (compiled against include/scamp_hip.h with the host C compiler)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "sz.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "scamp_hip.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu %zu\\n", sizeof(scamp_hip_params), '
                   'offsetof(scamp_hip_params, anchor_interval), '
                   'offsetof(scamp_hip_params, max_matches_per_column), '
                   'sizeof(scamp_hip_match), SCAMP_HIP_NOALLOC_SCRATCH_BYTES); '
                   'return 0; }\n')
    exe = tmp_path / "sz"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    size, off_anchor, off_k, match, scratch = (int(x) for x in out)
    assert size == ctypes.sizeof(C.Params)
    assert off_anchor == C.Params.anchor_interval.offset
    assert off_k == C.Params.max_matches_per_column.offset
    assert match == C.MATCH_DTYPE.itemsize == 12
    assert scratch == C.NOALLOC_SCRATCH_BYTES
