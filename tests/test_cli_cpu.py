"""The command-line tool's argument handling and input reader, on the CPU:
the checks and messages of src/main.cpp:102-170 (they fail before any GPU
work), the reader's token semantics (src/scamp_utils.cpp:28-74), and the
no-device failure of a valid invocation."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu-scrimp-large_amd", "SCAMP")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="SCAMP tool not built")


def run(*args, cwd=None):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60, cwd=cwd)
    return p.returncode, p.stdout + p.stderr


def test_argument_errors(tmp_path):
    f = tmp_path / "ts.txt"
    f.write_text("".join(f"{v}\n" for v in range(100)))
    cases = [
        (["--window=2"], "Subsequence length must be at least 3"),
        (["--window=10", "--max_tile_size=512"], "max tile size must be at least 1024"),
        (["--window=600", "--max_tile_size=1024"], "at least 2x larger than the window"),
        (["--window=10"], "primary input filename must be specified"),
        (["--window=10", "--mixed_precision", "--single_precision"],
         "only one precision flag can be enabled at a time"),
        (["--window=200", f"--input_a_file_name={f}"],
         "window size must be smaller than the timeseries length"),
        (["--window=10", "--bogus=3"], "unknown command line flag 'bogus'"),
        (["--window=ten"], "illegal value 'ten'"),
        (["--window=10", f"--input_a_file_name={f}", "--no_gpu"], "GPU only"),
        (["--window=10", f"--input_a_file_name={tmp_path / 'missing.txt'}"], "Unable to open"),
    ]
    for args, msg in cases:
        rc, out = run(*args)
        assert rc == 1, (args, out)
        assert msg in out, (args, out)


def test_flag_syntax_variants(tmp_path):
    # --name value, -name=value and --noname are accepted like gflags
    f = tmp_path / "ts.txt"
    f.write_text("".join(f"{v}\n" for v in range(100)))
    rc, out = run("--window", "200", "-input_a_file_name=" + str(f), "--nokeep_rows")
    assert rc == 1 and "window size must be smaller" in out


def test_reader_drops_an_unterminated_last_value(tmp_path):
    """The reference keeps a value only when a character follows it: with no
    final line end, 3 tokens give a 2-sample series (too short for m = 3)."""
    f = tmp_path / "ts.txt"
    f.write_text("1\n2\n3")
    rc, out = run("--window=3", f"--input_a_file_name={f}")
    assert rc == 1 and "window size must be smaller" in out
    f.write_text("1\n\n2\n  3\n")  # blank lines are skipped
    rc, out = run("--window=3", f"--input_a_file_name={f}")
    assert "window size must be smaller" not in out


def test_reader_rejects_garbage(tmp_path):
    f = tmp_path / "ts.txt"
    f.write_text("1\n2\nxyz\n4\n")
    rc, out = run("--window=3", f"--input_a_file_name={f}")
    assert rc == 1 and "FATAL ERROR: invalid argument" in out


def test_valid_invocation_without_a_device_fails_loudly(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "gpu-scrimp-large_amd"))
    from scamp_amd import _capi as C
    if C.device_count() > 0:
        pytest.skip("a HIP device is visible")
    f = tmp_path / "ts.txt"
    f.write_text("".join(f"{v * 0.37 % 1.0}\n" for v in range(300)))
    rc, out = run("--window=20", f"--input_a_file_name={f}", cwd=tmp_path)
    assert rc == 1 and "no HIP device is visible" in out
    assert not (tmp_path / "mp_columns_out").exists()
