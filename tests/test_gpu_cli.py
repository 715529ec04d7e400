"""The command-line tool end to end, as the reference's test/run_tests.sh
drives SCAMP's own: the sample inputs written to text, the tool run with
--window=100 at two tile sizes, and its output files checked against the
reference's sample outputs with the script's rules (test/difference.py:
NaN positions equal and |distance difference| <= 0.01; index lines differing
in at most 1 % of the rows).  Also: the aligned AB-join of a series with
itself, both AB orders, --keep_rows output, and the SUM_THRESH / 1NN /
MATRIX_SUMMARY file formats against brute force."""
import os
import subprocess

import numpy as np
import pytest

from helpers import corr_matrix, random_walk, reduce_1nn_index, reduce_matrix, reduce_sum, upper_pairs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu-scrimp-large_amd", "SCAMP")
GOLD = os.path.join(ROOT, "tests", "golden", "sample_io.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLD)


def write_series(path, a):
    with open(path, "w") as f:
        for v in a:
            f.write(f"{float(v)!r}\n")


def scamp(cwd, *args):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert p.returncode == 0, p.stdout + p.stderr
    return p


def read_floats(path):
    return np.loadtxt(path, dtype=np.float64, ndmin=1)


def check_against(cwd, mp_ref, mpi_ref, prefix="mp_columns_out"):
    mp = read_floats(cwd / prefix)
    mpi = np.loadtxt(cwd / (prefix + "_index"), dtype=np.int64, ndmin=1)
    assert mp.shape == mp_ref.shape and mpi.shape == mpi_ref.shape
    assert np.array_equal(np.isnan(mp), np.isnan(mp_ref))
    assert np.nanmax(np.abs(mp - mp_ref)) <= 0.01
    assert np.count_nonzero(mpi != mpi_ref) <= (mp_ref.size + 99) // 100


@pytest.mark.parametrize("name", ["randomwalk8K", "randomwalk16K", "randomwalk32K",
                                  "randomwalk64K_nan"])
def test_selfjoin_sample_outputs(golden, tmp_path, name):
    a = golden["in_" + name]
    write_series(tmp_path / "in.txt", a)
    for tile in (2000000, 8000):
        if tile >= 2 * a.size:
            continue
        scamp(tmp_path, "--window=100", f"--max_tile_size={tile}", "--input_a_file_name=in.txt")
        check_against(tmp_path, golden["mp_" + name], golden["mpi_" + name])


def test_aligned_abjoin_of_a_series_with_itself(golden, tmp_path):
    name = "randomwalk64K_nan"
    write_series(tmp_path / "in.txt", golden["in_" + name])
    scamp(tmp_path, "--aligned", "--window=100", "--input_a_file_name=in.txt",
          "--input_b_file_name=in.txt")
    check_against(tmp_path, golden["mp_" + name], golden["mpi_" + name])


def test_abjoin_sample_outputs_and_keep_rows(golden, tmp_path):
    a, b = golden["in_randomwalk16K"], golden["in_randomwalk32K"]
    write_series(tmp_path / "a.txt", a)
    write_series(tmp_path / "b.txt", b)
    scamp(tmp_path, "--window=100", "--input_a_file_name=a.txt", "--input_b_file_name=b.txt")
    check_against(tmp_path, golden["mp_randomwalk16K_randomwalk32K"],
                  golden["mpi_randomwalk16K_randomwalk32K"])
    scamp(tmp_path, "--window=100", "--input_a_file_name=b.txt", "--input_b_file_name=a.txt")
    check_against(tmp_path, golden["mp_randomwalk32K_randomwalk16K"],
                  golden["mpi_randomwalk32K_randomwalk16K"])
    # rows kept separately: the profile of b's windows over a
    scamp(tmp_path, "--window=100", "--keep_rows", "--input_a_file_name=a.txt",
          "--input_b_file_name=b.txt", "--output_b_file_name=rows",
          "--output_b_index_file_name=rows_index")
    check_against(tmp_path, golden["mp_randomwalk16K_randomwalk32K"],
                  golden["mpi_randomwalk16K_randomwalk32K"])
    check_against(tmp_path, golden["mp_randomwalk32K_randomwalk16K"],
                  golden["mpi_randomwalk32K_randomwalk16K"], prefix="rows")


def test_other_profile_types_and_pearson_output(tmp_path):
    a = random_walk(2000, 41, nan_runs=2)
    m = 64
    write_series(tmp_path / "in.txt", a)
    D = corr_matrix(a, None, m, True)
    # 1NN_INDEX as Pearson correlations
    scamp(tmp_path, f"--window={m}", "--input_a_file_name=in.txt", "--output_pearson")
    ref, _ = reduce_1nn_index(D)
    ref = np.where(ref <= -2, np.nan, ref)
    got = read_floats(tmp_path / "mp_columns_out")
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.nanmax(np.abs(got - ref)) <= 1e-6
    # 1NN: the same values, no index file needed
    scamp(tmp_path, f"--window={m}", "--input_a_file_name=in.txt", "--output_pearson",
          "--profile_type=1NN", "--output_a_file_name=nn")
    got1 = read_floats(tmp_path / "nn")
    assert np.array_equal(np.isnan(got1), np.isnan(got))
    assert np.nanmax(np.abs(got1 - got)) == 0.0
    # SUM_THRESH
    scamp(tmp_path, f"--window={m}", "--input_a_file_name=in.txt",
          "--profile_type=SUM_THRESH", "--threshold=0.5", "--output_a_file_name=sums")
    np.testing.assert_allclose(read_floats(tmp_path / "sums"), reduce_sum(D, 0.5),
                               rtol=1e-8, atol=1e-6)
    # MATRIX_SUMMARY: reduced_height lines of reduced_width values
    scamp(tmp_path, f"--window={m}", "--input_a_file_name=in.txt", "--output_pearson",
          "--profile_type=MATRIX_SUMMARY", "--reduced_height=5", "--reduced_width=7",
          "--output_a_file_name=summary")
    lines = (tmp_path / "summary").read_text().strip("\n").split("\n")
    assert len(lines) == 5 and all(len(l.split()) == 7 for l in lines)
    got = np.array([[float(x) for x in l.split()] for l in lines])
    ref = reduce_matrix(upper_pairs(D), 5, 7, 0.0)
    ref = np.where(ref <= -2, np.nan, ref)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.nanmax(np.abs(got - ref)) <= 1e-6


def test_mixed_precision_flag_and_gpu_list(golden, tmp_path):
    name = "randomwalk16K"
    write_series(tmp_path / "in.txt", golden["in_" + name])
    scamp(tmp_path, "--window=100", "--mixed_precision", "--gpus=0", "--anchor_interval=1024",
          "--input_a_file_name=in.txt")
    check_against(tmp_path, golden["mp_" + name], golden["mpi_" + name])
