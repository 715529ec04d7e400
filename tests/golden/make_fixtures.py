"""Write the golden fixtures used by the parity tests.

Run once in a tree that has the GPU-SCRIMP-LARGE reference mounted at
/root/reference (the fixtures are committed; nothing at test time reads the
reference tree):

    python tests/golden/make_fixtures.py

It converts the reference's own golden files -- inputs
``test/SampleInput/*.txt`` and the expected self-join / AB-join profiles
``test/SampleOutput/mp*_w100.txt`` (Euclidean distance) and
``mpi*_w100.txt`` (index), used by ``test/run_tests.sh:10-211`` with the
tolerances of ``test/difference.py:15-26`` and ``run_tests.sh:34-39`` -- into
one compressed ``.npz`` of plain arrays.  Only data is copied: inputs and
expected outputs.
"""
import os

import numpy as np

REF = "/root/reference/test"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sample_io.npz")

SELF = ["randomwalk8K", "randomwalk16K", "randomwalk32K", "randomwalk64K_nan"]
AB = [("randomwalk16K", "randomwalk32K"), ("randomwalk32K", "randomwalk16K")]
EXTRA_INPUTS = ["randomwalk1K_nan", "poorly_conditioned_test",
                # the input docs/source/cli.rst:127 names as failing in mixed/single
                "earthquake_precision_test100K"]


def _read(path, dtype=np.float64):
    # the reference writes 'nan' / '-nan' / 'inf' spellings; numpy parses them
    with open(path) as f:
        return np.array([float(x) for x in f.read().split()], dtype=dtype)


def main():
    arrays = {}
    names = set(SELF) | {x for p in AB for x in p} | set(EXTRA_INPUTS)
    for name in sorted(names):
        arrays["in_" + name] = _read(os.path.join(REF, "SampleInput", name + ".txt"))
    for name in SELF:
        arrays[f"mp_{name}"] = _read(
            os.path.join(REF, "SampleOutput", f"mp_{name}_w100.txt"), np.float64)
        arrays[f"mpi_{name}"] = _read(
            os.path.join(REF, "SampleOutput", f"mpi_{name}_w100.txt")).astype(np.int64)
    for a, b in AB:
        arrays[f"mp_{a}_{b}"] = _read(
            os.path.join(REF, "SampleOutput", f"mp_{a}_{b}_w100.txt"), np.float64)
        arrays[f"mpi_{a}_{b}"] = _read(
            os.path.join(REF, "SampleOutput", f"mpi_{a}_{b}_w100.txt")).astype(np.int64)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(arrays), "arrays")


if __name__ == "__main__":
    main()
