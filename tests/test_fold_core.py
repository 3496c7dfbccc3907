"""CPU checks of the exact-fold arithmetic the resampling kernels run.

tests/cpp/test_fold_core.cpp drives vsmc_amd/csrc/fold_core.hpp (the same
header the HIP kernels include) through the three-pass segmented evaluation
on random, heavy-tailed, tied, degenerate and decaying weight vectors and
compares every prefix value with the sequential left fold of
include/vsmc/resample/common.hpp:76-83, bit for bit.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def fold_core_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fold") / "test_fold_core")
    src = os.path.join(ROOT, "tests", "cpp", "test_fold_core.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", out, src], check=True)
    return out


def test_segmented_fold_is_exact(fold_core_bin):
    r = subprocess.run([fold_core_bin], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout


def test_fold_core_detects_a_broken_bound(fold_core_bin, tmp_path):
    """With the classification bound forced to zero the segmented evaluation
    must either still be exact or fail loudly (a transducer that does not
    apply); it must never produce a wrong value silently.  Planted in a CPU
    build of the same source."""
    src = os.path.join(ROOT, "tests", "cpp", "test_fold_core.cpp")
    out = str(tmp_path / "broken")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-DVSMCX_TEST_ZERO_EPS",
                    "-o", out, src], check=True)
    r = subprocess.run([out], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0  # the planted mistake does break cases ...
    # ... and every one is caught: rc 3 would be a silent mismatch; rc 1, 2
    # and 6 are transducers that refuse to apply
    assert "failed rc=" in r.stderr
    assert "rc=3" not in r.stderr, r.stderr[-2000:]
