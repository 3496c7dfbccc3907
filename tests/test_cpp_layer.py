"""CPU checks of the C++ host layer (include/vsmcx.hpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VSMC_INC = "/root/reference/include"


def test_header_compiles_standalone(tmp_path):
    """vsmcx.hpp compiles on its own (C++11) against the C ABI header."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "vsmcx.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only",
                    "-I" + os.path.join(ROOT, "include"), str(src)], check=True)


def test_engine_state_round_trip(tmp_path):
    """State read from vsmc::Threefry4x64 and written back continues its stream."""
    if not os.path.isdir(os.path.join(VSMC_INC, "vsmc")):
        pytest.skip("vSMC headers not present")
    out = str(tmp_path / "rt")
    subprocess.run(["g++", "-std=c++11", "-O1", "-I" + VSMC_INC, "-I" + os.path.join(ROOT, "include"),
                    "-o", out, os.path.join(ROOT, "tests", "cpp", "engine_roundtrip.cpp")],
                   check=True)
    r = subprocess.run([out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "failures 0" in r.stdout, r.stdout + r.stderr


def test_cpp_llh_const_is_the_student_t_normaliser(tmp_path):
    """vsmcx::Filter::llh_const (the constant the C++ layer puts in every
    filter descriptor) equals scipy's Student-t log density at 0."""
    from scipy import stats
    src = tmp_path / "c.cpp"
    src.write_text('#include <cstdio>\n#include "vsmcx.hpp"\n'
                   'int main() { std::printf("%.17g %.17g\\n", vsmcx::Filter::llh_const(10.0, 0.1),'
                   ' vsmcx::Filter::llh_const(3.0, 2.5)); return 0; }\n')
    out = str(tmp_path / "c")
    subprocess.run(["g++", "-std=c++11", "-O1", "-I" + os.path.join(ROOT, "include"), "-o", out,
                    str(src)], check=True)
    a, b = map(float, subprocess.run([out], capture_output=True, text=True, check=True).stdout.split())
    assert abs(a - stats.t.logpdf(0.0, df=10, scale=0.1)) < 1e-13
    assert abs(b - stats.t.logpdf(0.0, df=3, scale=2.5)) < 1e-13
