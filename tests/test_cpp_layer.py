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
