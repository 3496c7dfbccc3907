"""vsmc::Particle resampled by vSMC's own operator and by vsmcx::HipResample.

oracle/_ref/dropin_resample is built here from tests/cpp/dropin_resample.cpp
against the unmodified reference headers (oracle/Makefile, target dropin)
and links vsmc_amd/libvsmcx.so.  It runs Particle::resample
(include/vsmc/core/particle.hpp:230-250) twice on identical log-weights and
engines, once per operator, for all six schemes and N up to 2^20, and
requires identical ancestors and identical resample engines afterwards.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_resample")


@pytest.mark.gpu
def test_vsmc_particle_with_hip_resample():
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/dropin_resample not built (needs the vSMC tree at build time)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "failures 0" in r.stdout
