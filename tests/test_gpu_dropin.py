"""vSMC's own Particle<T> driven by the C++ adaptors of include/vsmcx.hpp.

1. vsmc::Particle resampled by vSMC's own operator and by vsmcx::HipResample.

oracle/_ref/dropin_resample is built here from tests/cpp/dropin_resample.cpp
against the unmodified reference headers (oracle/Makefile, target dropin)
and links vsmc_amd/libvsmcx.so.  It runs Particle::resample
(include/vsmc/core/particle.hpp:230-250) twice on identical log-weights and
engines, once per operator, for all six schemes and N up to 2^20, and
requires identical ancestors and identical resample engines afterwards.

2. vsmc::Particle<vsmcx::StateHIP<4, FP>> stepped by vsmcx::CvMoveHIP and
resampled by vsmcx::HipResample through the unmodified Particle::resample
(oracle/_ref/dropin_state, from tests/cpp/dropin_state.cpp): the particles in
HBM, the weights in vSMC's host WeightSet.  In FP64 every iteration's ESS,
log-evidence and MonitorMCMC mean match the CPU oracle's run of the same
filter within 1e-9, with the same resampling decisions; in FP32 the first
iteration (before any resample) is within 1e-5.
"""
import os
import subprocess

import numpy as np

import pytest

import vsmc_amd as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_resample")
BIN_STATE = os.path.join(ROOT, "oracle", "_ref", "dropin_state")


@pytest.mark.gpu
def test_vsmc_particle_with_hip_resample():
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/dropin_resample not built (needs the vSMC tree at build time)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "failures 0" in r.stdout


def _run_state(N, T, scheme, fp64, seed, rseed, threshold, obs):
    inp = "\n".join(f"{x:.17g} {y:.17g}" for x, y in obs) + "\n"
    r = subprocess.run([BIN_STATE, str(N), str(T), str(scheme), str(fp64), str(seed), str(rseed),
                        str(threshold)], input=inp, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [l.split() for l in r.stdout.strip().splitlines()]
    return {"ess": np.array([float(x[1]) for x in rows]), "log_z": np.array([float(x[2]) for x in rows]),
            "mean_x": np.array([float(x[3]) for x in rows]), "mean_y": np.array([float(x[4]) for x in rows]),
            "resampled": np.array([int(x[5]) for x in rows])}


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [V.SYSTEMATIC, V.STRATIFIED, V.RESIDUAL])
def test_vsmc_particle_with_state_hip(oracle, scheme):
    if not os.path.exists(BIN_STATE):
        pytest.skip("oracle/_ref/dropin_state not built (needs the vSMC tree at build time)")
    N, T, seed, rseed = 20000, 30, 42, 4242
    obs = V.cv_simulate(T, 8)
    got = _run_state(N, T, scheme, 1, seed, rseed, 0.5, obs)
    _, _, rec = oracle.cv_run(N, obs, seed, rseed, scheme, 0.5)
    assert (got["resampled"] == rec["resampled"]).all()
    for f in ("ess", "log_z", "mean_x", "mean_y"):
        np.testing.assert_allclose(got[f], rec[f], rtol=1e-9, atol=1e-9, err_msg=f)
    m = _run_state(N, 1, scheme, 0, seed, rseed, 0.5, obs[:1])
    assert abs(m["ess"][0] - rec["ess"][0]) <= 1e-5 * rec["ess"][0]
    assert abs(m["log_z"][0] - rec["log_z"][0]) <= 1e-5
