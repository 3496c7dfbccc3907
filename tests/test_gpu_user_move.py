"""Moves written as user kernels, on the GPU: tests/cpp/_bin/user_move runs
the OpenCL pf example's cv_init / cv_move (example/pf/pf_cl.cl:46-122) as
HIP kernels drawing from include/vsmcx_device.hpp (cburng4x32 = Threefry4x32,
normal01_4x32_53, key SEED + i, the counter of each slot kept across
iterations), over a vsmcx_state and the device weight set (vsmcx_weights_*:
set / add_log_weight from the kernel's device array, ESS test, exact-fold
resampling, in-place gather).  Nothing crosses PCIe inside its loop.

Against the oracle's restatement of the same program (oracle/vsmc_oracle.c,
orc_clpf_run: the same kernels on vSMC's WeightSet, resamplers and
StateMatrix::copy) over the pf example's data (tests/golden/pf.data, 100
steps), every scheme: resampling decisions and resample-engine draws equal,
ESS and log-evidence within 1e-9 relative, final particles within 1e-9.
The device's sqrt / log / log1p / sin / cos are not glibc's, so the values
agree to rounding rather than bit for bit.

Also the weight set's argument and sequence checks.
"""
import os
import subprocess

import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "user_move")
DATA = os.path.join(ROOT, "tests", "golden", "pf.data")


def _run(N, scheme, seed, rseed, threshold, obs):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/user_move not built (run __graft_entry__.build())")
    print("user_move", N, scheme, seed, rseed, threshold, flush=True)
    text = "\n".join(f"{x:.17g} {y:.17g}" for x, y in obs) + "\n"
    r = subprocess.run([BIN, str(N), str(scheme), str(seed), str(rseed), str(threshold)],
                       input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    T = len(obs)
    hist = np.array([[float(v) for v in line.split()] for line in lines[:T]])
    st = np.array([[float(v) for v in line.split()] for line in lines[T:T + N]])
    assert hist.shape == (T, 5) and st.shape == (N, 4)
    return hist, st


@pytest.mark.parametrize("scheme", range(6))
def test_user_move_equals_oracle(oracle, scheme):
    obs = np.loadtxt(DATA)
    N = 1000
    hist, st = _run(N, scheme, 101, 4242, 0.5, obs)
    ost, rec = oracle.clpf_run(N, obs, 101, 4242, scheme, 0.5)
    assert (hist[:, 0] == np.arange(len(obs))).all()
    assert (hist[:, 3] == rec["resampled"]).all()
    assert (hist[:, 4] == rec["draws"]).all()
    assert rec["resampled"].sum() > 5
    np.testing.assert_allclose(hist[:, 1], rec["ess"], rtol=1e-9)
    np.testing.assert_allclose(hist[:, 2], rec["log_z"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-12)


def test_user_move_larger_and_never_resampling(oracle):
    obs = np.loadtxt(DATA)[:40]
    for N, threshold in ((20000, 0.5), (3000, 0.0)):
        hist, st = _run(N, V.STRATIFIED, 7, 99, threshold, obs)
        ost, rec = oracle.clpf_run(N, obs, 7, 99, V.STRATIFIED, threshold)
        assert (hist[:, 3] == rec["resampled"]).all()
        if threshold == 0.0:
            assert rec["resampled"].sum() == 0
        np.testing.assert_allclose(hist[:, 2], rec["log_z"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-12)


def test_weights_argument_checks():
    ctx = V.Context(0)
    with pytest.raises(V.VsmcxError):
        V.Weights(ctx, 100, scheme=9)
    with pytest.raises(V.VsmcxError):
        V.Weights(ctx, 0)
    w = V.Weights(ctx, 100, max_steps=2)
    buf = w.buffer()
    assert buf
    with pytest.raises(V.VsmcxError):
        w.add_log(buf)  # before set_log
    with pytest.raises(V.VsmcxError):
        w.resample()  # before set_log
    with pytest.raises(V.VsmcxError):
        w.set_log(None)
    s = V.State(ctx, 50, 4, V.FP64_FUSED)
    assert w.steps() == 0
    w.close()
    s.close()
    ctx.close()
