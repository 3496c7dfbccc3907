"""vsmcx::PfSampler (include/vsmcx.hpp) on the GPU: the Sampler-shaped driver
of the device filter and the table Sampler::print writes
(include/vsmc/core/sampler.hpp:593-721), as the pf example saves it
(example/pf/include/pf_mpi_do.hpp:36-64) and example/pf/pf.R reads it.

tests/cpp/_bin/pf_sampler runs the driver over the pf example's own
observation file (tests/golden/pf.data, a copy of example/pf/pf.data: 100
(x, y) pairs) with N = 1000 (the example's ParticleNum) and threshold 0.5,
and prints the table.  The table is parsed as pf.R parses it
(read.table(header = TRUE): fields split on white space) and checked:
  * layout: the header Size Resampled Accept.0 ESS pos.x pos.y, every field
    followed by a tab, one row per iteration (100);
  * values, FP64 pipeline: Size and Accept.0 = N (the pf init and move return
    1 per particle); Resampled, ESS and pos.x / pos.y equal the CPU oracle's
    run of the same filter (oracle/vsmc_oracle.c) within the 6 significant
    digits the stream prints;
  * the mixed pipeline: the same layout, decisions in {0, 1}, ESS in (0, N].

The print format itself is pinned by reading Sampler::print: the reference's
Sampler does not compile with this toolchain (core/path.hpp:141,144 convert
double pointers to size_t pointers), so no reference-made table exists.
"""
import os
import subprocess

import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "pf_sampler")
DATA = os.path.join(ROOT, "tests", "golden", "pf.data")
HEADER = ["Size", "Resampled", "Accept.0", "ESS", "pos.x", "pos.y"]


def _run(N, scheme, fp64, seed=42, rseed=4242, threshold=0.5):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/pf_sampler not built (run __graft_entry__.build())")
    print("pf_sampler", N, scheme, fp64, flush=True)
    r = subprocess.run([BIN, DATA, str(N), str(scheme), str(int(fp64)), str(seed), str(rseed),
                        str(threshold)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _parse(text):
    lines = text.split("\n")
    # Sampler::print ends each row with '\n'; the example adds std::endl
    assert lines[-1] == "" and lines[-2] == ""
    body = lines[:-2]
    for line in body:
        assert line.endswith("\t"), repr(line[-20:])
    header = body[0].split()
    rows = [line.split() for line in body[1:]]
    return header, rows


@pytest.mark.parametrize("scheme", range(6))
def test_pf_sampler_table_fp64(oracle, scheme):
    N = 1000
    obs = np.loadtxt(DATA)
    T = obs.shape[0]
    header, rows = _parse(_run(N, scheme, True))
    assert header == HEADER
    assert len(rows) == T and all(len(r) == len(HEADER) for r in rows)
    _, _, rec = oracle.cv_run(N, obs, 42, 4242, scheme, 0.5, want_state=False)
    size = np.array([int(r[0]) for r in rows])
    res = np.array([int(r[1]) for r in rows])
    acc = np.array([int(r[2]) for r in rows])
    vals = np.array([[float(v) for v in r[3:]] for r in rows])
    assert (size == N).all() and (acc == N).all()
    assert (res == rec["resampled"]).all()
    # six significant digits, as the stream prints them
    for j, f in enumerate(("ess", "mean_x", "mean_y")):
        np.testing.assert_allclose(vals[:, j], rec[f], rtol=6e-6, atol=1e-12, err_msg=f)
    # each field is the stream's default rendering of a double (%g, 6 digits)
    for r in rows:
        for v in r[3:]:
            assert v == "%g" % float(v), v


def test_pf_sampler_table_mixed():
    N = 1000
    header, rows = _parse(_run(N, V.SYSTEMATIC, False))
    assert header == HEADER and len(rows) == 100
    res = np.array([int(r[1]) for r in rows])
    ess = np.array([float(r[3]) for r in rows])
    assert set(res.tolist()) <= {0, 1}
    assert (ess > 0).all() and (ess <= N * (1 + 1e-6)).all()
