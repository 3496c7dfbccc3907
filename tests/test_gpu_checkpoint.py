"""Checkpoint / resume of the device filter (vsmcx_filter_save / _load): a
filter saved after k steps and loaded into a new filter continues exactly as
an uninterrupted run -- history (ESS, log-evidence, means, decisions, engine
draws) and final particles and log-weights bit for bit -- for both fused
pipelines, across a resample and a non-resample step boundary.  Also the
load-time checks: a file that is not a checkpoint, a truncated one, and a
save from the plain pipeline are refused.
"""
import os

import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def _run(ctx, desc, obs, steps):
    f = V.Filter(ctx, desc)
    f.set_obs(obs)
    f.run(steps)
    return f


@pytest.mark.parametrize("prec", [V.MIXED, V.FP64_FUSED])
@pytest.mark.parametrize("scheme,k", [(V.SYSTEMATIC, 13), (V.RESIDUAL, 1), (V.STRATIFIED, 29)])
def test_checkpoint_resume_is_exact(ctx, tmp_path, prec, scheme, k):
    N, T = 6000, 30
    obs = V.cv_simulate(T, 4)
    desc = V.cv_desc(N, T, precision=prec, scheme=scheme, seed=11, resample_seed=12)
    print("checkpoint", prec, scheme, k, flush=True)
    a = _run(ctx, desc, obs, T)
    a.finish()
    ha, (sa, lwa) = a.history(), a.state()
    a.close()
    b = _run(ctx, desc, obs, k)
    path = os.path.join(tmp_path, "f.ck")
    b.save(path)
    b.close()
    c = V.Filter.load(ctx, path)
    assert c.steps() == k
    c.run(T - k)
    c.finish()
    hc, (sc, lwc) = c.history(), c.state()
    c.close()
    assert ha.tobytes() == hc.tobytes()
    assert (sa == sc).all() and (lwa == lwc).all()


def test_checkpoint_rejects_bad_files(ctx, tmp_path):
    obs = V.cv_simulate(5, 1)
    p = os.path.join(tmp_path, "x.ck")
    with open(p, "wb") as fh:
        fh.write(b"not a checkpoint at all, just some bytes" * 4)
    with pytest.raises(V.VsmcxError):
        V.Filter.load(ctx, p)
    f = _run(ctx, V.cv_desc(500, 5), obs, 3)
    f.save(p)
    f.close()
    data = open(p, "rb").read()
    with open(p, "wb") as fh:
        fh.write(data[: len(data) // 2])
    with pytest.raises(V.VsmcxError):
        V.Filter.load(ctx, p)
    g = _run(ctx, V.cv_desc(500, 5, precision=V.FP64_PLAIN), obs, 2)
    with pytest.raises(V.VsmcxError):
        g.save(p)
    g.close()
    with pytest.raises(V.VsmcxError):
        V.Filter.load(ctx, os.path.join(tmp_path, "missing.ck"))
