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


@pytest.mark.parametrize("scheme,k", [(V.SYSTEMATIC, 17), (V.MULTINOMIAL, 8),
                                      (V.RESIDUAL_STRATIFIED, 25)])
def test_resumed_fp64_filter_matches_oracle(ctx, oracle, tmp_path, scheme, k):
    """A FP64 filter saved after k steps, loaded into a new filter and run to
    the end equals the FP64 oracle's uninterrupted run of vSMC's algorithm:
    identical resampling decisions and engine draws, ESS, log-evidence,
    means, final particles and log-weights within 1e-9 relative."""
    N, T = 6000, 40
    obs = V.cv_simulate(T, 9)
    desc = V.cv_desc(N, T, precision=V.FP64_FUSED, scheme=scheme, seed=42, resample_seed=4242)
    print("checkpoint vs oracle", scheme, k, flush=True)
    b = _run(ctx, desc, obs, k)
    path = os.path.join(tmp_path, "o.ck")
    b.save(path)
    b.close()
    c = V.Filter.load(ctx, path)
    c.run(T - k)
    c.finish()
    h = c.history()
    st, lw = c.state()
    c.close()
    ost, olw, rec = oracle.cv_run(N, obs, 42, 4242, scheme, 0.5)
    assert (h["resampled"] == rec["resampled"]).all()
    assert (h["draws"] == rec["draws"]).all()
    for f in ("ess", "log_z", "mean_x", "mean_y"):
        np.testing.assert_allclose(h[f], rec[f], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)


def test_checkpoint_header_checked_before_allocation(ctx, tmp_path):
    """A header whose particle count or horizon is corrupt, or whose sizes do
    not match the file, is refused before any allocation from those sizes;
    a save replaces an earlier checkpoint only once the new one is whole."""
    obs = V.cv_simulate(5, 1)
    p = os.path.join(tmp_path, "h.ck")
    f = _run(ctx, V.cv_desc(500, 5), obs, 3)
    f.save(p)
    good = open(p, "rb").read()
    # vsmcx_filter_desc starts after magic (8 B) and four uint32 (16 B): n, then horizon
    for off, val in ((24, 1 << 40), (24, 0), (32, 1 << 50)):
        bad = bytearray(good)
        bad[off:off + 8] = int(val).to_bytes(8, "little")
        with open(p, "wb") as fh:
            fh.write(bytes(bad))
        with pytest.raises(V.VsmcxError):
            V.Filter.load(ctx, p)
    with open(p, "wb") as fh:
        fh.write(good + b"trailing")
    with pytest.raises(V.VsmcxError):
        V.Filter.load(ctx, p)
    # a save into a directory that cannot hold the temporary file fails and
    # leaves the earlier checkpoint as it was
    with open(p, "wb") as fh:
        fh.write(good)
    with pytest.raises(V.VsmcxError):
        f.save(os.path.join(tmp_path, "no_such_dir", "x.ck"))
    assert open(p, "rb").read() == good
    f.save(p)
    assert not os.path.exists(p + ".tmp")
    g = V.Filter.load(ctx, p)
    assert g.steps() == 3
    g.close()
    f.close()
