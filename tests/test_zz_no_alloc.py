"""vsmcx_resample_noalloc against vsmcx_resample, through the C ABI: the same
replication counts, copy_from layout and engine state bit for bit, on two
calls in a row with the second smaller (the fixed scratch array keeps the
first call's contents), for every scheme, and the size limit enforced."""
import sys

import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 18


def say(*a):
    print("[no_alloc]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def _weights(kind, n, rng):
    if kind == "heavy":
        w = rng.random(n) ** 8
    elif kind == "zeros":
        w = rng.random(n)
        w[rng.random(n) < 0.7] = 0
        if not w.any():
            w[0] = 1
    else:
        w = rng.random(n)
    return w / w.sum()


def _pair(ctx, scheme, w, seed, N=None):
    e1 = V.ThreefryState.seeded(seed)
    e2 = V.ThreefryState.seeded(seed)
    rep, cf = ctx.resample(scheme, w, e1, N=N)
    rep2, cf2 = ctx.resample(scheme, w, e2, N=N, noalloc=True)
    return (rep, cf, e1), (rep2, cf2, e2)


def _same(a, b):
    rep, cf, e1 = a
    rep2, cf2, e2 = b
    assert (rep == rep2).all()
    assert (cf == cf2).all()
    assert list(e1.ctr) == list(e2.ctr) and list(e1.key) == list(e2.key)
    assert int(e1.index) == int(e2.index) and list(e1.buffer) == list(e2.buffer)


@pytest.mark.parametrize("scheme", range(6))
def test_noalloc_equals_resample_two_calls(ctx, scheme):
    rng = np.random.default_rng(500 + scheme)
    # the largest size first, then a smaller one over what the first call left
    for n, kind in ((NOALLOC_MAX, "heavy"), (50001, "zeros"), (4097, "flat"), (1, "flat")):
        say("scheme", scheme, "n", n, kind)
        w = _weights(kind, n, rng)
        seed = int(rng.integers(1, 2**31))
        a, b = _pair(ctx, scheme, w, seed)
        _same(a, b)


def test_noalloc_m_ne_n_and_engine_carry(ctx):
    rng = np.random.default_rng(9)
    e1 = V.ThreefryState.seeded(77)
    e2 = V.ThreefryState.from_buffer_copy(e1)
    for r in range(5):
        m = int(rng.integers(1, 20000))
        n = int(rng.integers(1, 20000))
        w = _weights("heavy", m, rng)
        say("carry", r, m, n)
        rep, cf = ctx.resample(V.STRATIFIED, w, e1, N=n)
        rep2, cf2 = ctx.resample(V.STRATIFIED, w, e2, N=n, noalloc=True)
        assert (rep == rep2).all() and (cf == cf2).all()
        assert list(e1.ctr) == list(e2.ctr) and int(e1.index) == int(e2.index)


def test_noalloc_allocates_nothing_and_rejects_oversize(ctx):
    before = ctx.allocated()
    w = np.full(1000, 1e-3)
    e = V.ThreefryState.seeded(3)
    ctx.resample(V.SYSTEMATIC, w, e, noalloc=True)
    assert ctx.allocated() == before
    big = np.full(NOALLOC_MAX + 1, 1.0 / (NOALLOC_MAX + 1))
    with pytest.raises(V.VsmcxError) as ei:
        ctx.resample(V.SYSTEMATIC, big, V.ThreefryState.seeded(3), noalloc=True)
    assert ei.value.code == -1 and "VSMCX_NOALLOC_MAX" in str(ei.value)
    with pytest.raises(V.VsmcxError):
        ctx.resample(V.SYSTEMATIC, w, V.ThreefryState.seeded(3), N=NOALLOC_MAX + 1,
                     noalloc=True)
