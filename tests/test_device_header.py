"""The public device header include/vsmcx_device.hpp (the RNG user kernels
include), compiled for the host as tests/cpp/_bin/libvsmcx_device_host.so:

  * block functions: the Random123 known-answer vectors (the table in
    test_oracle.py) -- the header is the library's own copy of them;
  * the twelve u01 conversions: bit for bit against the reference's own
    include/vsmc/rng/u01.h compiled here (oracle/_ref/libvsmcref.so), over
    edge integers and random ones, and against their closed forms at 0 and
    at the largest integer;
  * urng streams (urng.h:179-226) and normal01 streams (normal01.h:118-139):
    against the oracle's restatement; beyond the block function that is a
    reading of the source -- Random123 is absent, so urng.h and normal01.h
    cannot be compiled here (parity unpinned for their conventions).  The
    first four draws are checked to be one block's words last to first;
  * vsmcx::Philox4x32State (include/vsmcx.hpp) against vSMC's own Philox4x32
    (oracle/_ref/philox_state, built against the reference headers).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle_ctypes import RefHarness, ref_available
from test_oracle import KAT_PHILOX, KAT_TF32, KAT_TF64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "cpp", "_bin", "libvsmcx_device_host.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


@pytest.fixture(scope="module")
def dh():
    if not os.path.exists(SO):
        pytest.fail("tests/cpp/_bin/libvsmcx_device_host.so not built (run __graft_entry__.build())")
    L = C.CDLL(SO)
    L.dh_philox4x32_block.argtypes = [u32p, u32p, u32p]
    L.dh_threefry4x32_block.argtypes = [u32p, u32p, u32p]
    L.dh_threefry4x64_block.argtypes = [u64p, u64p, u64p]
    L.dh_urng_stream.argtypes = [C.c_int, C.c_uint64, C.c_size_t, u64p]
    L.dh_normal01_stream.argtypes = [C.c_int, C.c_uint64, C.c_size_t, f64p]
    L.dh_u01_32.argtypes = [u32p, C.c_size_t, f64p]
    L.dh_u01_64.argtypes = [u64p, C.c_size_t, f64p]
    return L


@pytest.mark.parametrize("ctr,key,want", KAT_PHILOX)
def test_header_philox_kat(dh, ctr, key, want):
    out = np.zeros(4, np.uint32)
    dh.dh_philox4x32_block(np.array(ctr, np.uint32), np.array(key, np.uint32), out)
    assert list(out) == want


@pytest.mark.parametrize("ctr,key,want", KAT_TF32)
def test_header_threefry4x32_kat(dh, ctr, key, want):
    out = np.zeros(4, np.uint32)
    dh.dh_threefry4x32_block(np.array(ctr, np.uint32), np.array(key, np.uint32), out)
    assert list(out) == want


@pytest.mark.parametrize("ctr,key,want", KAT_TF64)
def test_header_threefry4x64_kat(dh, ctr, key, want):
    out = np.zeros(4, np.uint64)
    dh.dh_threefry4x64_block(np.array(ctr, np.uint64), np.array(key, np.uint64), out)
    assert [int(x) for x in out] == want


def _ints32():
    rng = np.random.default_rng(5)
    edge = [0, 1, 2, 63, 64, 65, 127, 128, 255, 256, 511, 512, (1 << 23) - 1, 1 << 23,
            (1 << 24) - 1, 1 << 24, (1 << 31) - 1, 1 << 31, 0x7fffffc0, 0xffffff00,
            0xffffff80, 0xfffffffe, 0xffffffff]
    return np.concatenate([np.array(edge, np.uint32),
                           rng.integers(0, 1 << 32, 4000, dtype=np.uint64).astype(np.uint32)])


def _ints64():
    rng = np.random.default_rng(6)
    edge = [0, 1, 2, 511, 512, 513, 1023, 1024, 2047, 2048, (1 << 53) - 1, 1 << 53,
            (1 << 63) - 1, 1 << 63, 0x7ffffffffffffe00, 0xfffffffffffff800,
            0xfffffffffffffc00, 0xfffffffffffffffe, 0xffffffffffffffff]
    return np.concatenate([np.array(edge, np.uint64),
                           rng.integers(0, 1 << 63, 4000, dtype=np.uint64) * 2 + 1])


def _dh_u01(dh):
    u32, u64 = _ints32(), _ints64()
    a = np.zeros((u32.size, 8))
    b = np.zeros((u64.size, 4))
    dh.dh_u01_32(u32, u32.size, a)
    dh.dh_u01_64(u64, u64.size, b)
    return u32, u64, a, b


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (needs the vSMC tree)")
def test_header_u01_equals_reference(dh):
    u32, u64, a, b = _dh_u01(dh)
    ref = RefHarness()
    assert (a == ref.u01_32(u32)).all()
    assert (b == ref.u01_64(u64)).all()


def test_header_u01_closed_forms(dh):
    """Ends of each interval: closed ends reach 0 or 1 exactly, open ends stop
    one half or one cell short."""
    u32 = np.array([0, 0xffffffff], np.uint32)
    u64 = np.array([0, 0xffffffffffffffff], np.uint64)
    a = np.zeros((2, 8))
    b = np.zeros((2, 4))
    dh.dh_u01_32(u32, 2, a)
    dh.dh_u01_64(u64, 2, b)
    # 32-bit to float: closed_closed, closed_open, open_closed, open_open
    assert list(a[0, :4]) == [0.0, 0.0, 2.0 ** -24, 2.0 ** -24]
    assert list(a[1, :4]) == [1.0, 1 - 2.0 ** -24, 1.0, 1 - 2.0 ** -24]
    # 32-bit to double
    assert list(a[0, 4:]) == [0.0, 0.0, 2.0 ** -32, 2.0 ** -33]
    assert list(a[1, 4:]) == [1.0, 1 - 2.0 ** -32, 1.0, 1 - 2.0 ** -33]
    # 64-bit to double
    assert list(b[0]) == [0.0, 0.0, 2.0 ** -53, 2.0 ** -53]
    assert list(b[1]) == [1.0, 1 - 2.0 ** -53, 1.0, 1 - 2.0 ** -53]
    # every value in [0, 1], open ends excluded
    u32, u64, a, b = _dh_u01(dh)
    assert (a >= 0).all() and (a <= 1).all() and (b >= 0).all() and (b <= 1).all()
    assert (a[:, [2, 3, 6, 7]] > 0).all() and (a[:, [1, 3, 5, 7]] < 1).all()
    assert (b[:, [2, 3]] > 0).all() and (b[:, [1, 3]] < 1).all()


def test_header_urng_stream(dh, oracle):
    n = 1003
    out = np.zeros(n, np.uint64)
    dh.dh_urng_stream(1, 77, n, out)  # threefry4x32 = the default cburng4x32
    assert (out.astype(np.uint32) == oracle.cburng4x32_stream(77, n)).all()
    # the first block at counter {1, 0, 0, 0}, words last to first
    blk = oracle.threefry4x32_block([1, 0, 0, 0], [77, 0, 0, 0])
    assert list(out[:4]) == [int(x) for x in blk[::-1]]
    # Philox4x32 and Threefry4x64 engines follow the same convention
    for kind, block, key in ((0, oracle.philox4x32_block, [77, 0]),
                             (2, oracle.threefry4x64_block, [77, 0, 0, 0])):
        out = np.zeros(9, np.uint64)
        dh.dh_urng_stream(kind, 77, 9, out)
        b1, b2 = block([1, 0, 0, 0], key), block([2, 0, 0, 0], key)
        want = [int(x) for x in b1[::-1]] + [int(x) for x in b2[::-1]] + [int(block([3, 0, 0, 0], key)[3])]
        assert [int(x) for x in out] == want


def test_header_normal01_stream(dh, oracle):
    n = 501
    out = np.zeros(n)
    dh.dh_normal01_stream(1, 9, n, out)
    np.testing.assert_array_equal(out, oracle.normal01_4x32_53_stream(9, n))
    # the FP32 and the 64-bit versions: N(0, 1) draws, the same order of use
    for kind in (0, 2):
        z = np.zeros(20000)
        dh.dh_normal01_stream(kind, 11, z.size, z)
        assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.03


PHILOX_STATE = os.path.join(ROOT, "oracle", "_ref", "philox_state")


@pytest.mark.skipif(not os.path.exists(PHILOX_STATE),
                    reason="oracle/_ref/philox_state not built (needs the vSMC tree at build time)")
def test_philox_state_text_round_trip():
    """vSMC's PhiloxEngine writes ctr key buffer index but reads buffer ctr key
    index (philox.hpp:455-484): its own pair does not round-trip a mid-block
    state.  vsmcx::Philox4x32State does, and the engine it restores continues
    the original's stream."""
    r = subprocess.run([PHILOX_STATE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    ref_pair, state_pair, same_stream = (int(v) for v in r.stdout.split())
    assert ref_pair == 0
    assert state_pair == 1 and same_stream == 1
