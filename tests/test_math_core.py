"""CPU checks of the kernels' own scalar arithmetic (vsmc_amd/csrc/math_core.hpp),
compiled for the host as tests/cpp/_bin/libvsmcx_math_host.so.  The host build
runs the same IEEE operations as the device build (fma, rint, ldexp and a bit
construction), so these pin the device's values too.

  * exp_lse, the FP64 exp of the normalised weights (W_i = e^(lw_i - m) / sum)
    and of the log-sum-exp partials: within one unit in the last place of the
    correctly rounded e^d over [-746, 0] (random, log-spaced, subnormal range)
    against an 80-bit long-double exp, spot-checked with mpmath at 50 digits,
    plus its special values;
  * u01_oo_24_bits, the FP32 uniform of the Box-Muller normals: equal bit for
    bit to vSMC's u01_open_open_32_24 formula ((u >> 9) + 0.5) 2^-23
    (include/vsmc/rng/u01.h:116-118) for every one of the 2^23 values of u >> 9.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "cpp", "_bin", "libvsmcx_math_host.so")

f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")


@pytest.fixture(scope="module")
def mh():
    if not os.path.exists(SO):
        pytest.fail("tests/cpp/_bin/libvsmcx_math_host.so not built (run __graft_entry__.build())")
    L = C.CDLL(SO)
    L.mh_exp_lse.argtypes = [f64p, C.c_size_t, f64p]
    L.mh_u01_oo_24.argtypes = [u32p, C.c_size_t, f32p]
    return L


def exp_lse(mh, d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.empty_like(d)
    mh.mh_exp_lse(d, d.size, out)
    return out


def ulp_error(got, d):
    """|got - e^d| in units of the last place of the correctly rounded e^d,
    with e^d from the 64-bit-significand long double exp."""
    ref = np.exp(d.astype(np.longdouble))
    near = ref.astype(np.float64)
    ulp = np.spacing(near)  # subnormal spacing below 2^-1022
    return np.abs((got.astype(np.longdouble) - ref) / ulp.astype(np.longdouble)).astype(np.float64)


def test_exp_lse_within_one_ulp(mh):
    rng = np.random.default_rng(5)
    parts = [
        -746.0 * rng.random(2_000_000),                      # the whole domain
        -np.exp(rng.uniform(np.log(1e-300), np.log(50.0), 500_000)),  # small |d|
        rng.uniform(-745.2, -708.0, 200_000),                # subnormal results
        -np.arange(0, 746, 0.5),                             # half-integers, ties of k
        -np.log(2.0) * np.arange(0, 1076) + 0.5 * np.log(2.0),  # |r| = ln2/2 edges
    ]
    d = np.concatenate(parts)
    err = ulp_error(exp_lse(mh, d), d)
    worst = float(err.max())
    assert worst <= 1.0, (worst, float(d[err.argmax()]))
    # and most values are the correctly rounded ones
    assert float((err < 0.5).mean()) > 0.9


def test_exp_lse_against_mpmath(mh):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    rng = np.random.default_rng(11)
    d = np.concatenate([-746.0 * rng.random(300), [-1e-12, -0.34657359027997264, -0.5, -1.0,
                                                    -708.3964185322641, -745.1332191019411]])
    got = exp_lse(mh, d)
    for x, g in zip(d, got):
        e = mpmath.exp(mpmath.mpf(float(x)))
        near = float(e)
        ulp = np.spacing(near) if near != 0 else 2.0 ** -1074
        assert abs(mpmath.mpf(float(g)) - e) <= ulp, (x, g, near)


def test_exp_lse_special_values(mh):
    d = np.array([0.0, -0.0, -np.inf, np.nan, -746.5, -1000.0, -1e308, -5e-324, -1e-300])
    e = exp_lse(mh, d)
    assert e[0] == 1.0 and e[1] == 1.0
    assert e[2] == 0.0 and np.isnan(e[3])
    assert e[4] == 0.0 and e[5] == 0.0 and e[6] == 0.0
    assert e[7] == 1.0 and e[8] == 1.0
    # the smallest subnormal near the end of the range, as e^d rounds there
    assert exp_lse(mh, np.array([-745.1332191019411]))[0] == 5e-324
    # monotone over a fine grid of the reduction's boundaries
    g = np.linspace(-40.0, 0.0, 400_001)
    assert np.all(np.diff(exp_lse(mh, g)) >= 0)


def test_u01_oo_24_bits_equals_reference_formula(mh):
    m = np.arange(1 << 23, dtype=np.uint32)
    rng = np.random.default_rng(3)
    u = (m << np.uint32(9)) | rng.integers(0, 1 << 9, m.size, dtype=np.uint32)
    got = np.empty(u.size, dtype=np.float32)
    mh.mh_u01_oo_24(u, u.size, got)
    # vSMC's formula in FP32 arithmetic; every step is exact
    want = (np.float32(0.5) + m.astype(np.float32)) * np.float32(1.0 / 8388608.0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got.min() > 0 and got.max() < 1
