"""Device arithmetic the FP64 parity depends on.

sqrt and division must be correctly rounded, as glibc's are.  The contact
force's integer power (repulsion.cpp:15, std::pow with an int exponent) is
evaluated on the device as the correctly rounded power; glibc's pow is within
0.52 ulp and differs from the correctly rounded value in well under 1% of
arguments, which the second test bounds.
"""
import ctypes
import ctypes.util
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sqrt_div_correctly_rounded():
    from softflow_amd import capi
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.uniform(0, 10, 200000), 10.0 ** rng.uniform(-30, 30, 100000)])
    b = np.concatenate([rng.uniform(0.1, 10, 200000), 10.0 ** rng.uniform(-30, 30, 100000)])
    assert np.array_equal(capi.probe_math(0, a), np.sqrt(a))
    assert np.array_equal(capi.probe_math(1, a, b), a / b)


def test_contact_power_correctly_rounded():
    from softflow_amd import capi
    rng = np.random.default_rng(1)
    x = rng.uniform(0.28, 1e3, 4000)        # sigma / r over the contact range
    for p in (2, 3, 4, 6, 12):
        got = capi.probe_math(3, x, np.full_like(x, p))
        exact = np.array([float(Fraction(v) ** p) for v in x])   # one rounding of the exact power
        assert np.array_equal(got, exact), p


def test_contact_power_vs_glibc_pow():
    from softflow_amd import capi
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.pow.restype = ctypes.c_double
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(2)
    x = rng.uniform(0.28, 1e3, 20000)
    got = capi.probe_math(3, x, np.full_like(x, 4.0))
    ref = np.array([libm.pow(v, 4.0) for v in x])
    ulps = np.abs(got.view(np.int64) - ref.view(np.int64))
    assert ulps.max() <= 1
    assert np.count_nonzero(ulps) < 0.01 * x.size


def test_copy_ceiling_is_a_plausible_hbm_rate():
    # the streaming-copy reference beside the 8 TB/s peak: a real HBM rate,
    # below the peak, and above what any cache-resident copy would hide
    from softflow_amd import capi
    g = capi.copy_ceiling(1 << 28, 5)
    assert 1000.0 < g < 8000.0, g
