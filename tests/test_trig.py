"""The Newton-step kernel's own sin/cos (Cody-Waite reduction + Taylor series, olf-gpu_amd/csrc/trig.hpp), built as
host code with g++ and checked against the C library's correctly reduced sin and cos (numpy): absolute error within
4e-16 over the arguments power-flow angles take and far beyond, exact quadrant values, odd/even symmetry, NaN for
non-finite input, and bounded finite output for absurd arguments."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def trig(tmp_path_factory):
    out = tmp_path_factory.mktemp("trig") / "libtrig.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    os.path.join(HERE, "native", "trig_host.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    p = ctypes.POINTER(ctypes.c_double)
    lib.trig_sincos.argtypes = [ctypes.c_int, p, p, p]

    def f(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s = np.empty_like(x)
        c = np.empty_like(x)
        lib.trig_sincos(x.size, x.ctypes.data_as(p), s.ctypes.data_as(p), c.ctypes.data_as(p))
        return s, c
    return f


def test_accuracy_against_the_c_library(trig):
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-4, 4, 200000), rng.uniform(-100, 100, 100000),
                        rng.uniform(-1e6, 1e6, 50000), rng.uniform(-1e12, 1e12, 20000),
                        np.arange(-64, 65) * (np.pi / 4), np.arange(-64, 65) * (np.pi / 2),
                        np.nextafter(np.pi / 4, 0) * np.array([1, -1, 3, -3, 5])])
    s, c = trig(x)
    assert np.max(np.abs(s - np.sin(x))) <= 4e-16
    assert np.max(np.abs(c - np.cos(x))) <= 4e-16


def test_exact_points_and_symmetry(trig):
    s, c = trig(np.array([0.0, -0.0, 1e-300, -1e-300]))
    assert s[0] == 0.0 and c[0] == 1.0 and s[1] == 0.0 and c[1] == 1.0
    assert s[2] == 1e-300 and s[3] == -1e-300
    rng = np.random.default_rng(12)
    x = rng.uniform(0, 50, 10000)
    sp, cp = trig(x)
    sm, cm = trig(-x)
    np.testing.assert_array_equal(sp, -sm)
    np.testing.assert_array_equal(cp, cm)
    # sin^2 + cos^2 = 1 to rounding
    assert np.max(np.abs(sp * sp + cp * cp - 1.0)) <= 6e-16


def test_non_finite_and_huge_arguments(trig):
    s, c = trig(np.array([np.nan, np.inf, -np.inf]))
    assert np.all(np.isnan(s)) and np.all(np.isnan(c))
    x = np.array([1e16, -3e20, 1e100, -1e300, np.finfo(float).max])
    s, c = trig(x)
    assert np.all(np.isfinite(s)) and np.all(np.isfinite(c))
    assert np.all(np.abs(s) <= 1.0 + 1e-15) and np.all(np.abs(c) <= 1.0 + 1e-15)
