"""CPU check of the FP64 sine/cosine used for the w-phase screens
(curig_amd/csrc/trig.h): the same function, built for the host, against the C
library's sin/cos (numpy) and against exact quadrant values."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "trig_check.cpp")
CSRC = os.path.join(ROOT, "curig_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("trig") / "libtrig.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, "-o", out, SRC])
    L = ctypes.cdll.LoadLibrary(out)
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.trig_check_run.argtypes = [ctypes.c_long, f64, f64, f64]
    L.trig_check_run.restype = None
    return L


def run(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib.trig_check_run(len(x), x, s, c)
    return s, c


@pytest.mark.parametrize("scale", [1.0, 10.0, 1e3, 1e5, 3e7])
def test_against_libm(lib, scale):
    rng = np.random.default_rng(int(scale))
    x = (rng.random(200000) * 2 - 1) * scale
    s, c = run(lib, x)
    # absolute error against the C library, in units of 2^-53 (both functions
    # round the reduced argument, so the bound grows like ulp(x) / 2^53 + 1e-16)
    tol = 2.5e-16 + 2.0 ** -52 * np.abs(x) * 2.0 ** -52
    assert np.all(np.abs(s - np.sin(x)) <= tol)
    assert np.all(np.abs(c - np.cos(x)) <= tol)


def test_exact_points_and_identity(lib):
    k = np.arange(-2000, 2001, dtype=np.float64)
    s, c = run(lib, k * np.pi / 2)
    # sin(k pi/2), cos(k pi/2) in {0, +-1} up to the rounding of k pi / 2 itself
    np.testing.assert_allclose(s, np.sin(k * np.pi / 2), atol=1e-12)
    np.testing.assert_allclose(c, np.cos(k * np.pi / 2), atol=1e-12)
    x = np.linspace(-50, 50, 100001)
    s, c = run(lib, x)
    assert np.max(np.abs(s * s + c * c - 1)) <= 3 * 2.0 ** -52  # rounding of the check itself: 2 ulp
    s2, c2 = run(lib, 2 * x)
    assert np.max(np.abs(s2 - 2 * s * c)) < 1e-15
    assert np.max(np.abs(c2 - (c * c - s * s))) < 1e-15


def test_small_arguments_relative(lib):
    x = np.logspace(-300, -1, 3000)
    s, c = run(lib, np.concatenate([x, -x]))
    xs = np.concatenate([x, -x])
    assert np.all(np.abs(s - np.sin(xs)) <= 2.3e-16 * np.abs(np.sin(xs)))
    assert np.all(np.abs(c - np.cos(xs)) <= 2.3e-16)


@pytest.mark.parametrize("scale", [1.0, 1e3, 1e6])
def test_reduced_fp32_phase(lib, scale):
    """FP64 reduction, FP32 polynomial: absolute error within a few FP32 ulps of 1."""
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C")
    lib.trig_check_rf.argtypes = [ctypes.c_long, f64, f32, f32]
    lib.trig_check_rf.restype = None
    rng = np.random.default_rng(7)
    x = np.ascontiguousarray((rng.random(200000) * 2 - 1) * scale)
    s = np.empty(len(x), np.float32)
    c = np.empty(len(x), np.float32)
    lib.trig_check_rf(len(x), x, s, c)
    assert np.max(np.abs(s - np.sin(x))) < 1.5e-7
    assert np.max(np.abs(c - np.cos(x))) < 1.5e-7
