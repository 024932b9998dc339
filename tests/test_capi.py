"""The C-ABI library without a GPU: it loads, exports every symbol the headers
declare, its host-only plan math equals the oracle's restatement, and argument
errors are reported before any device work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from curig_amd import _lib
from oracle import nufft_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    names = set()
    for h in ("cugridder.h", "curig_ext.h", "cufft_nufft.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|const char \*)\s*\*?\s*([A-Za-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_c_functions()
    cxx = set(_lib.CXX_EXPORTS)
    assert declared - cxx == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    for name in _lib.EXPORTS:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in cxx:
        assert re.search(r"\b_Z\d+%s" % name, out), name


@pytest.mark.parametrize("nx,ny,fov,eps,sign,w0,w1", [
    (256, 256, 2.0, 1e-6, -1, -900.0, 1100.0),
    (2048, 2048, 4.0, 1e-6, -1, -4000.0, 3500.0),
    (8192, 8192, 6.0, 1e-5, -1, -34000.0, 34600.0),
    (30, 250, 10.0, 5e-7, 1, -0.1, 0.15),
    (17, 33, 30.0, 1e-12, -1, 0.0, 0.0),
])
def test_plan_info_matches_oracle(nx, ny, fov, eps, sign, w0, w1):
    freq = 1.4e9
    uvw = np.array([[0.0, 0.0, w0], [0.0, 0.0, w1]])
    want = R.plan_params(2, nx, ny, fov, freq, uvw, eps, 2.0, sign)
    got = _lib.plan_info(2, nx, ny, fov, freq, eps, 2.0, sign, min(w0, w1), max(w0, w1))
    for k in ("kw", "nf1", "nf2", "nf3"):
        assert got[k] == want[k], k
    for k in ("beta", "c", "gamma", "h", "i_center", "o_center", "pixelsize_x", "pixelsize_y"):
        assert abs(got[k] - want[k]) <= 1e-12 * max(1.0, abs(want[k])), k


@pytest.mark.parametrize("eps", [2e-1, 1e-4, 1e-6, 1e-12])
@pytest.mark.parametrize("nf", [14, 60, 512])
def test_kernel_fseries_matches_oracle(eps, nf):
    out = np.zeros(nf // 2 + 1)
    assert _lib.load().curig_kernel_fseries(eps, 2.0, nf, out) == 0
    _, kw, beta, c = R.setup_conv_opts(eps, 2.0)
    want = R.fseries_grid(nf, kw, beta, c) * np.exp(-beta)   # normalised kernel (peak 1)
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-13 * abs(want[0]))


@pytest.mark.parametrize("n", [1, 4, 22, 56])
def test_gauss_legendre_export(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert _lib.load().curig_gauss_legendre(n, x, w) == 0
    X, W = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, X, atol=1e-14)
    np.testing.assert_allclose(w, W, atol=1e-14)


def _call(fn, nrow=4, nx=8, ny=8, eps=1e-6, sigma=2.0, sign=-1):
    uvw = np.zeros((max(nrow, 1), 3))
    vis = np.zeros(max(nrow, 1), dtype=np.complex128)
    dirty = np.zeros((max(nx, 1), max(ny, 1)), dtype=np.complex128)
    return fn(nrow, nx, ny, 2.0, 1e9, uvw, vis, dirty, eps, sigma, sign)


def test_error_codes_before_device_work():
    lib = _lib.load()
    for fn in (lib.ms2dirty_1, lib.dirty2ms_1):
        assert _call(fn, eps=1e-17) == 1       # reference: setup_conv_opts -> 1, wrapper returns
        assert _call(fn, sigma=1.0) == 2       # invalid upsampling factor
        assert _call(fn, nx=0) == 3
        assert _call(fn, nrow=-1) == 4
    assert "epsilon" in lib.curig_last_error().decode() or lib.curig_last_error()
    assert lib.curig_set_tuning(-1, 0, 0, 0) == 5
    assert lib.curig_set_tuning(0, 0, 0, 0) == 0


def test_python_api_mirrors_reference_names():
    import curig_amd

    for name in ("ms2dirty", "dirty2ms", "vis2dirty", "dirty2vis"):
        assert callable(getattr(curig_amd, name))
