"""GPU: the C++ plan APIs (include/cugridder.h, include/cufft_nufft.h) run end
to end by a hipcc-built driver (tests/plan_driver.cpp) that uses them as the
reference's drivers do (test/w_s_gridder_test.cu:166-271,
test/w_s_degridder_test.cu:170-276, test/nufft_2d_1_test.cu:180-327), checked
against the oracle's direct sums.

Bars as in the other parity tests: 3 eps + 1e-11 (FP64), + 3e-6 (FP32-mixed)."""
import ctypes
import os
import sys
import zlib

import numpy as np
import pytest

from oracle import nufft_ref as R
from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_bin", "libcurig_plan_driver.so")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_drv = None


def driver():
    global _drv
    if _drv is None:
        from curig_amd import _lib

        _lib.load()  # the library (and torch's HIP runtime) first, as the driver links to it
        if not os.path.exists(DRIVER):
            raise RuntimeError(f"{DRIVER} not built (python -m curig_amd.build builds it)")
        d = ctypes.cdll.LoadLibrary(DRIVER)
        f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
        c128 = np.ctypeslib.ndpointer(np.complex128, flags="C")
        ci, cd, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        d.drv_gridder.argtypes = [ci, ci, ci, ci, cd, cd, f64, f64, f64, ci, ci, vp, c128, c128, cd, cd, ci, ci]
        d.drv_gridder.restype = ci
        d.drv_nufft.argtypes = [ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, ci, cd, cd, ci, ci, c128, c128]
        d.drv_nufft.restype = ci
        _drv = d
    return _drv


def _err(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(a))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _last_error():
    from curig_amd import _lib

    return _lib.load().curig_last_error()


# name, nrow, nx, ny, fov, eps, pirange, sign, weights, kerevalmeth, precision
GRID_CASES = [
    ("fp64_w_pirange0_sign-1", 600, 64, 64, 4.0, 1e-7, 0, -1, True, 0, 0),
    ("fp64_now_pirange1_sign+1", 500, 48, 80, 6.0, 1e-9, 1, 1, False, 0, 0),
    ("fp64_horner_lds", 800, 128, 128, 3.0, 1e-6, 0, -1, True, 1, 0),
    ("fp32_lds_sign+1", 900, 128, 128, 3.0, 1e-5, 0, 1, True, 0, 1),
    ("fp32_horner_pirange1", 700, 64, 48, 5.0, 1e-5, 1, -1, True, 1, 1),
]


@pytest.mark.parametrize("case", GRID_CASES, ids=[c[0] for c in GRID_CASES])
def test_gridder_plan_api_both_directions(case):
    name, nrow, nx, ny, fov, eps, pirange, sign, use_w, kerev, prec = case
    say(f"[plan api gridder] {name}")
    d = driver()
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    f0 = 1.2e9
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * nx / 2 * 0.9
    uvw[:, 2] *= 4
    wgt = rng.uniform(0.5, 1.0, nrow) if use_w else None
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    s = 2 * np.pi if pirange else 1.0
    u, v, w = (np.ascontiguousarray(uvw[:, k] * s) for k in range(3))
    wp = None if wgt is None else np.ascontiguousarray(wgt)
    tol = 3 * eps + 1e-11 + (3e-6 if prec else 0)
    dirty = np.zeros((nx, ny), dtype=np.complex128)
    code = d.drv_gridder(1, nrow, nx, ny, fov, f0, u, v, w, pirange, sign, _ptr(wp), vis.copy(), dirty, eps, 2.0,
                         kerev, prec)
    assert code == 0, (code, _last_error())
    want = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, sign)
    e1 = _err(want, dirty)
    out = np.zeros(nrow, dtype=np.complex128)
    code = d.drv_gridder(0, nrow, nx, ny, fov, f0, u, v, w, pirange, sign, _ptr(wp), out, img.copy(), eps, 2.0,
                         kerev, prec)
    assert code == 0, (code, _last_error())
    want_v = O.dft_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, sign)
    e2 = _err(want_v, out)
    say(f"  ms2dirty err {e1:.2e}, dirty2ms err {e2:.2e}, tol {tol:.1e}")
    assert e1 < tol and e2 < tol


def test_gridder_plan_api_rejects_bad_options():
    d = driver()
    n = 32
    u = v = w = np.zeros(4)
    vis = np.zeros(4, complex)
    img = np.zeros((n, n), complex)
    # w_term_method other than 1 is not provided; kerevalmeth 2 does not exist;
    # epsilon below 1.1e-16 returns 1 as the reference's setup does
    assert d.drv_gridder(1, 4, n, n, 2.0, 1e9, u, v, w, 0, -1, None, vis, img, 1e-6, 2.0, 2, 0) != 0
    assert d.drv_gridder(1, 4, n, n, 2.0, 1e9, u, v, w, 0, -1, None, vis, img, 1e-17, 2.0, 0, 0) == 1


NUFFT_CASES = [  # type, N, precision, kerevalmeth, direction
    (1, (64, 64), 0, 0, 1), (2, (64, 64), 0, 0, 0), (1, (48, 40), 1, 1, 0), (2, (100,), 0, 0, 1),
    (1, (20, 16, 6), 0, 0, 1), (2, (20, 16, 6), 1, 0, 0),
]


@pytest.mark.parametrize("case", NUFFT_CASES, ids=[f"t{c[0]}_{'x'.join(map(str, c[1]))}_p{c[2]}_k{c[3]}_d{c[4]}"
                                                  for c in NUFFT_CASES])
def test_nufft_plan_api(case):
    type_, N, prec, kerev, direction = case
    say(f"[plan api nufft] type {type_} N {N} precision {prec} kereval {kerev} direction {direction}")
    d = driver()
    dim = len(N)
    M = 1500
    rng = np.random.default_rng(7 * type_ + dim)
    coords = [np.ascontiguousarray(rng.uniform(-np.pi, np.pi, M) * 1.3) for _ in range(dim)]
    ptrs = [_ptr(coords[k]) if k < dim else None for k in range(3)]
    Nf = tuple(N) + (1,) * (3 - dim)
    eps = 1e-6 if prec == 0 else 1e-5
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fk = rng.standard_normal(tuple(N[::-1])) + 1j * rng.standard_normal(tuple(N[::-1]))
    iflag = 1 if direction > 0 else -1
    if type_ == 1:
        out = np.zeros(tuple(N[::-1]), dtype=np.complex128)
        code = d.drv_nufft(1, dim, Nf[0], Nf[1], Nf[2], M, *ptrs, None, direction, eps, 2.0, kerev, prec, c.copy(),
                           out)
        want = R.nufft_direct(1, N, coords, c, iflag)
    else:
        out = np.zeros(M, dtype=np.complex128)
        code = d.drv_nufft(2, dim, Nf[0], Nf[1], Nf[2], M, *ptrs, None, direction, eps, 2.0, kerev, prec, out,
                           np.ascontiguousarray(fk))
        want = R.nufft_direct(2, N, coords, fk, iflag)
    assert code == 0, (code, _last_error())
    err = _err(want, out)
    tol = 3 * eps + 1e-11 + (3e-6 if prec else 0)
    say(f"  err {err:.2e} tol {tol:.1e}")
    assert err < tol


@pytest.mark.parametrize("direction", [1, 0])
def test_nufft_type3_plan_api(direction):
    """The reference driver's type-3 case (test/nufft_1d_3_1_test.cu,
    nufft_1d_3_2_test.cu: x_j = 2 + j pi / M, s_i = -i / N)."""
    say(f"[plan api nufft type 3] direction {direction}")
    d = driver()
    M, N, eps = 1000, 200, 1e-6
    x = np.ascontiguousarray(2.0 + np.arange(M) * np.pi / M)
    s = np.ascontiguousarray(-np.arange(N) / N)
    rng = np.random.default_rng(3 + direction)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    iflag = 1 if direction == 1 else -1
    if direction == 1:
        out = np.zeros(N, dtype=np.complex128)
        code = d.drv_nufft(3, 1, N, 1, 1, M, _ptr(x), None, None, _ptr(s), 1, eps, 2.0, 0, 0, c.copy(), out)
        want = R.nufft3_direct(x, s, c, iflag, 1)
    else:
        out = np.zeros(M, dtype=np.complex128)
        code = d.drv_nufft(3, 1, N, 1, 1, M, _ptr(x), None, None, _ptr(s), 0, eps, 2.0, 0, 0, out, f.copy())
        want = R.nufft3_direct(x, s, f, iflag, 0)
    assert code == 0, (code, _last_error())
    err = _err(want, out)
    say(f"  err {err:.2e}")
    assert err < 3 * eps + 1e-11
