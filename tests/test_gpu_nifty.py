"""GPU parity of the NIFTY form (curig_nifty_ms2dirty / curig_nifty_dirty2ms and
the Python API curig_amd.cursl) against the oracle's NIFTY direct sum: several
channels with their own frequencies, weights and a mask, rectangular pixels,
given grid sizes and the 2-D form (do_wstacking=False), in FP64 and FP32-mixed,
on the rocFFT path (non power-of-two grids) and the LDS path (power-of-two).
Bars as in the other parity tests: 3 eps + 1e-11 for FP64 (+3e-6 FP32-mixed),
10 eps where a given grid sets another upsampling factor than 2; adjointness
at rtol 1e-12 in FP64."""
import sys

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(seed, nrow, nchan, nx, ny, dx, dy, wmax=300.0):
    rng = np.random.default_rng(seed)
    freq = np.linspace(1.0e9, 1.4e9, nchan)
    C = 299792458.0
    uvw = np.empty((nrow, 3))
    # u, v up to the grid's band limit at the highest frequency
    uvw[:, 0] = rng.uniform(-1, 1, nrow) * 0.5 / dx * C / freq[-1]
    uvw[:, 1] = rng.uniform(-1, 1, nrow) * 0.5 / dy * C / freq[-1]
    uvw[:, 2] = rng.uniform(-wmax, wmax, nrow)
    ms = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    wgt = rng.uniform(0.5, 1.0, (nrow, nchan))
    mask = (rng.random((nrow, nchan)) > 0.25).astype(np.uint8)
    return uvw, freq, ms, wgt, mask, rng


def _tol(grid, eps, precision):
    # 3 eps at upsampling 2; the kernel chosen for another upsampling factor
    # (FINUFFT's width formula, conv_interp_invoker.cu:46-116) reaches ~4 eps
    nx, ny, nu, nv = grid
    sigma = 2.0 if nu == 0 else min(nu / nx, nv / ny)
    return (3 if sigma == 2.0 else 10) * eps + 1e-11 + (3e-6 if precision else 0)


def _lib():
    from curig_amd import _lib as L

    return L


def nifty(to_image, uvw, freq, data, wgt, mask, nx, ny, dx, dy, nu, nv, eps, wstack, precision):
    L = _lib()
    lib = L.load()
    nrow, nchan = uvw.shape[0], freq.shape[0]
    cdt, rdt = (np.complex128, np.float64) if precision == 0 else (np.complex64, np.float32)
    w = None if wgt is None else np.ascontiguousarray(wgt, dtype=rdt)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    wp = None if w is None else w.ctypes.data
    mp = None if m is None else m.ctypes.data
    if to_image:
        ms = np.ascontiguousarray(data, dtype=cdt)
        out = np.zeros((nx, ny), dtype=cdt)
        code = lib.curig_nifty_ms2dirty(nrow, nchan, uvw, freq, ms.ctypes.data, wp, mp, nx, ny, dx, dy, nu, nv, eps,
                                        int(wstack), out.ctypes.data, precision)
    else:
        img = np.ascontiguousarray(data, dtype=cdt)
        out = np.zeros((nrow, nchan), dtype=cdt)
        code = lib.curig_nifty_dirty2ms(nrow, nchan, uvw, freq, out.ctypes.data, wp, mp, nx, ny, dx, dy, nu, nv, eps,
                                        int(wstack), img.ctypes.data, precision)
    assert code == 0, lib.curig_last_error()
    return out.astype(np.complex128)


# (nx, ny, nu, nv): rocFFT path with the default grid, a given non-square grid
# (upsampling min(160/64, 120/48) = 2.5), and the LDS path (power-of-two grid)
GRIDS = [(64, 48, 0, 0), (64, 48, 160, 120), (64, 64, 128, 128)]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("wstack", [True, False])
@pytest.mark.parametrize("grid", GRIDS)
def test_nifty_ms2dirty_matches_direct_sum(grid, wstack, precision):
    nx, ny, nu, nv = grid
    dx, dy, eps = 2.5e-4, 3.5e-4, 1e-6
    uvw, freq, ms, wgt, mask, _ = _case(11, 3000, 3, nx, ny, dx, dy)
    say(f"[nifty m2d] grid {grid} wstack {wstack} precision {precision}")
    got = nifty(True, uvw, freq, ms, wgt, mask, nx, ny, dx, dy, nu, nv, eps, wstack, precision)
    want = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy, apply_w=wstack)
    err = O.l2error(want, got)
    tol = _tol(grid, eps, precision)
    say(f"  err {err:.3e} tol {tol:.1e}")
    assert err < tol
    if not wstack:
        # the w terms are large enough here that ignoring them is visible
        full = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy, apply_w=True)
        assert O.l2error(full, got) > 100 * tol


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("wstack", [True, False])
@pytest.mark.parametrize("grid", GRIDS)
def test_nifty_dirty2ms_matches_direct_sum(grid, wstack, precision):
    nx, ny, nu, nv = grid
    dx, dy, eps = 2.5e-4, 3.5e-4, 1e-6
    uvw, freq, _, wgt, mask, rng = _case(12, 2000, 3, nx, ny, dx, dy)
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    say(f"[nifty d2m] grid {grid} wstack {wstack} precision {precision}")
    got = nifty(False, uvw, freq, img, wgt, mask, nx, ny, dx, dy, nu, nv, eps, wstack, precision)
    want = O.dft_nifty_dirty2ms(uvw, freq, img, wgt, mask, dx, dy, apply_w=wstack)
    assert np.all(got[mask == 0] == 0)
    err = O.l2error(want, got)
    tol = _tol(grid, eps, precision)
    say(f"  err {err:.3e} tol {tol:.1e}")
    assert err < tol


@pytest.mark.parametrize("wstack", [True, False])
@pytest.mark.parametrize("grid", GRIDS)
def test_nifty_adjoint(grid, wstack):
    nx, ny, nu, nv = grid
    dx, dy, eps = 2.5e-4, 3.5e-4, 1e-6
    uvw, freq, ms, wgt, mask, rng = _case(13, 2500, 2, nx, ny, dx, dy)
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    say(f"[nifty adjoint] grid {grid} wstack {wstack}")
    a = nifty(True, uvw, freq, ms * mask, wgt, mask, nx, ny, dx, dy, nu, nv, eps, wstack, 0)
    b = nifty(False, uvw, freq, img, wgt, mask, nx, ny, dx, dy, nu, nv, eps, wstack, 0)
    lhs, rhs = np.vdot(img, a), np.vdot(b, ms * mask)
    say(f"  rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_nifty_python_api():
    from curig_amd import cursl

    nx, ny, dx, dy, eps = 64, 48, 2.5e-4, 3.5e-4, 1e-6
    uvw, freq, ms, wgt, mask, rng = _case(14, 2000, 2, nx, ny, dx, dy)
    say("[nifty api] ms2dirty / dirty2ms, 2 channels, mask")
    img = cursl.ms2dirty(uvw, freq, ms, wgt, nx, ny, dx, dy, 0, 0, eps, True, mask=mask)
    assert img.shape == (nx, ny) and img.dtype == np.float64
    want = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy).real
    assert O.l2error(want, img) < 3 * eps + 1e-11
    real_img = rng.standard_normal((nx, ny))
    vis = cursl.dirty2ms(uvw, freq, real_img, wgt, dx, dy, 0, 0, eps, True, mask=mask)
    assert vis.shape == (uvw.shape[0], 2) and vis.dtype == np.complex128
    want = O.dft_nifty_dirty2ms(uvw, freq, real_img + 0j, wgt, mask, dx, dy)
    assert O.l2error(want, vis) < 3 * eps + 1e-11
    # the reference's single-channel call: ms (nrow,), freq[0] used
    one = cursl.ms2dirty(uvw, freq, ms[:, 0], None, nx, ny, dx, dy, 0, 0, eps, True)
    want = O.dft_nifty_ms2dirty(uvw, freq[0], ms[:, 0], None, None, nx, ny, dx, dy).real
    assert O.l2error(want, one) < 3 * eps + 1e-11
    # complex64 input selects the FP32-mixed path
    img32 = cursl.ms2dirty(uvw, freq, ms.astype(np.complex64), wgt, nx, ny, dx, dy, 0, 0, 1e-5, True, mask=mask)
    assert img32.dtype == np.float32
    want = O.dft_nifty_ms2dirty(uvw, freq, ms.astype(np.complex64), wgt.astype(np.float32), mask, nx, ny, dx,
                                dy).real
    assert O.l2error(want, img32) < 3e-5 + 3e-6
