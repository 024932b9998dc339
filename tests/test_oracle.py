"""Pins the oracle (oracle/dft.c and oracle/nufft_ref.py) before it is trusted.

Each check is chosen so that a plausible mistake fails it:
* the DFT reduces to numpy.fft.fft2 on a lattice with w = 0 (pins the u<->l,
  v<->m pairing, the sign of the exponent and the integer-halved pixel centre);
* a single row at u = v = 0 isolates the w term (sign * w * (n-1), division by n);
* exact adjointness of the DFT pair and of the gridder restatement (the
  reference asserts rtol 1e-12, python/curagridder/test_wst.py:77);
* the gridder restatement (grid + FFT + w-DFT + corrections) agrees with the
  direct sum (two unrelated algorithms) at the requested epsilon;
* the committed golden fixtures (tests/golden/) hold the DFT's outputs.
"""
import os

import numpy as np
import pytest

from oracle import nufft_ref as R
from tests import oracle_helpers as O

C = 299792458.0
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rand_case(rng, nrow, nx, scale=1.0, f0=1e9):
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * scale * nx / 2
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.random(nrow) + 0.5
    return uvw, vis, wgt, f0


def test_dft_lattice_reduces_to_fft():
    # x_j = 2 pi f/c u_j dx = 2 pi a_j / nx with integer a_j, w = 0:
    # dirty[i][k] = sum_j V_j e^{+2 pi i (a_j (i - nx/2)/nx + b_j (k - ny/2)/ny)} / n_ik
    rng = np.random.default_rng(3)
    nx, ny, fov, f0 = 16, 12, 3.0, 1.3e9
    dx = fov * np.pi / 180 / nx
    dy = fov * np.pi / 180 / ny
    nrow = 40
    a = rng.integers(-nx // 2, nx // 2, nrow)
    b = rng.integers(-ny // 2, ny // 2, nrow)
    uvw = np.zeros((nrow, 3))
    uvw[:, 0] = a / nx / dx / (f0 / C)
    uvw[:, 1] = b / ny / dy / (f0 / C)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    got = O.dft_ms2dirty(uvw, vis, None, nx, ny, fov, f0, sign=-1)
    grid = np.zeros((nx, ny), dtype=np.complex128)
    np.add.at(grid, (a % nx, b % ny), vis)
    img = np.fft.ifft2(grid) * nx * ny                      # e^{+2 pi i (a i / nx + b k / ny)}
    i = np.arange(nx)[:, None]
    k = np.arange(ny)[None, :]
    # shift the origin to the integer-halved pixel centre: index i -> i - nx/2
    img = np.roll(img, (nx // 2, ny // 2), axis=(0, 1))
    l = (i - nx // 2) * dx
    m = (k - ny // 2) * dy
    n = np.sqrt(1 - l * l - m * m)
    np.testing.assert_allclose(got, img / n, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_w_term_single_row(sign):
    nx, ny, fov, f0 = 10, 14, 20.0, 1.4e9
    w_m = 37.5
    uvw = np.array([[0.0, 0.0, w_m]])
    vis = np.array([0.3 - 1.1j])
    wgt = np.array([0.7])
    got = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, sign=sign)
    i = np.arange(nx)[:, None]
    k = np.arange(ny)[None, :]
    l = (i - nx // 2) * fov * np.pi / 180 / nx
    m = (k - ny // 2) * fov * np.pi / 180 / ny
    n = np.sqrt(1 - l * l - m * m)
    want = 0.7 * vis[0] * np.exp(2j * np.pi * f0 / C * sign * w_m * (n - 1)) / n
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # the image centre is exactly wgt * V (n = 1 there)
    assert abs(got[nx // 2, ny // 2] - 0.7 * vis[0]) < 1e-15


def test_dft_uvw_zero_is_constant_over_n():
    rng = np.random.default_rng(5)
    nrow, nx, ny, fov = 25, 8, 6, 7.0
    uvw = np.zeros((nrow, 3))
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.random(nrow)
    got = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, 1e9)
    l = (np.arange(nx)[:, None] - nx // 2) * fov * np.pi / 180 / nx
    m = (np.arange(ny)[None, :] - ny // 2) * fov * np.pi / 180 / ny
    np.testing.assert_allclose(got, np.sum(wgt * vis) / np.sqrt(1 - l * l - m * m), rtol=1e-13)


@pytest.mark.parametrize("use_wgt", [False, True])
def test_dft_adjointness(use_wgt):
    rng = np.random.default_rng(7)
    nx, ny = 18, 22
    uvw, vis, wgt, f0 = _rand_case(rng, 60, nx, scale=3.0)
    wgt = wgt if use_wgt else None
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    d = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, 5.0, f0)
    v = O.dft_dirty2ms(uvw, img, wgt, nx, ny, 5.0, f0)
    np.testing.assert_allclose(np.vdot(vis, v), np.vdot(d, img), rtol=1e-12)


def test_dft_pixels_subset_matches_full():
    rng = np.random.default_rng(9)
    nx, ny = 20, 16
    uvw, vis, wgt, f0 = _rand_case(rng, 50, nx, scale=2.0)
    full = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, 4.0, f0)
    pix = rng.choice(nx * ny, 37, replace=False)
    sub = O.dft_ms2dirty_pixels(uvw, vis, wgt, nx, ny, 4.0, f0, pix)
    np.testing.assert_allclose(sub, full.ravel()[pix], rtol=0, atol=1e-13)


CASES = [
    # nrow, nx, ny, fov (deg), eps, uv scale (x Nyquist), sign, use_wgt
    (100, 32, 48, 10.0, 1e-6, 1.0, -1, True),
    (200, 64, 64, 5.0, 1e-12, 1.0, -1, False),
    (300, 40, 40, 20.0, 1e-4, 2.9, 1, True),      # |x| up to ~2.9 pi (reference's single wrap)
    (150, 24, 36, 2.0, 1e-5, 4.7, -1, True),      # |x| beyond 3 pi: folded (DFT definition)
    (50, 30, 250, 1.0, 5e-7, 1.0, -1, False),
    (80, 31, 17, 8.0, 1e-6, 1.5, -1, True),       # odd sizes, integer-halved centre
]


@pytest.mark.parametrize("case", CASES)
def test_gridder_restatement_vs_dft(case):
    nrow, nx, ny, fov, eps, scale, sign, use_wgt = case
    rng = np.random.default_rng(nrow + nx)
    uvw, vis, wgt, f0 = _rand_case(rng, nrow, nx, scale)
    wgt = wgt if use_wgt else None
    d = R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps, 2.0, sign)
    truth = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, sign)
    tol = 3 * eps + 1e-11
    assert O.l2error(truth, d) < tol
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    v = R.dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps, 2.0, sign)
    vtruth = O.dft_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, sign)
    assert O.l2error(vtruth, v) < tol
    # exact adjointness of the discrete operator pair
    np.testing.assert_allclose(np.vdot(vis, v), np.vdot(d, img), rtol=1e-12)


def test_gridder_restatement_fold_invariance():
    # x -> x + 2 pi (u -> u + 1/(f/c dx)) leaves every output unchanged
    rng = np.random.default_rng(11)
    nx, ny, fov, eps = 24, 24, 6.0, 1e-8
    uvw, vis, wgt, f0 = _rand_case(rng, 60, nx, 0.8)
    d0 = R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps, 2.0, -1)
    uvw2 = uvw.copy()
    uvw2[:7, 0] += 1.0 / (f0 / C * fov * np.pi / 180 / nx)
    uvw2[3:9, 1] -= 2.0 / (f0 / C * fov * np.pi / 180 / ny)
    d1 = R.ms2dirty(uvw2, vis, wgt, nx, ny, fov, f0, eps, 2.0, -1)
    assert O.l2error(d0, d1) < 1e-11


def test_golden_fixtures_hold_dft_outputs():
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files, "no golden fixtures committed"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f), allow_pickle=False)
        nx, ny = int(z["nx"]), int(z["ny"])
        fov, freq, sign = float(z["fov"]), float(z["freq"]), int(z["sign"])
        wgt = z["wgt"] if z["wgt"].size else None
        d = O.dft_ms2dirty(z["uvw"], z["vis"], wgt, nx, ny, fov, freq, sign)
        np.testing.assert_allclose(d, z["dirty"], rtol=0, atol=1e-12 * np.abs(z["dirty"]).max())
        v = O.dft_dirty2ms(z["uvw"], z["image"], wgt, nx, ny, fov, freq, sign)
        np.testing.assert_allclose(v, z["vis_out"], rtol=0, atol=1e-12 * np.abs(z["vis_out"]).max())
