"""NIFTY form without a GPU: the oracle's NIFTY direct sum pinned against the
reference-form direct sum, closed forms and adjointness; the library's host-only
plan for pixel sizes, grid sizes and the 2-D form against the oracle's plan
math; argument errors returned before any device work."""
import numpy as np
import pytest

from curig_amd import _lib
from oracle import nufft_ref as R
from tests import oracle_helpers as O

C = 299792458.0


def _case(seed, nrow, nchan, umax=600.0, wmax=300.0):
    rng = np.random.default_rng(seed)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.uniform(-umax, umax, (nrow, 2))
    uvw[:, 2] = rng.uniform(-wmax, wmax, nrow)
    freq = np.linspace(1.0e9, 1.4e9, nchan)
    ms = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    wgt = rng.uniform(0.5, 1.0, (nrow, nchan))
    mask = (rng.random((nrow, nchan)) > 0.3).astype(np.uint8)
    return uvw, freq, ms, wgt, mask, rng


def test_oracle_nifty_single_channel_equals_reference_form():
    uvw, freq, ms, wgt, _, _ = _case(1, 40, 1)
    nx, ny, fov = 12, 10, 1.5
    dx, dy = fov * np.pi / 180 / nx, fov * np.pi / 180 / ny
    a = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, None, nx, ny, dx, dy)
    b = O.dft_ms2dirty(uvw, ms[:, 0], wgt[:, 0], nx, ny, fov, freq[0], sign=-1)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
    img = np.random.default_rng(2).standard_normal((nx, ny)) + 0j
    v1 = O.dft_nifty_dirty2ms(uvw, freq, img, wgt, None, dx, dy)
    v2 = O.dft_dirty2ms(uvw, img, wgt[:, 0], nx, ny, fov, freq[0], sign=-1)
    np.testing.assert_allclose(v1[:, 0], v2, rtol=0, atol=1e-12 * np.abs(v2).max())


def test_oracle_nifty_channels_and_mask():
    uvw, freq, ms, wgt, mask, _ = _case(3, 30, 3)
    nx, ny, dx, dy = 10, 14, 2e-4, 3e-4
    whole = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy)
    parts = sum(O.dft_nifty_ms2dirty(uvw, freq[c], ms[:, c] * mask[:, c], wgt[:, c], None, nx, ny, dx, dy)
                for c in range(3))
    np.testing.assert_allclose(whole, parts, rtol=0, atol=1e-12 * np.abs(parts).max())
    img = np.random.default_rng(4).standard_normal((nx, ny)) + 0j
    v = O.dft_nifty_dirty2ms(uvw, freq, img, wgt, mask, dx, dy)
    assert np.all(v[mask == 0] == 0)
    for c in range(3):
        vc = O.dft_nifty_dirty2ms(uvw, freq[c], img, wgt[:, c], None, dx, dy)[:, 0]
        np.testing.assert_allclose(v[:, c], vc * mask[:, c], rtol=0, atol=1e-12 * np.abs(vc).max())


@pytest.mark.parametrize("apply_w", [True, False])
def test_oracle_nifty_plane_wave_closed_form(apply_w):
    # one sample: a plane wave e^{i 2pi f/c (u l + v m - w (n-1))} / n, or without
    # the w term and the 1/n in the 2-D form; rectangular pixels and image so that
    # a swapped axis or pixel size fails
    u, v, w, f = 350.0, -120.0, 230.0, 1.2e9
    nx, ny, dx, dy = 9, 6, 2.0e-3, 3.5e-3
    vis = 0.7 - 0.4j
    got = O.dft_nifty_ms2dirty(np.array([[u, v, w]]), f, vis, None, None, nx, ny, dx, dy, apply_w=apply_w)
    l = (np.arange(nx) - nx // 2)[:, None] * dx
    m = (np.arange(ny) - ny // 2)[None, :] * dy
    n = np.sqrt(1 - l ** 2 - m ** 2)
    if apply_w:
        want = vis * np.exp(2j * np.pi * f / C * (u * l + v * m - w * (n - 1))) / n
    else:
        want = vis * np.exp(2j * np.pi * f / C * (u * l + v * m)) * np.ones_like(n)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("apply_w", [True, False])
def test_oracle_nifty_adjoint(apply_w):
    uvw, freq, ms, wgt, mask, rng = _case(5, 25, 2)
    nx, ny, dx, dy = 8, 12, 3e-4, 2e-4
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    a = O.dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy, apply_w=apply_w)
    b = O.dft_nifty_dirty2ms(uvw, freq, img, wgt, mask, dx, dy, apply_w=apply_w)
    lhs = np.vdot(img, a)
    rhs = np.vdot(b, ms)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_nifty_plan_defaults_equal_fov_plan():
    for nx, ny, fov, eps in ((256, 256, 2.0, 1e-6), (2048, 1024, 4.0, 1e-5), (30, 250, 10.0, 5e-7)):
        dx, dy = fov * np.pi / 180 / nx, fov * np.pi / 180 / ny
        a = _lib.plan_info(10, nx, ny, fov, 1.4e9, eps, 2.0, -1, -800.0, 1100.0)
        b = _lib.nifty_plan_info(10, nx, ny, dx, dy, 1.4e9, eps, 0, 0, True, -800.0, 1100.0)
        for k in a:
            assert b[k] == pytest.approx(a[k], rel=1e-13, abs=1e-15), k


@pytest.mark.parametrize("nx,ny,nu,nv", [(64, 48, 160, 96), (64, 64, 96, 96), (100, 60, 150, 130)])
def test_nifty_plan_grid_size(nx, ny, nu, nv):
    dx, dy = 2.5e-4, 3.5e-4
    eps, freq, w0, w1 = 1e-6, 1.3e9, -300.0, 250.0
    got = _lib.nifty_plan_info(10, nx, ny, dx, dy, freq, eps, nu, nv, True, w0, w1)
    assert (got["nf2"], got["nf1"]) == (nu, nv)
    sigma = min(nu / nx, nv / ny)
    # kernel and w planes from the oracle's plan math at that upsampling factor
    _, kw, beta, c = R.setup_conv_opts(eps, sigma)
    assert got["kw"] == kw
    assert got["beta"] == pytest.approx(beta, rel=1e-13)
    assert got["c"] == pytest.approx(c, rel=1e-13)
    # P depends on S (from both pixel sizes) and X: recompute with the oracle's routine
    lmin, mmin = -0.5 * dx * nx, -0.5 * dy * ny
    S = abs((np.sqrt(1 - lmin ** 2 - mmin ** 2) - 1) / 2)
    fc = 2 * np.pi * freq / C
    X = abs(w1 - w0) * fc / 2
    P, h, gam = R.set_nhg_type3(S, X, kw, sigma)
    assert (got["nf3"], got["h"], got["gamma"]) == (P, pytest.approx(h, rel=1e-13), pytest.approx(gam, rel=1e-13))


def test_nifty_plan_without_wstacking():
    # the 2-D form: every w taken as 0, one plane (the plain 2-D NUFFT)
    nx, ny, dx, dy, eps = 128, 96, 2e-4, 2.5e-4, 1e-6
    got = _lib.nifty_plan_info(10, nx, ny, dx, dy, 1.2e9, eps, 0, 0, False, -500.0, 700.0)
    want = _lib.nifty_plan_info(10, nx, ny, dx, dy, 1.2e9, eps, 0, 0, True, -500.0, 700.0)
    assert got["i_center"] == 0.0
    assert (got["nf3"], got["nbatch"]) == (1, 1)
    for k in ("kw", "nf1", "nf2", "beta", "c", "pixelsize_x", "pixelsize_y"):
        assert got[k] == want[k], k


def test_nifty_argument_errors():
    lib = _lib.load()
    z = np.zeros((4, 3))
    f = np.array([1e9])
    ms = np.zeros((4, 1), complex)
    d = np.zeros((8, 8), complex)

    def call(**kw):
        a = dict(nrow=4, nchan=1, nx=8, ny=8, dx=1e-3, dy=1e-3, nu=0, nv=0, eps=1e-6, ws=1)
        a.update(kw)
        return lib.curig_nifty_ms2dirty(a["nrow"], a["nchan"], z, f, ms.ctypes.data, None, None, a["nx"], a["ny"],
                                        a["dx"], a["dy"], a["nu"], a["nv"], a["eps"], a["ws"], d.ctypes.data, 0)

    assert call(eps=1e-17) == 1
    assert call(nu=17, nv=16) == 3      # odd grid size
    assert call(nu=8, nv=16) == 3       # grid not larger than the image
    assert call(nu=16, nv=0) == 3       # one axis only
    assert call(dx=0.0) == 9            # pixel size
    assert call(nchan=0) == 5
    assert call(dx=0.3, dy=0.3) == 5    # beyond the horizon
