"""Generic NUFFT (types 1 and 2) without a GPU: the oracle's direct sums pinned
against numpy's FFT on uniform points, against the oracle's NIFTY direct sum
(the same 2-D transform reached through pixel sizes and frequency), against
separability in 3-D, 2 pi periodicity and exact adjointness; the library's
argument errors before any device work."""
import numpy as np
import pytest

from curig_amd import _lib
from oracle import nufft_ref as R
from tests import oracle_helpers as O


@pytest.mark.parametrize("iflag", [1, -1])
@pytest.mark.parametrize("N", [16, 15])
def test_type1_uniform_points_is_fft(N, iflag):
    rng = np.random.default_rng(N)
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    u = 2 * np.pi * np.arange(N) / N
    got = R.nufft_direct(1, (N,), (u,), c, iflag)
    k = np.arange(N) - N // 2
    full = np.fft.ifft(c) * N if iflag > 0 else np.fft.fft(c)
    np.testing.assert_allclose(got, full[k % N], rtol=1e-12, atol=1e-12)


def test_type1_2d_equals_nifty_form_without_w():
    # u_nufft pairs with the fast index (N1 = ny), v_nufft with the slow one (N2 = nx)
    rng = np.random.default_rng(3)
    M, N1, N2 = 40, 10, 7
    u, v = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    dx = dy = 1e-3
    freq = 299792458.0 / (2 * np.pi * dx)      # 2 pi f/c dx = 1: metres are radians per pixel
    uvw = np.stack([v, u, np.zeros(M)], axis=1)
    want = O.dft_nifty_ms2dirty(uvw, freq, c, None, None, N2, N1, dx, dy, apply_w=False)
    got = R.nufft_direct(1, (N1, N2), (u, v), c, 1)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


def test_type1_3d_separable_and_periodic():
    rng = np.random.default_rng(4)
    M, N = 30, (6, 5, 4)
    u, v = rng.uniform(-np.pi, np.pi, M), rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f3 = R.nufft_direct(1, N, (u, v, np.zeros(M)), c, -1)
    f2 = R.nufft_direct(1, N[:2], (u, v), c, -1)
    for k3 in range(N[2]):
        np.testing.assert_allclose(f3[k3], f2, rtol=1e-12, atol=1e-12)
    w = rng.uniform(-np.pi, np.pi, M)
    a = R.nufft_direct(1, N, (u, v, w), c, 1)
    b = R.nufft_direct(1, N, (u + 2 * np.pi, v - 4 * np.pi, w + 6 * np.pi), c, 1)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10 * np.abs(a).max())


@pytest.mark.parametrize("N", [(9,), (8, 5), (4, 6, 3)])
def test_type1_type2_adjoint(N):
    rng = np.random.default_rng(len(N))
    M = 25
    coords = tuple(rng.uniform(-np.pi, np.pi, M) for _ in N)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fk = rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1])
    for iflag in (1, -1):
        a = R.nufft_direct(1, N, coords, c, iflag)
        b = R.nufft_direct(2, N, coords, fk, -iflag)   # the adjoint flips the sign
        assert abs(np.vdot(fk, a) - np.vdot(b, c)) <= 1e-12 * abs(np.vdot(fk, a))


def test_nufft_argument_errors():
    lib = _lib.load()

    def call(type_=1, dim=2, N=(8, 8, 1), M=4, iflag=1, eps=1e-6, sigma=2.0, prec=0, fk=1):
        return lib.curig_nufft_dev(type_, dim, N[0], N[1], N[2], M, 8, 8, 8, 8, fk, iflag, eps, sigma, prec, None)

    assert call(eps=1e-17) == 1
    assert call(sigma=1.0) == 2
    assert call(type_=3) == 5
    assert call(dim=4) == 5
    assert call(iflag=0) == 5
    assert call(M=-1) == 4
    assert call(N=(0, 8, 1)) == 3
    assert call(dim=3, N=(8, 8, 0)) == 3
    assert call(prec=2) == 5
    assert call(fk=None) == 5


@pytest.mark.parametrize("iflag", [1, -1])
def test_type3_on_integer_outputs_is_type1(iflag):
    # outputs at the centred integer modes turn type 3 into type 1 (and its
    # adjoint direction into type 2)
    rng = np.random.default_rng(9)
    M, N = 30, 11
    x = rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    s = np.arange(N) - N // 2
    np.testing.assert_allclose(R.nufft3_direct(x, s, c, iflag, 1), R.nufft_direct(1, (N,), (x,), c, iflag),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(R.nufft3_direct(x, s, f, iflag, 0), R.nufft_direct(2, (N,), (x,), f, iflag),
                               rtol=1e-12, atol=1e-12)


def test_type3_directions_adjoint():
    rng = np.random.default_rng(10)
    x, s = rng.uniform(-30, 30, 20), rng.uniform(-5, 5, 13)
    c = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    f = rng.standard_normal(13) + 1j * rng.standard_normal(13)
    a = R.nufft3_direct(x, s, c, 1, 1)
    b = R.nufft3_direct(x, s, f, -1, 0)
    assert abs(np.vdot(f, a) - np.vdot(b, c)) <= 1e-12 * abs(np.vdot(f, a))


def test_type3_argument_errors():
    lib = _lib.load()

    def call(M=4, N=5, iflag=1, eps=1e-6, prec=0, s=8, f=8):
        return lib.curig_nufft3_dev(M, N, 8, s, 8, f, iflag, 1, eps, 2.0, prec, None)

    assert call(eps=1e-17) == 1
    assert call(iflag=2) == 5
    assert call(M=-1) == 4
    assert call(N=0) == 3
    assert call(s=None) == 5
    assert call(prec=3) == 5
