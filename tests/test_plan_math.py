"""Host plan math: the oracle's restatement pinned by closed forms and library
routines, and the C++ library's host-only plan functions checked against it.

No GPU is touched here (the library's plan functions are pure host code)."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import nufft_ref as R


@pytest.mark.parametrize("eps,kw,beta", [
    (1e-1, 2, 4.40), (1e-2, 3, 6.78), (1e-3, 4, 9.52), (1e-4, 5, 11.5), (1e-5, 6, 13.8),
    (1e-6, 7, 16.1), (1e-12, 13, 29.9), (5e-7, 8, 18.4), (2e-1, 2, 4.40)])
def test_kw_beta_table(eps, kw, beta):
    # kw = ceil(-log10(eps/10)), beta = betaoverns*kw  (conv_interp_invoker.cu:84-112)
    ier, k, b, c = R.setup_conv_opts(eps, 2.0)
    assert ier == 0 and k == kw and abs(b - beta) < 1e-9 and abs(c - 4.0 / kw ** 2) < 1e-15


def test_conv_opts_errors():
    assert R.setup_conv_opts(1e-17, 2.0)[0] == 1      # clamped eps -> warning code 1
    assert R.setup_conv_opts(1e-6, 1.0)[0] == 2       # sigma <= 1 -> code 2
    ier, kw, beta, c = R.setup_conv_opts(1e-6, 1.25)  # custom sigma formula
    assert ier == 0 and kw == math.ceil(-math.log(1e-6) / (math.pi * math.sqrt(1 - 1 / 1.25)))
    assert abs(beta - 0.97 * math.pi * (1 - 1 / 2.5) * kw) < 1e-12


def _smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def test_next235beven_brute_force():
    for n in range(0, 3000):
        got = R.next235beven(n, 1)
        want = next(m for m in range(max(n, 2), 10 ** 6) if m % 2 == 0 and _smooth(m))
        assert got == want, n


def test_get_num_cells():
    assert R.get_num_cells(256, 7, 2.0) == 512
    assert R.get_num_cells(8192, 6, 2.0) == 16384
    assert R.get_num_cells(30, 7, 2.0) == 60
    assert R.get_num_cells(3, 7, 2.0) == 16            # nf >= 2 kw = 14 -> 2-3-5 smooth 16
    assert R.get_num_cells(250, 8, 2.0) == 500


def test_set_nhg_type3_closed_form():
    # nf = next235even(max(int(2 sigma S X / pi + kw + 1), 2 kw)), h = 2pi/nf, gamma = nf/(2 sigma S)
    S, X, kw = 0.00137, 1.02e6, 6
    nf, h, g = R.set_nhg_type3(S, X, kw, 2.0)
    base = int(2 * 2.0 * S * X / math.pi + kw + 1)
    assert nf == R.next235beven(base) and abs(h - 2 * math.pi / nf) < 1e-15
    assert abs(g - nf / (4.0 * S)) < 1e-9 * g
    # safeguards (common.cpp:178-184)
    nf, h, g = R.set_nhg_type3(0.0, 0.0, 7, 2.0)
    assert nf == 16 and abs(g - 16 / 4.0) < 1e-15
    nf, h, g = R.set_nhg_type3(0.01, 0.0, 7, 2.0)      # X -> 1/S
    assert nf == R.next235beven(max(int(4 * 0.01 * 100 / math.pi + 8), 14))
    nf, h, g = R.set_nhg_type3(1e-9, 5.0, 7, 2.0)      # S -> 1/X
    assert abs(g - nf / (4.0 * 0.2)) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 22, 34, 60])
def test_gauss_legendre_vs_numpy(n):
    x, w = R.gauss_legendre(n)
    X, W = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, X, atol=1e-14)
    np.testing.assert_allclose(w, W, atol=1e-14)


@pytest.mark.parametrize("eps", [1e-4, 1e-6, 1e-12])
def test_fseries_grid_vs_quad(eps):
    _, kw, beta, c = R.setup_conv_opts(eps, 2.0)
    nf = 64
    fk = R.fseries_grid(nf, kw, beta, c)
    # the 2p-node rule resolves the kernel's edge to ~1e-4 eps (reference design,
    # deconv.cu:75), far inside the gridder's own error budget
    tol = max(1e-3 * eps, 1e-13) * abs(fk[0])
    for k in [0, 1, 5, 17, 32]:
        f = lambda x: math.exp(beta * math.sqrt(1 - c * x * x)) * math.cos(2 * math.pi * k * x / nf)
        val, _ = integrate.quad(f, -kw / 2, kw / 2, epsabs=0, epsrel=1e-13, limit=200)
        assert abs(fk[k] - (-1) ** k * val) < tol


@pytest.mark.parametrize("eps", [1e-5, 1e-12])
def test_fseries_at_vs_quad(eps):
    _, kw, beta, c = R.setup_conv_opts(eps, 2.0)
    for z in [0.0, 0.3, 1.1, math.pi / 2]:
        f = lambda x: math.exp(beta * math.sqrt(1 - c * x * x)) * math.cos(x * z)
        val, _ = integrate.quad(f, -kw / 2, kw / 2, epsabs=0, epsrel=1e-13, limit=200)
        got = float(R.fseries_at(np.array([z]), kw, beta, c)[0])
        assert abs(got - val) < max(1e-3 * eps, 1e-13) * abs(val)
