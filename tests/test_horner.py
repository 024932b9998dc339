"""The piecewise-polynomial kernel (Horner evaluation) without a GPU: the
library's plan-time fit checked against the oracle's ES kernel on a fine grid
of every interval, at the tolerance the plan uses (a quarter of epsilon, not
below 1e-15 in FP64 or 3e-8 in FP32-mixed)."""
import numpy as np
import pytest

from curig_amd import _lib
from oracle import nufft_ref as R


def _fit(eps, sigma, precision):
    lib = _lib.load()
    deg = np.zeros(1, dtype=np.int32)
    coef = np.zeros(28 * 25)
    assert lib.curig_kernel_horner(eps, sigma, precision, deg, coef, coef.size) == 0, lib.curig_last_error()
    return int(deg[0]), coef


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("eps,sigma", [(1e-2, 2.0), (1e-4, 2.0), (1e-5, 2.0), (1e-6, 2.0), (1e-9, 2.0),
                                       (1e-12, 2.0), (1e-6, 1.5), (1e-5, 1.25)])
def test_horner_fit_matches_kernel(eps, sigma, precision):
    _, kw, beta, c = R.setup_conv_opts(eps, sigma)
    deg, coef = _fit(eps, sigma, precision)
    assert kw - 1 <= deg <= 24
    coef = coef[: kw * (deg + 1)].reshape(kw, deg + 1)
    tol = max(0.25 * eps, 1e-15 if precision == 0 else 3e-8)
    t = np.linspace(0.0, 1.0, 1001)
    worst = 0.0
    for i in range(kw):
        x = -kw / 2 + i + t
        want = R.es_kernel(x, kw, beta, c) * np.exp(-beta)   # normalised: peak 1
        got = np.polyval(coef[i], 2 * t - 1)
        worst = max(worst, float(np.max(np.abs(got - want))))
    assert worst < tol, (deg, worst, tol)


def test_kernel_eval_switch():
    lib = _lib.load()
    assert lib.curig_set_kernel_eval(2) == 5
    assert lib.curig_set_kernel_eval(1) == 0
    assert lib.curig_set_kernel_eval(0) == 0
