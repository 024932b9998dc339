"""Differentiation form: adjointness through the torch interface, agreement of
autograd gradients with the direct-sum adjoint, and central finite differences
of the imaging loss ||W^1/2 (A x - V)||^2 + lam TV(x) on random pixels."""
import sys

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(torch, nrow=20000, n=128, seed=3, eps=1e-7):
    from curig_amd.autograd import GridderGeometry

    rng = np.random.default_rng(seed)
    f0 = 1.4e9
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * n * 0.6
    uvw[:, 2] *= 8.0
    geo = GridderGeometry(uvw, f0, n, n, 3.0, epsilon=eps, wgt=rng.uniform(0.5, 1, nrow))
    return geo, uvw, rng, f0


def _array_case(torch, n=256, eps=1e-7):
    """64 antennas in a 3 km disk, 64 snapshots over +-4 h (129,024 rows), 1.4 GHz,
    fov 2 deg: the array geometry of the differentiation workload, scaled down."""
    from curig_amd.autograd import GridderGeometry

    sys.path.insert(0, __import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    from bench import make_uvw

    uvw = make_uvw(64, 3000.0, 64, 4.0, seed=5)
    rng = np.random.default_rng(3)
    f0 = 1.4e9
    geo = GridderGeometry(uvw, f0, n, n, 2.0, epsilon=eps, wgt=rng.uniform(0.5, 1, len(uvw)))
    return geo, uvw, rng, f0


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
@pytest.mark.parametrize("geometry", ["array", "dense"])
def test_adjointness_through_torch(dtype, geometry):
    """<A x, y> = <x, A^H y> for random complex x, y.  FP64: 1e-12.  FP32-mixed:
    1e-6 on array coverage.  The "dense" case packs 20,000 random baselines into
    the inner 0.6 % of the uv plane, so every grid cell there sums thousands of
    FP32 contributions: the FP32 spread's rounding (~ sqrt(count) x 6e-8) sets a
    1e-5 bound there."""
    torch = _torch()
    from curig_amd.autograd import dirty2ms, ms2dirty

    say(f"[autograd] adjointness {dtype} {geometry}")
    geo, uvw, rng, f0 = _array_case(torch) if geometry == "array" else _case(torch)
    cd = torch.complex128 if dtype == "fp64" else torch.complex64
    x = torch.tensor(rng.standard_normal((geo.nx, geo.ny)) + 1j * rng.standard_normal((geo.nx, geo.ny)),
                     dtype=cd, device="cuda")
    y = torch.tensor(rng.standard_normal(geo.nrow) + 1j * rng.standard_normal(geo.nrow), dtype=cd, device="cuda")
    ax = dirty2ms(x, geo)
    ahy = ms2dirty(y, geo)
    lhs = torch.vdot(ax.to(torch.complex128), y.to(torch.complex128)).item()
    rhs = torch.vdot(x.reshape(-1).to(torch.complex128), ahy.reshape(-1).to(torch.complex128)).item()
    rel = abs(lhs - rhs) / abs(lhs)
    say(f"[autograd] adjointness {dtype} {geometry}: {rel:.2e}")
    tol = 1e-12 if dtype == "fp64" else (1e-6 if geometry == "array" else 1e-5)
    assert rel < tol


def test_gradient_is_direct_sum_adjoint():
    """d/dx of Re<g, A x> for real x is Re(A^H g): compare with the oracle."""
    torch = _torch()
    from curig_amd.autograd import dirty2ms

    say("[autograd] gradient vs direct-sum adjoint")
    geo, uvw, rng, f0 = _case(torch, nrow=3000, n=64)
    x = torch.tensor(rng.standard_normal((64, 64)), dtype=torch.float64, device="cuda", requires_grad=True)
    g = rng.standard_normal(geo.nrow) + 1j * rng.standard_normal(geo.nrow)
    v = dirty2ms(x, geo)
    (torch.vdot(torch.tensor(g, device="cuda"), v).real).backward()
    want = O.dft_ms2dirty(uvw, g, geo.wgt.cpu().numpy(), 64, 64, 3.0, f0, -1).real
    got = x.grad.cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-6


def tv(x, eps=1e-8):
    dx = x[1:, :-1] - x[:-1, :-1]
    dy = x[:-1, 1:] - x[:-1, :-1]
    return (dx * dx + dy * dy + eps).sqrt().sum()


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_finite_differences_imaging_loss(dtype):
    torch = _torch()
    from curig_amd.autograd import dirty2ms

    say(f"[autograd] central finite differences {dtype}")
    geo, uvw, rng, f0 = _case(torch, nrow=20000, n=128)
    rd = torch.float64 if dtype == "fp64" else torch.float32
    cd = torch.complex128 if dtype == "fp64" else torch.complex64
    data = torch.tensor(rng.standard_normal(geo.nrow) + 1j * rng.standard_normal(geo.nrow), dtype=cd, device="cuda")
    sw = torch.tensor(np.sqrt(rng.uniform(0.5, 1, geo.nrow)), dtype=torch.float64, device="cuda")
    lam = 1e-3

    def loss(x):
        r = (dirty2ms(x, geo) - data).to(torch.complex128) * sw
        return (r.real ** 2 + r.imag ** 2).sum() + lam * tv(x.to(torch.float64))

    x0 = torch.tensor(rng.standard_normal((geo.nx, geo.ny)), dtype=rd, device="cuda")
    x = x0.clone().requires_grad_(True)
    loss(x).backward()
    grad = x.grad.to(torch.float64).cpu().numpy().reshape(-1)
    h = 1e-3
    pix = rng.choice(geo.nx * geo.ny, 100, replace=False)
    fd = np.empty(100)
    with torch.no_grad():
        for q, p in enumerate(pix):
            e = torch.zeros(geo.nx * geo.ny, dtype=rd, device="cuda")
            e[p] = h
            e = e.reshape(geo.nx, geo.ny)
            fd[q] = (loss(x0 + e).item() - loss(x0 - e).item()) / (2 * h)
    err = np.linalg.norm(fd - grad[pix]) / np.linalg.norm(grad[pix])
    say(f"[autograd] FD relative error {err:.2e}")
    # FP32: the differences of two FP32-rounded losses, h = 1e-3, measure
    # about 1.2e-3 here; the bar leaves a factor of four
    assert err < (1e-6 if dtype == "fp64" else 5e-3)
