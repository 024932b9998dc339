"""GPU parity of the generic NUFFT (curig_nufft_dev, the work behind the plan
API's cunufft_exec) against the oracle's direct sums: types 1 and 2 in 1, 2
and 3 dimensions, both signs, FP64 and FP32-mixed, on the rocFFT path and the
LDS path (power-of-two grids), with coordinates beyond [-pi, pi).  Bars: 3 eps
+ 1e-11 (FP64), 3 eps + 3e-6 (FP32-mixed); adjointness of types 1 and 2 at
rtol 1e-12 in FP64."""
import sys

import numpy as np
import pytest

from oracle import nufft_ref as R

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(type_, N, coords, data, iflag, eps, precision):
    import torch

    from curig_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda:0")
    dim = len(N)
    Nf = tuple(N) + (1,) * (3 - dim)
    M = len(coords[0])
    cd = torch.complex128 if precision == 0 else torch.complex64
    xs = [torch.tensor(np.asarray(x, dtype=np.float64), device=dev) for x in coords]
    ptr = lambda i: xs[i].data_ptr() if i < dim else None
    if type_ == 1:
        c = torch.tensor(np.asarray(data), dtype=cd, device=dev)
        fk = torch.zeros(N[::-1], dtype=cd, device=dev)
    else:
        fk = torch.tensor(np.asarray(data), dtype=cd, device=dev)
        c = torch.zeros(M, dtype=cd, device=dev)
    code = lib.curig_nufft_dev(type_, dim, Nf[0], Nf[1], Nf[2], M, ptr(0), ptr(1), ptr(2), c.data_ptr(),
                               fk.data_ptr(), iflag, eps, 2.0, precision,
                               torch.cuda.current_stream().cuda_stream)
    assert code == 0, lib.curig_last_error()
    torch.cuda.synchronize()
    out = fk if type_ == 1 else c
    return out.cpu().numpy().astype(np.complex128)


def _case(seed, N, M):
    rng = np.random.default_rng(seed)
    coords = tuple(rng.uniform(-np.pi, np.pi, M) * (1.0 if i else 1.5) for i in range(len(N)))  # u beyond +-pi
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fk = rng.standard_normal(tuple(N[::-1])) + 1j * rng.standard_normal(tuple(N[::-1]))
    return coords, c, fk


def _err(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(a))


# 1-D; 2-D on the rocFFT path (grid 100 x 96) and the LDS path (128 x 128); 3-D
SHAPES = [(100,), (48, 50), (64, 64), (24, 20, 6)]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("iflag", [1, -1])
@pytest.mark.parametrize("type_", [1, 2])
@pytest.mark.parametrize("N", SHAPES)
def test_nufft_matches_direct_sum(N, type_, iflag, precision):
    eps = 1e-6 if precision == 0 else 1e-5
    coords, c, fk = _case(sum(N) + type_, N, 1500)
    say(f"[nufft] type {type_} N {N} iflag {iflag} precision {precision}")
    data = c if type_ == 1 else fk
    got = _run(type_, N, coords, data, iflag, eps, precision)
    want = R.nufft_direct(type_, N, coords, data, iflag)
    err = _err(want, got)
    tol = 3 * eps + 1e-11 + (3e-6 if precision else 0)
    say(f"  err {err:.3e} tol {tol:.1e}")
    assert err < tol


@pytest.mark.parametrize("N", SHAPES)
def test_nufft_adjoint(N):
    coords, c, fk = _case(7 + sum(N), N, 1200)
    say(f"[nufft adjoint] N {N}")
    a = _run(1, N, coords, c, 1, 1e-9, 0)
    b = _run(2, N, coords, fk, -1, 1e-9, 0)
    lhs, rhs = np.vdot(fk, a), np.vdot(b, c)
    say(f"  rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def _run3(M, N, x, s, data, iflag, direction, eps, precision):
    import torch

    from curig_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda:0")
    cd = torch.complex128 if precision == 0 else torch.complex64
    dx = torch.tensor(np.asarray(x, dtype=np.float64), device=dev)
    ds = torch.tensor(np.asarray(s, dtype=np.float64), device=dev)
    if direction == 1:
        c = torch.tensor(np.asarray(data), dtype=cd, device=dev)
        f = torch.zeros(N, dtype=cd, device=dev)
    else:
        f = torch.tensor(np.asarray(data), dtype=cd, device=dev)
        c = torch.zeros(M, dtype=cd, device=dev)
    code = lib.curig_nufft3_dev(M, N, dx.data_ptr(), ds.data_ptr(), c.data_ptr(), f.data_ptr(), iflag, direction,
                                eps, 2.0, precision, torch.cuda.current_stream().cuda_stream)
    assert code == 0, lib.curig_last_error()
    torch.cuda.synchronize()
    return (f if direction == 1 else c).cpu().numpy().astype(np.complex128)


# (x range, s range): a wide product (hundreds of w planes) and the reference
# driver's case (x = 2 + j pi / M, s = -i / N)
T3 = [("wide", 1500, 300), ("reference", 1000, 200)]


def _case3(kind, M, N, seed):
    rng = np.random.default_rng(seed)
    if kind == "wide":
        x, s = rng.uniform(-20, 20, M), rng.uniform(-15, 15, N)
    else:
        x, s = 2.0 + np.arange(M) * np.pi / M, -np.arange(N) / N
    return x, s, rng.standard_normal(M) + 1j * rng.standard_normal(M), rng.standard_normal(N) + 1j * rng.standard_normal(N)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("iflag", [1, -1])
@pytest.mark.parametrize("direction", [1, 0])
@pytest.mark.parametrize("case", T3, ids=[t[0] for t in T3])
def test_nufft3_matches_direct_sum(case, direction, iflag, precision):
    kind, M, N = case
    eps = 1e-6 if precision == 0 else 1e-5
    x, s, c, f = _case3(kind, M, N, 21 + direction)
    say(f"[nufft3] {kind} M {M} N {N} direction {direction} iflag {iflag} precision {precision}")
    data = c if direction == 1 else f
    got = _run3(M, N, x, s, data, iflag, direction, eps, precision)
    want = R.nufft3_direct(x, s, data, iflag, direction)
    err = _err(want, got)
    tol = 3 * eps + 1e-11 + (3e-6 if precision else 0)
    say(f"  err {err:.3e} tol {tol:.1e}")
    assert err < tol


def test_nufft3_adjoint():
    M, N = 1500, 300
    x, s, c, f = _case3("wide", M, N, 31)
    say("[nufft3 adjoint]")
    a = _run3(M, N, x, s, c, 1, 1, 1e-9, 0)
    b = _run3(M, N, x, s, f, -1, 0, 1e-9, 0)
    lhs, rhs = np.vdot(f, a), np.vdot(b, c)
    say(f"  rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
