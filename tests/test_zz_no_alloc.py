"""GPU: the entry point that allocates nothing (curig_ms2dirty_dev_noalloc)
against the allocating one it mirrors (curig_ms2dirty_dev), through the C ABI,
on two calls in a row with the second smaller, in both precisions; and its
clear rejections (a working set beyond the fixed scratch, a grid off the LDS
FFT path)."""
import ctypes
import sys

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from curig_amd import _lib

    lib = _lib.load()
    # the argument types, set here as well: a pointer passed without them would
    # be truncated to 32 bits
    f = lib.curig_ms2dirty_dev_noalloc
    v, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    f.argtypes = [i, i, i, d, d, v, v, v, v, d, d, i, i, v]
    f.restype = i
    return torch, lib


def _case(torch, rng, nrow, n, precision):
    f0 = 1.3e9
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * n * 0.9
    uvw[:, 2] *= 6
    cd = torch.complex128 if precision == 0 else torch.complex64
    rd = torch.float64 if precision == 0 else torch.float32
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1.0, nrow)
    return (uvw, vis, wgt, f0, torch.from_numpy(uvw).cuda(), torch.tensor(vis, dtype=cd, device="cuda"),
            torch.tensor(wgt, dtype=rd, device="cuda"))


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry(precision):
    torch, lib = _setup()
    rng = np.random.default_rng(40 + precision)
    cd = torch.complex128 if precision == 0 else torch.complex64
    fov, eps = 3.0, 1e-6 if precision == 0 else 1e-5
    for call, (nrow, n) in enumerate([(200_000, 1024), (30_000, 256)]):
        say(f"[no-alloc] precision {precision} call {call}: nrow {nrow}, {n}^2")
        uvw, vis, wgt, f0, d_uvw, d_vis, d_wgt = _case(torch, rng, nrow, n, precision)
        ref = torch.zeros((n, n), dtype=cd, device="cuda")
        out = torch.full((n, n), float("nan"), dtype=cd, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        assert lib.curig_ms2dirty_dev(nrow, n, n, fov, f0, d_uvw.data_ptr(), d_vis.data_ptr(), d_wgt.data_ptr(),
                                      ref.data_ptr(), eps, 2.0, -1, precision, st) == 0, lib.curig_last_error()
        assert lib.curig_ms2dirty_dev_noalloc(nrow, n, n, fov, f0, d_uvw.data_ptr(), d_vis.data_ptr(),
                                              d_wgt.data_ptr(), out.data_ptr(), eps, 2.0, -1, precision, st) == 0, \
            lib.curig_last_error()
        torch.cuda.synchronize()
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        pix = rng.choice(n * n, 64, replace=False)
        want = O.dft_ms2dirty_pixels(uvw, vis, wgt, n, n, fov, f0, pix, -1)
        e = float(np.linalg.norm(a.reshape(-1)[pix] - want) / np.linalg.norm(want))
        say(f"  vs allocating entry {rel:.1e}, vs direct sum {e:.1e}")
        assert rel < (1e-12 if precision == 0 else 1e-6)
        assert e < 3 * eps + (1e-11 if precision == 0 else 3e-6)


def test_noalloc_rejects_what_it_cannot_hold():
    torch, lib = _setup()
    rng = np.random.default_rng(9)
    uvw, vis, wgt, f0, d_uvw, d_vis, d_wgt = _case(torch, rng, 1000, 8192, 1)
    big = torch.zeros((8192, 8192), dtype=torch.complex64, device="cuda")
    code = lib.curig_ms2dirty_dev_noalloc(1000, 8192, 8192, 3.0, f0, d_uvw.data_ptr(), d_vis.data_ptr(),
                                          d_wgt.data_ptr(), big.data_ptr(), 1e-5, 2.0, -1, 1, None)
    msg = lib.curig_last_error().decode()
    say(f"[no-alloc] 8192^2: code {code}: {msg}")
    assert code == 10 and "scratch" in msg
    odd = torch.zeros((300, 300), dtype=torch.complex64, device="cuda")
    code = lib.curig_ms2dirty_dev_noalloc(1000, 300, 300, 3.0, f0, d_uvw.data_ptr(), d_vis.data_ptr(),
                                          d_wgt.data_ptr(), odd.data_ptr(), 1e-5, 2.0, -1, 1, None)
    msg = lib.curig_last_error().decode()
    say(f"[no-alloc] 300^2 (grid 600, not a power of two): code {code}: {msg}")
    assert code == 10 and "LDS" in msg
    torch.cuda.synchronize()
