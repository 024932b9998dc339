"""GPU parity: the HIP gridder, called through the C ABI, against the oracle.

Two oracles, two bars:
* the numpy restatement of the reference's algorithm (oracle/nufft_ref.py):
  the same discrete operator, so FP64 results agree to ~1e-10 relative;
* the FP64 direct-sum DFT (oracle/dft.c): agreement to the requested epsilon
  (3 eps + 1e-11 for FP64; FP32-mixed adds 3e-6 of single-precision roundoff).
Adjointness is asserted at rtol 1e-12 in FP64 as the reference does
(python/curagridder/test_wst.py:77)."""
import sys
import zlib

import numpy as np
import pytest

from oracle import nufft_ref as R
from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0


def _lib():
    from curig_amd import _lib as L

    return L.load()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_case(seed, nrow, nx, scale=1.0, f0=1e9, wscale=1.0):
    rng = np.random.default_rng(seed)
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * scale * nx / 2
    uvw[:, 2] *= wscale
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1.0, nrow)
    return uvw, vis, wgt, f0, rng


def gpu_ms2dirty(uvw, vis, wgt, nx, ny, fov, freq, eps, sigma=2.0, sign=-1):
    lib = _lib()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128)
    dirty = np.zeros((nx, ny), dtype=np.complex128)
    if wgt is None:
        code = lib.ms2dirty_1(uvw.shape[0], nx, ny, fov, freq, uvw, vis, dirty, eps, sigma, sign)
    else:
        code = lib.ms2dirty_2(uvw.shape[0], nx, ny, fov, freq, uvw, vis, np.ascontiguousarray(wgt, np.float64),
                              dirty, eps, sigma, sign)
    assert code == 0, lib.curig_last_error()
    return dirty


def gpu_dirty2ms(uvw, img, wgt, nx, ny, fov, freq, eps, sigma=2.0, sign=-1):
    lib = _lib()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.zeros(uvw.shape[0], dtype=np.complex128)
    img = np.ascontiguousarray(img, dtype=np.complex128)
    if wgt is None:
        code = lib.dirty2ms_1(uvw.shape[0], nx, ny, fov, freq, uvw, vis, img, eps, sigma, sign)
    else:
        code = lib.dirty2ms_2(uvw.shape[0], nx, ny, fov, freq, uvw, vis, np.ascontiguousarray(wgt, np.float64),
                              img, eps, sigma, sign)
    assert code == 0, lib.curig_last_error()
    return vis


def gpu_f32(to_image, uvw, data, wgt, nx, ny, fov, freq, eps, sign=-1, sigma=2.0):
    lib = _lib()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    w = None if wgt is None else np.ascontiguousarray(wgt, np.float32)
    wp = None if w is None else w.ctypes.data
    if to_image:
        vis = np.ascontiguousarray(data, dtype=np.complex64)
        dirty = np.zeros((nx, ny), dtype=np.complex64)
        code = lib.ms2dirty_f32(uvw.shape[0], nx, ny, fov, freq, uvw.ctypes.data, vis.ctypes.data, wp,
                                dirty.ctypes.data, eps, sigma, sign)
        assert code == 0, lib.curig_last_error()
        return dirty
    img = np.ascontiguousarray(data, dtype=np.complex64)
    vis = np.zeros(uvw.shape[0], dtype=np.complex64)
    code = lib.dirty2ms_f32(uvw.shape[0], nx, ny, fov, freq, uvw.ctypes.data, vis.ctypes.data, wp, img.ctypes.data,
                            eps, sigma, sign)
    assert code == 0, lib.curig_last_error()
    return vis


CASES = [
    # name, nrow, nx, ny, fov, eps, uv scale (x Nyquist), w scale, sign, use_wgt, sigma
    ("square_1e-6", 300, 64, 64, 10.0, 1e-6, 1.0, 1.0, -1, True, 2.0),
    ("rect_1e-12", 200, 48, 80, 5.0, 1e-12, 1.0, 3.0, -1, False, 2.0),
    ("wrap_2.9pi", 400, 40, 40, 20.0, 1e-4, 2.9, 1.0, 1, True, 2.0),
    ("fold_4.7pi", 250, 36, 24, 2.0, 1e-5, 4.7, 20.0, -1, True, 2.0),
    ("odd_sizes", 180, 31, 17, 8.0, 1e-6, 1.5, 1.0, -1, True, 2.0),
    ("wide_w", 300, 64, 48, 30.0, 1e-7, 1.0, 40.0, -1, True, 2.0),
    ("eps_0.2", 120, 32, 32, 10.0, 2e-1, 1.0, 1.0, -1, False, 2.0),
    ("sigma_1.25", 150, 40, 40, 6.0, 1e-6, 1.0, 2.0, -1, True, 1.25),
    ("big_tiles", 2000, 256, 192, 3.0, 1e-6, 1.3, 5.0, -1, True, 2.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_parity(case):
    name, nrow, nx, ny, fov, eps, scale, wsc, sign, use_wgt, sigma = case
    say(f"[fp64 parity] {name}: nrow={nrow} image={nx}x{ny} fov={fov} eps={eps} sigma={sigma}")
    uvw, vis, wgt, f0, rng = rand_case(zlib.crc32(name.encode()), nrow, nx, scale, wscale=wsc)
    wgt = wgt if use_wgt else None
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    d_gpu = gpu_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps, sigma, sign)
    d_ref = R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps, sigma, sign)
    d_dft = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, sign)
    assert O.l2error(d_ref, d_gpu) < 1e-10, "vs restatement of the reference algorithm"
    assert O.l2error(d_dft, d_gpu) < 3 * eps + 1e-11, "vs direct sum"
    v_gpu = gpu_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps, sigma, sign)
    v_ref = R.dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps, sigma, sign)
    v_dft = O.dft_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, sign)
    assert O.l2error(v_ref, v_gpu) < 1e-10
    assert O.l2error(v_dft, v_gpu) < 3 * eps + 1e-11
    np.testing.assert_allclose(np.vdot(vis, v_gpu), np.vdot(d_gpu, img), rtol=1e-12)


@pytest.mark.parametrize("case", CASES[:6], ids=[c[0] for c in CASES[:6]])
def test_fp32_mixed_vs_dft(case):
    name, nrow, nx, ny, fov, eps, scale, wsc, sign, use_wgt, sigma = case
    eps = max(eps, 1e-6)
    say(f"[fp32 parity] {name}: nrow={nrow} image={nx}x{ny} eps={eps}")
    uvw, vis, wgt, f0, rng = rand_case(zlib.crc32(name.encode()) + 1, nrow, nx, scale, wscale=wsc)
    wgt = wgt.astype(np.float32) if use_wgt else None
    vis = vis.astype(np.complex64)
    img = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
    d = gpu_f32(True, uvw, vis, wgt, nx, ny, fov, f0, eps, sign, sigma)
    d_dft = O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, sign)
    assert O.l2error(d_dft, d) < 3 * eps + 3e-6
    v = gpu_f32(False, uvw, img, wgt, nx, ny, fov, f0, eps, sign, sigma)
    v_dft = O.dft_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, sign)
    assert O.l2error(v_dft, v) < 3 * eps + 3e-6
    lhs = np.vdot(vis.astype(np.complex128), v.astype(np.complex128))
    rhs = np.vdot(d.astype(np.complex128), img.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


# the reference's own test grid (python/curagridder/test_wst.py:36-83) through the
# NIFTY-style API, with the adjointness it asserts and the DFT error it prints
REF_GRID = [(nrow, nx, ny, fov, eps, w)
            for nrow in (2, 27, 100) for (nx, ny) in ((30, 128), (128, 250), (16, 64))
            for fov in (1, 10, 20) for eps in (2e-1, 5e-3, 5e-5, 5e-7, 5e-12) for w in (False, True)]


def test_reference_test_grid_nifty_api():
    import curig_amd

    say(f"[reference grid] {len(REF_GRID)} cases")
    worst = 0.0
    for (nrow, nx, ny, fov, eps, use_wgt) in REF_GRID:
        xpixsize = fov * np.pi / 180 / nx
        ypixsize = fov * np.pi / 180 / ny
        rng = np.random.default_rng(42)
        f0 = 1e9
        freq = f0 + np.arange(1) * f0
        uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C)
        ms = rng.random((nrow, 1)) - 0.5 + 1j * (rng.random((nrow, 1)) - 0.5)
        dirty = rng.random((nx, ny)) - 0.5
        wgt = rng.random((nrow, 1)) if use_wgt else None
        d2 = curig_amd.ms2dirty(uvw, freq, ms, wgt, nx, ny, xpixsize, ypixsize, 0, 0, eps, True)
        ms2 = curig_amd.dirty2ms(uvw, freq, dirty, wgt, xpixsize, ypixsize, 0, 0, eps, True)
        np.testing.assert_allclose(np.vdot(ms, ms2).real, np.vdot(d2, dirty), rtol=1e-12)
        truth = O.dft_ms2dirty(uvw, ms[:, 0], None if wgt is None else wgt[:, 0], nx, ny, fov, f0, -1).real
        err = O.l2error(truth, d2)
        worst = max(worst, err / max(eps, 1e-11))
        assert err < 3 * eps + 1e-10, (nrow, nx, ny, fov, eps, use_wgt, err)
    say(f"[reference grid] worst error / eps = {worst:.3f}")


def test_nrow_zero_and_single_row():
    say("[edge] nrow=0, nrow=1 plane wave, uvw=0")
    nx, ny, fov, f0 = 32, 24, 5.0, 1.2e9
    d = gpu_ms2dirty(np.zeros((0, 3)), np.zeros(0, np.complex128), None, nx, ny, fov, f0, 1e-6)
    assert np.all(d == 0)
    v = gpu_dirty2ms(np.zeros((0, 3)), np.ones((nx, ny)), None, nx, ny, fov, f0, 1e-6)
    assert v.shape == (0,)
    uvw = np.array([[123.0, -45.0, 67.0]])
    vis = np.array([0.5 - 2.0j])
    d = gpu_ms2dirty(uvw, vis, np.array([0.8]), nx, ny, fov, f0, 1e-9)
    i = np.arange(nx)[:, None]
    k = np.arange(ny)[None, :]
    l = (i - nx // 2) * fov * np.pi / 180 / nx
    m = (k - ny // 2) * fov * np.pi / 180 / ny
    n = np.sqrt(1 - l * l - m * m)
    want = 0.8 * vis[0] * np.exp(2j * np.pi * f0 / C * (l * 123.0 - m * 45.0 - 67.0 * (n - 1))) / n
    assert O.l2error(want, d) < 3e-9
    d0 = gpu_ms2dirty(np.zeros((5, 3)), np.arange(5) * (1 + 1j), None, nx, ny, fov, f0, 1e-6)
    assert O.l2error(10 * (1 + 1j) / n, d0) < 3e-6


def test_row_permutation_invariance():
    say("[edge] permutation invariance")
    uvw, vis, wgt, f0, rng = rand_case(5, 3000, 128, 2.0, wscale=10.0)
    nx = ny = 128
    d0 = gpu_ms2dirty(uvw, vis, wgt, nx, ny, 4.0, f0, 1e-7)
    p = rng.permutation(len(vis))
    d1 = gpu_ms2dirty(uvw[p], vis[p], wgt[p], nx, ny, 4.0, f0, 1e-7)
    assert O.l2error(d0, d1) < 1e-12


def test_error_codes_on_device_path():
    say("[edge] error codes")
    lib = _lib()
    nx = ny = 16
    uvw = np.array([[1.0, 2.0, np.nan]])
    vis = np.ones(1, np.complex128)
    dirty = np.zeros((nx, ny), np.complex128)
    assert lib.ms2dirty_1(1, nx, ny, 2.0, 1e9, uvw, vis, dirty, 1e-6, 2.0, -1) == 9
    assert lib.ms2dirty_1(1, nx, ny, 2.0, 1e9, np.zeros((1, 3)), vis, dirty, 1e-6, 2.0, 3) == 5
    assert lib.ms2dirty_1(1, nx, ny, 200.0, 1e9, np.zeros((1, 3)), vis, dirty, 1e-6, 2.0, -1) == 5


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_explicit_gpu_dft_vs_cpu_oracle():
    torch = _torch()
    say("[explicit] GPU direct sum vs CPU oracle")
    lib = _lib()
    uvw, vis, wgt, f0, rng = rand_case(9, 700, 40, 3.0, wscale=30.0)
    nx, ny, fov = 40, 56, 7.0
    pix = np.ascontiguousarray(rng.choice(nx * ny, 300, replace=False).astype(np.int64))
    d_uvw = torch.from_numpy(uvw).cuda()
    d_vis = torch.from_numpy(vis).cuda()
    d_wgt = torch.from_numpy(wgt).cuda()
    d_pix = torch.from_numpy(pix).cuda()
    out = torch.zeros(300, dtype=torch.complex128, device="cuda")
    code = lib.curig_explicit_ms2dirty_dev(700, nx, ny, fov, f0, d_uvw.data_ptr(), d_vis.data_ptr(),
                                           d_wgt.data_ptr(), 0, 300, d_pix.data_ptr(), out.data_ptr(), -1, None)
    assert code == 0
    torch.cuda.synchronize()
    want = O.dft_ms2dirty_pixels(uvw, vis, wgt, nx, ny, fov, f0, pix, -1)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_multichannel_flags_vs_per_channel():
    torch = _torch()
    say("[multi] 3 channels with flags, summed and per-channel images")
    lib = _lib()
    nrow, nchan, nx, ny, fov = 900, 3, 64, 64, 5.0
    rng = np.random.default_rng(21)
    uvw = (rng.random((nrow, 3)) - 0.5) / (1e9 / C) * 40
    freq = np.array([1.0e9, 1.2e9, 1.5e9])
    vis = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    wgt = rng.uniform(0.5, 1, (nrow, nchan))
    flag = (rng.random((nrow, nchan)) < 0.3).astype(np.uint8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d_uvw, d_vis, d_wgt, d_flag = t(uvw), t(vis), t(wgt), t(flag)
    per = torch.zeros((nchan, nx, ny), dtype=torch.complex128, device="cuda")
    tot = torch.zeros((nx, ny), dtype=torch.complex128, device="cuda")
    for out, pc in ((per, 1), (tot, 0)):
        code = lib.curig_ms2dirty_multi_dev(nrow, nchan, nx, ny, fov, freq, d_uvw.data_ptr(), d_vis.data_ptr(),
                                            d_wgt.data_ptr(), d_flag.data_ptr(), out.data_ptr(), pc, 1e-8, 2.0, -1,
                                            0, None)
        assert code == 0
    torch.cuda.synchronize()
    per_h = per.cpu().numpy()
    for c in range(nchan):
        keep = flag[:, c] == 0
        want = O.dft_ms2dirty(uvw[keep], vis[keep, c], wgt[keep, c], nx, ny, fov, freq[c], -1)
        assert O.l2error(want, per_h[c]) < 3e-8
    assert O.l2error(per_h.sum(0), tot.cpu().numpy()) < 1e-13
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    d_img = t(img)
    vout = torch.zeros((nrow, nchan), dtype=torch.complex128, device="cuda")
    code = lib.curig_dirty2ms_multi_dev(nrow, nchan, nx, ny, fov, freq, d_uvw.data_ptr(), vout.data_ptr(),
                                        d_wgt.data_ptr(), d_flag.data_ptr(), d_img.data_ptr(), 0, 1e-8, 2.0, -1, 0,
                                        None)
    assert code == 0
    torch.cuda.synchronize()
    vh = vout.cpu().numpy()
    for c in range(nchan):
        keep = flag[:, c] == 0
        want = O.dft_dirty2ms(uvw[keep], img, wgt[keep, c], nx, ny, fov, freq[c], -1)
        assert O.l2error(want, vh[keep, c]) < 3e-8
        assert np.all(vh[~keep, c] == 0)


def _c3_like_rows(nrow, seed=4):
    import bench

    uvw = bench.make_uvw(256, 20000.0, 768, 4.0, seed=1)
    rng = np.random.default_rng(seed)
    return uvw[rng.choice(len(uvw), nrow, replace=False)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_long_baseline_many_planes(precision):
    """The long-baseline geometry (20 km array, fov 6 deg, 1.4 GHz: ~1800 w planes,
    dozens of plane batches, |x| folded over ~150 periods) on a 256^2 image,
    against the direct sum at 3000 pixels."""
    nrow, nx, ny, fov, f0, eps = 30000, 256, 256, 6.0, 1.4e9, 1e-5
    say(f"[long baseline] {precision}: nrow={nrow} image={nx}x{ny} fov={fov} eps={eps}")
    uvw = _c3_like_rows(nrow)
    rng = np.random.default_rng(11)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1, nrow)
    pix = rng.choice(nx * ny, 3000, replace=False)
    truth = O.dft_ms2dirty_pixels(uvw, vis, wgt, nx, ny, fov, f0, pix, -1)
    if precision == "fp64":
        d = gpu_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps)
        tol = 3 * eps
    else:
        d = gpu_f32(True, uvw, vis, wgt, nx, ny, fov, f0, eps)
        tol = 3 * eps + 3e-6
    from curig_amd import _lib as L

    info = L.plan_info(nrow, nx, ny, fov, f0, eps, 2.0, -1, uvw[:, 2].min(), uvw[:, 2].max(),
                       0 if precision == "fp64" else 1)
    err = O.l2error(truth, d.reshape(-1)[pix])
    say(f"[long baseline] planes={info['nf3']} batches={info['nbatch']} err={err:.3e}")
    assert info["nf3"] > 1000 and info["nbatch"] > 10
    assert err < tol
    # dirty2ms at 300 rows against the direct sum
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    sel = rng.choice(nrow, 300, replace=False)
    if precision == "fp64":
        v = gpu_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps)
    else:
        v = gpu_f32(False, uvw, img, wgt, nx, ny, fov, f0, eps)
    vt = O.dft_dirty2ms(uvw[sel], img, wgt[sel], nx, ny, fov, f0, -1)
    verr = O.l2error(vt, v[sel])
    say(f"[long baseline] dirty2ms err={verr:.3e}")
    assert verr < tol


@pytest.mark.parametrize("planes_per_batch", [1, 3, 5])
def test_small_plane_batches_match_restatement(planes_per_batch):
    """Plane-batch boundaries (spread source batches, interpolation ranges, the
    w-accumulate across batches and the cyclic plane index) against the
    restatement with every batch size."""
    lib = _lib()
    say(f"[plane batches] {planes_per_batch}")
    assert lib.curig_set_tuning(0, 0, planes_per_batch, 0) == 0
    try:
        uvw, vis, wgt, f0, rng = rand_case(77, 400, 48, 1.7, wscale=25.0)
        nx, ny, fov, eps = 48, 40, 12.0, 1e-7
        img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
        d = gpu_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps)
        assert O.l2error(R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps), d) < 1e-10
        v = gpu_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps)
        assert O.l2error(R.dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps), v) < 1e-10
        d32 = gpu_f32(True, uvw, vis, wgt, nx, ny, fov, f0, eps)
        assert O.l2error(R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps), d32) < 3e-6
    finally:
        lib.curig_set_tuning(0, 0, 0, 0)


@pytest.mark.parametrize("nx,ny", [(128, 128), (256, 128), (128, 512), (512, 512)])
def test_lds_fft_path_matches_restatement(nx, ny):
    """Power-of-two grids run the pruned LDS FFT (row pass + four-step column
    passes); FP64 against the restatement, FP32 against the direct sum."""
    from curig_amd import _lib as L

    say(f"[lds fft] {nx}x{ny}")
    uvw, vis, wgt, f0, rng = rand_case(nx * 7 + ny, 600, nx, 1.4, wscale=6.0)
    fov, eps = 3.0, 1e-7
    info = L.plan_info(600, nx, ny, fov, f0, eps, 2.0, -1, uvw[:, 2].min(), uvw[:, 2].max(), 0)
    assert info["fft_lds"] == 1, info
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    d = gpu_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps)
    assert L.stats()["fft_lds"] in (0, 1)
    assert O.l2error(R.ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, eps), d) < 1e-10
    v = gpu_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps)
    assert O.l2error(R.dirty2ms(uvw, img, wgt, nx, ny, fov, f0, eps), v) < 1e-10
    np.testing.assert_allclose(np.vdot(vis, v), np.vdot(d, img), rtol=1e-12)
    d32 = gpu_f32(True, uvw, vis, wgt, nx, ny, fov, f0, 1e-6)
    assert O.l2error(O.dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, f0, -1), d32) < 3e-6 + 3e-6
    v32 = gpu_f32(False, uvw, img, wgt, nx, ny, fov, f0, 1e-6)
    assert O.l2error(O.dft_dirty2ms(uvw, img, wgt, nx, ny, fov, f0, -1), v32) < 6e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lds_fft_path_equals_rocfft_path(precision):
    """The two FFT paths on one 2048^2 problem (many rows, deep w range)."""
    lib = _lib()
    nx = ny = 2048
    say(f"[lds vs rocfft] {precision} {nx}x{ny}")
    uvw, vis, wgt, f0, rng = rand_case(99, 200000, nx, 1.2, wscale=40.0)
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    fn_d = gpu_ms2dirty if precision == "fp64" else (lambda *a: gpu_f32(True, *a))
    fn_v = gpu_dirty2ms if precision == "fp64" else (lambda *a: gpu_f32(False, *a))
    d_lds = fn_d(uvw, vis, wgt, nx, ny, 2.0, f0, 1e-6)
    v_lds = fn_v(uvw, img, wgt, nx, ny, 2.0, f0, 1e-6)
    assert lib.curig_set_fft_path(1) == 0
    try:
        d_roc = fn_d(uvw, vis, wgt, nx, ny, 2.0, f0, 1e-6)
        v_roc = fn_v(uvw, img, wgt, nx, ny, 2.0, f0, 1e-6)
    finally:
        lib.curig_set_fft_path(0)
    tol = 1e-11 if precision == "fp64" else 2e-6
    assert O.l2error(d_roc, d_lds) < tol
    assert O.l2error(v_roc, v_lds) < tol


@pytest.mark.parametrize("precision,n", [("fp64", 4096), ("fp64", 8192), ("fp32", 8192), ("fp32", 8191)])
def test_split_rows_equal_rocfft_path(precision, n):
    """Grids whose rows exceed half a CU's LDS (16384-point FP32, 8192- and
    16384-point FP64) run each row as several shorter workgroups (four by
    default), and FP64 columns of 16384 (more than a CU's LDS) run in registers
    with one-component LDS exchanges; the result equals the rocFFT path and, on
    a few visibilities, the FP64 direct sum."""
    lib = _lib()
    say(f"[split rows] {precision} {n}x{n}")
    rng = np.random.default_rng(7)
    f0, nrow = 1e9, 20000
    # u, v over the whole grid (every row and column of it is touched), small w
    dx = 1.0 * np.pi / 180 / n
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = (rng.random((nrow, 2)) - 0.5) * 0.98 / dx * C / f0
    uvw[:, 2] = rng.uniform(-1, 1, nrow) * 30.0
    wgt = rng.uniform(0.5, 1.0, nrow)
    img = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    fn_d = gpu_ms2dirty if precision == "fp64" else (lambda *a: gpu_f32(True, *a))
    fn_v = gpu_dirty2ms if precision == "fp64" else (lambda *a: gpu_f32(False, *a))
    eps = 1e-6
    info = _lib_plan(lib, len(uvw), n, n, 1.0, f0, eps, uvw, precision)
    assert info["fft_lds"] == 1 and info["rsplit"] == 4
    d_lds = fn_d(uvw, vis, wgt, n, n, 1.0, f0, eps)
    v_lds = fn_v(uvw, img, wgt, n, n, 1.0, f0, eps)
    assert lib.curig_set_fft_path(1) == 0
    try:
        d_roc = fn_d(uvw, vis, wgt, n, n, 1.0, f0, eps)
        v_roc = fn_v(uvw, img, wgt, n, n, 1.0, f0, eps)
    finally:
        lib.curig_set_fft_path(0)
    tol = 1e-11 if precision == "fp64" else 2e-6
    assert O.l2error(d_roc, d_lds) < tol
    assert O.l2error(v_roc, v_lds) < tol
    pix = rng.choice(n * n, 64, replace=False)
    want_d = O.dft_ms2dirty_pixels(uvw, vis, wgt, n, n, 1.0, f0, pix, -1)
    assert O.l2error(want_d, np.asarray(d_lds).reshape(-1)[pix]) < 3 * eps + (3e-6 if precision == "fp32" else 0)
    sel = rng.choice(len(uvw), 64, replace=False)
    want = O.dft_dirty2ms(uvw[sel], img, None if wgt is None else wgt[sel], n, n, 1.0, f0, -1)
    assert O.l2error(want, np.asarray(v_lds)[sel]) < 3 * eps + (3e-6 if precision == "fp32" else 0)


def test_split_rows_dense_and_empty_rows():
    """FP32 rows of 16384 (8192^2 images) split in four parts, with visibilities
    that leave whole bands of rows empty (the zero-output shortcut), put
    thousands of records into a few rows (several per thread of a row-part
    workgroup) and spread the rest over the grid: both directions equal the
    rocFFT path and the FP64 direct sums."""
    lib = _lib()
    torch = _torch()
    n, f0, eps = 8192, 1e9, 1e-6
    say(f"[split rows, dense and empty rows] fp32 {n}x{n}")
    rng = np.random.default_rng(17)
    dx = 1.0 * np.pi / 180 / n
    scale = C / f0 / dx
    spread_rows = (rng.random((15000, 2)) - 0.5) * 0.98
    dense = np.stack([(rng.random(15000) - 0.5) * 0.98, 0.2 + rng.random(15000) * 0.002], axis=1)
    uv = np.concatenate([spread_rows, dense])
    uv = uv[np.abs(uv[:, 0] - 0.1) > 0.05]  # an empty band of grid rows
    nrow = len(uv)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = uv * scale
    uvw[:, 2] = rng.uniform(-1, 1, nrow) * 30.0
    wgt = rng.uniform(0.5, 1.0, nrow)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    img = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    info = _lib_plan(lib, nrow, n, n, 1.0, f0, eps, uvw, "fp32")
    assert info["fft_lds"] == 1 and info["rsplit"] == 4

    def both(fft_path):
        assert lib.curig_set_fft_path(fft_path) == 0
        try:
            return (gpu_f32(True, uvw, vis, wgt, n, n, 1.0, f0, eps),
                    gpu_f32(False, uvw, img, wgt, n, n, 1.0, f0, eps))
        finally:
            lib.curig_set_fft_path(0)

    d_lds, v_lds = both(0)
    d_roc, v_roc = both(1)
    e1 = O.l2error(d_roc, d_lds)
    pix = rng.choice(n * n, 64, replace=False)
    want = O.dft_ms2dirty_pixels(uvw, vis, wgt, n, n, 1.0, f0, pix, -1)
    e2 = O.l2error(want, np.asarray(d_lds).reshape(-1)[pix])
    say(f"  ms2dirty: vs rocFFT path {e1:.2e}, vs direct sum {e2:.2e}")
    # dirty2ms against the GPU FP64 direct sum (pinned to the CPU oracle in
    # test_gpu_configs.py)
    e3 = O.l2error(v_roc, v_lds)
    rows = np.sort(rng.choice(nrow, 200, replace=False))
    d_rows = torch.from_numpy(rows.astype(np.int64)).cuda()
    d_out = torch.zeros(len(rows), dtype=torch.complex128, device="cuda")
    d_uvw = torch.from_numpy(np.ascontiguousarray(uvw)).cuda()
    d_img = torch.from_numpy(img).cuda()
    d_wgt = torch.from_numpy(wgt.astype(np.float32)).cuda()
    assert lib.curig_explicit_dirty2ms_dev(nrow, n, n, 1.0, f0, d_uvw.data_ptr(), d_img.data_ptr(), d_wgt.data_ptr(),
                                           1, len(rows), d_rows.data_ptr(), d_out.data_ptr(), -1, None) == 0
    torch.cuda.synchronize()
    e4 = O.l2error(d_out.cpu().numpy(), v_lds[rows])
    del d_img
    say(f"  dirty2ms: vs rocFFT path {e3:.2e}, vs direct sum {e4:.2e}")
    assert e1 < 2e-6 and e2 < 3 * eps + 3e-6
    assert e3 < 2e-6 and e4 < 3 * eps + 3e-6


@pytest.mark.parametrize("n", [2048, 8192])
def test_row_spread_block_scaling_over_magnitudes(n):
    """The FP32 row spread (4096-point rows whole at 2048^2, row parts of 4096
    at 8192^2) accumulates each 16-cell block as int32 scaled by a power of
    two chosen from a bound on the block's records: the image must scale with
    the visibilities over the whole FP32 range (1e-36 .. 1e30, where a fixed
    scale would lose or overflow bits), match the rocFFT path (FP32 float
    atomics) on data whose rows mix magnitudes 1e-6 and 1e6, and turn a NaN
    visibility into NaNs."""
    lib = _lib()
    f0, eps = 1e9, 1e-6
    say(f"[row spread, block scaling] fp32 {n}x{n}")
    rng = np.random.default_rng(23)
    dx = 1.0 * np.pi / 180 / n
    nrow = 20000
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = (rng.random((nrow, 2)) - 0.5) * 0.98 * (C / f0 / dx)
    uvw[:, 2] = rng.uniform(-1, 1, nrow) * 30.0
    wgt = np.ones(nrow)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    info = _lib_plan(lib, nrow, n, n, 1.0, f0, eps, uvw, "fp32")
    assert info["fft_lds"] == 1 and info["rsplit"] == (4 if n == 8192 else 0)
    d1 = np.asarray(gpu_f32(True, uvw, vis, wgt, n, n, 1.0, f0, eps))
    for s in (1e-36, 1e30):
        ds = np.asarray(gpu_f32(True, uvw, vis * s, wgt, n, n, 1.0, f0, eps))
        e = O.l2error(d1.astype(np.complex128) * s, ds.astype(np.complex128))
        say(f"  scale {s:g}: {e:.2e}")
        assert np.all(np.isfinite(ds)) and e < 1e-6
    mixed = vis * np.where(rng.random(nrow) < 0.5, 1e-6, 1e6)
    assert lib.curig_set_fft_path(1) == 0
    try:
        d_roc = np.asarray(gpu_f32(True, uvw, mixed, wgt, n, n, 1.0, f0, eps))
    finally:
        lib.curig_set_fft_path(0)
    d_lds = np.asarray(gpu_f32(True, uvw, mixed, wgt, n, n, 1.0, f0, eps))
    e = O.l2error(d_roc, d_lds)
    say(f"  mixed 1e-6 / 1e6 vs rocFFT path: {e:.2e}")
    assert e < 2e-6
    bad = vis.copy()
    bad[nrow // 2] = np.nan
    assert np.isnan(np.asarray(gpu_f32(True, uvw, bad, wgt, n, n, 1.0, f0, eps))).any()


@pytest.mark.parametrize("precision,n", [("fp64", 2048), ("fp32", 2048), ("fp64", 8192), ("fp32", 8192)])
def test_wide_kernels_on_plane_loop_rows(precision, n):
    """Kernel widths above 8 (epsilon 1e-10: 16 taps per axis stored) on the
    plane-loop row kernels, whole rows of 4096 (2048^2) and row parts of 4096
    (8192^2), both directions, equal the rocFFT path."""
    lib = _lib()
    say(f"[wide kernels, plane-loop rows] {precision} {n}x{n}")
    rng = np.random.default_rng(31)
    f0, nrow, eps = 1e9, 3000, 1e-10
    dx = 1.0 * np.pi / 180 / n
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = (rng.random((nrow, 2)) - 0.5) * 0.98 / dx * C / f0
    uvw[:, 2] = rng.uniform(-1, 1, nrow) * 10.0
    wgt = rng.uniform(0.5, 1.0, nrow)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    img = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    info = _lib_plan(lib, nrow, n, n, 1.0, f0, eps, uvw, precision)
    assert info["fft_lds"] == 1 and info["kw"] > 8
    assert info["rsplit"] == (4 if n == 8192 else 0)
    if precision == "fp64":
        fd = lambda: gpu_ms2dirty(uvw, vis, wgt, n, n, 1.0, f0, eps)
        fv = lambda: gpu_dirty2ms(uvw, img, wgt, n, n, 1.0, f0, eps)
    else:
        fd = lambda: gpu_f32(True, uvw, vis, wgt, n, n, 1.0, f0, eps)
        fv = lambda: gpu_f32(False, uvw, img, wgt, n, n, 1.0, f0, eps)
    d_lds, v_lds = fd(), fv()
    assert lib.curig_set_fft_path(1) == 0
    try:
        d_roc, v_roc = fd(), fv()
    finally:
        lib.curig_set_fft_path(0)
    e1, e2 = O.l2error(d_roc, d_lds), O.l2error(v_roc, v_lds)
    say(f"  kw {info['kw']}: ms2dirty {e1:.2e}, dirty2ms {e2:.2e}")
    tol = 1e-11 if precision == "fp64" else 2e-6
    assert e1 < tol and e2 < tol


def _lib_plan(lib, nrow, nx, ny, fov, f0, eps, uvw, precision):
    from curig_amd import _lib as L

    return L.plan_info(nrow, nx, ny, fov, f0, eps, 2.0, -1, float(uvw[:, 2].min()), float(uvw[:, 2].max()),
                       0 if precision == "fp64" else 1)


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_mode_matches_channel_loop(precision):
    """Batched form: the hipGraph-captured multi-channel sequence equals the plain
    per-channel loop, on repeated replays with new data between them."""
    torch = _torch()
    lib = _lib()
    say(f"[graph] precision {precision}")
    nrow, nchan, nx, ny, fov = 4000, 5, 128, 128, 3.0
    rng = np.random.default_rng(31 + precision)
    uvw = (rng.random((nrow, 3)) - 0.5) / (1e9 / C) * 60
    uvw[:, 2] *= 10
    freq = np.linspace(1.0e9, 1.5e9, nchan)
    cd = torch.complex128 if precision == 0 else torch.complex64
    rd = torch.float64 if precision == 0 else torch.float32
    d_uvw = torch.from_numpy(uvw).cuda()
    d_wgt = torch.tensor(rng.uniform(0.5, 1, (nrow, nchan)), dtype=rd, device="cuda")
    d_flag = torch.tensor((rng.random((nrow, nchan)) < rng.uniform(0, 0.4, nchan)).astype(np.uint8), device="cuda")
    d_vis = torch.zeros((nrow, nchan), dtype=cd, device="cuda")
    img_g = torch.zeros((nchan, nx, ny), dtype=cd, device="cuda")
    img_l = torch.zeros_like(img_g)
    v_g = torch.zeros((nrow, nchan), dtype=cd, device="cuda")
    v_l = torch.zeros_like(v_g)

    def call(graph, img, vout):
        assert lib.curig_set_graph_mode(graph) == 0
        assert lib.curig_ms2dirty_multi_dev(nrow, nchan, nx, ny, fov, freq, d_uvw.data_ptr(), d_vis.data_ptr(),
                                            d_wgt.data_ptr(), d_flag.data_ptr(), img.data_ptr(), 1, 1e-6, 2.0, -1,
                                            precision, None) == 0
        assert lib.curig_dirty2ms_multi_dev(nrow, nchan, nx, ny, fov, freq, d_uvw.data_ptr(), vout.data_ptr(),
                                            d_wgt.data_ptr(), d_flag.data_ptr(), img.data_ptr(), 1, 1e-6, 2.0, -1,
                                            precision, None) == 0

    try:
        for rep in range(3):
            d_vis.copy_(torch.tensor(rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan)),
                                     dtype=cd))
            call(1, img_g, v_g)
            torch.cuda.synchronize()
            call(0, img_l, v_l)
            torch.cuda.synchronize()
            tol = 1e-12 if precision == 0 else 1e-6
            assert O.l2error(img_l.cpu().numpy(), img_g.cpu().numpy()) < tol, rep
            assert O.l2error(v_l.cpu().numpy(), v_g.cpu().numpy()) < tol, rep
        # and the channels are right: one channel against the direct sum
        c = 2
        keep = d_flag[:, c].cpu().numpy() == 0
        want = O.dft_ms2dirty(uvw[keep], d_vis[:, c].cpu().numpy()[keep].astype(np.complex128),
                              d_wgt[:, c].cpu().numpy()[keep].astype(np.float64), nx, ny, fov, freq[c], -1)
        assert O.l2error(want, img_g[c].cpu().numpy()) < (3e-6 if precision == 0 else 6e-6)
    finally:
        lib.curig_set_graph_mode(1)
