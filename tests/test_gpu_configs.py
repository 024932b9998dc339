"""GPU parity on the five benchmark workloads at their full sizes (synthetic,
seeded; the generator is bench.py's):

* c1: 32 antennas in a 1 km disk x 64 snapshots (31,744 rows), 256^2, fov 2
  deg, 1.4 GHz, epsilon 1e-6, FP64, weights: both directions against the
  oracle's direct sums over the whole image and every row;
* c2: 128 antennas in a 5 km disk x 256 snapshots (2,080,768 rows) x 4
  channels over 1.0-1.4 GHz, a 50-source sky plus noise, 2048^2, fov 4 deg,
  epsilon 1e-6, FP64 round trip: 2,000 pixels and 400 rows against the GPU
  direct sums, a subset of those against the CPU oracle, and adjointness;
* c3: 256 antennas in a 20 km disk x 768 snapshots (25,067,520 rows) x 2
  channels (1.0, 1.4 GHz), 8192^2, fov 6 deg, FP32-mixed, one round trip at
  epsilon 1e-6: 10,000 pixels and 1,000 rows per channel against the GPU
  direct sums (themselves pinned to the CPU oracle in test_gpu_parity.py);
* c4: 64 channels over 1.0-1.5 GHz, flag fraction U[0, 0.4] per channel,
  4096^2 per channel, fov 3 deg, epsilon 1e-5, FP32-mixed: the hipGraph
  sequence against the per-channel loop, two channels against the direct sums;
* c5: 64 antennas in a 3 km disk x 512 snapshots (1,032,192 rows), 1024^2,
  fov 2 deg, FP32: adjointness <A x, y> = <x, A^H y> to 1e-6, the autograd
  gradient against the direct-sum adjoint on sampled pixels, and central
  finite differences on 100 pixels at h = 1e-3.

Bars: epsilon-scaled as in the other parity tests (3 eps + 1e-11 in FP64,
+ 3e-6 for FP32-mixed), computed against FP64 direct sums."""
import os
import sys

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _lib():
    from curig_amd import _lib as L

    return L.load()


def _err(got, want):
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def _gpu_pixels(torch, lib, d_uvw, d_vis, d_wgt, nrow, n, fov, f, pix, precision):
    """FP64 direct sum of one channel's image at pixels pix (GPU)."""
    d_pix = torch.from_numpy(np.ascontiguousarray(pix, dtype=np.int64)).cuda()
    out = torch.zeros(len(pix), dtype=torch.complex128, device="cuda")
    assert lib.curig_explicit_ms2dirty_dev(nrow, n, n, fov, f, d_uvw.data_ptr(), d_vis.data_ptr(),
                                           None if d_wgt is None else d_wgt.data_ptr(), precision, len(pix),
                                           d_pix.data_ptr(), out.data_ptr(), -1, None) == 0, lib.curig_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _gpu_rows(torch, lib, d_uvw, d_img, d_wgt, nrow, n, fov, f, rows, precision):
    """FP64 direct sum of one channel's dirty2ms of d_img at rows (GPU)."""
    d_rows = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.int64)).cuda()
    out = torch.zeros(len(rows), dtype=torch.complex128, device="cuda")
    assert lib.curig_explicit_dirty2ms_dev(nrow, n, n, fov, f, d_uvw.data_ptr(), d_img.data_ptr(),
                                           None if d_wgt is None else d_wgt.data_ptr(), precision, len(rows),
                                           d_rows.data_ptr(), out.data_ptr(), -1, None) == 0, lib.curig_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_explicit_dirty2ms_gpu_vs_cpu_oracle():
    """The GPU adjoint direct sum the large tests compare against, pinned to the
    CPU oracle (both precisions of the image)."""
    torch = _torch()
    lib = _lib()
    rng = np.random.default_rng(12)
    nrow, n, fov, f0 = 3000, 96, 4.0, 1.2e9
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C) * n * 0.8
    uvw[:, 2] *= 6
    wgt = rng.uniform(0.5, 1, nrow)
    img = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    rows = rng.choice(nrow, 300, replace=False)
    want = O.dft_dirty2ms(uvw[rows], img, wgt[rows], n, n, fov, f0, -1)
    d_uvw = torch.from_numpy(uvw).cuda()
    for precision, cd, rd in ((0, torch.complex128, torch.float64), (1, torch.complex64, torch.float32)):
        d_img = torch.tensor(img, dtype=cd, device="cuda")
        d_wgt = torch.tensor(wgt, dtype=rd, device="cuda")
        got = _gpu_rows(torch, lib, d_uvw, d_img, d_wgt, nrow, n, fov, f0, rows, precision)
        e = _err(got, want)
        say(f"[explicit dirty2ms] precision {precision}: {e:.2e}")
        assert e < (1e-12 if precision == 0 else 3e-7)


def test_c1_full_image_fp64():
    import bench as B

    lib = _lib()
    cfg = B.make_config("c1", "fp64")
    uvw, vis, wgt = cfg["uvw"], np.ascontiguousarray(cfg["vis"][:, 0]), np.ascontiguousarray(cfg["wgt"][:, 0])
    n, fov, f0, eps = cfg["nx"], cfg["fov"], float(cfg["freqs"][0]), cfg["eps"]
    say(f"[c1] {len(uvw)} rows, {n}^2, fov {fov}, eps {eps}")
    assert len(uvw) == 31744
    dirty = np.zeros((n, n), dtype=np.complex128)
    assert lib.ms2dirty_2(len(uvw), n, n, fov, f0, np.ascontiguousarray(uvw), vis, wgt, dirty, eps, 2.0, -1) == 0
    want = O.dft_ms2dirty(uvw, vis, wgt, n, n, fov, f0, -1)
    e1 = _err(dirty, want)
    rng = np.random.default_rng(1)
    img = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out = np.zeros(len(uvw), dtype=np.complex128)
    assert lib.dirty2ms_2(len(uvw), n, n, fov, f0, np.ascontiguousarray(uvw), out, wgt, img, eps, 2.0, -1) == 0
    want_v = O.dft_dirty2ms(uvw, img, wgt, n, n, fov, f0, -1)
    e2 = _err(out, want_v)
    adj = abs(np.vdot(vis, out) - np.vdot(dirty, img)) / abs(np.vdot(vis, out))
    say(f"  ms2dirty {e1:.2e} (whole image), dirty2ms {e2:.2e} (every row), adjointness {adj:.1e}")
    assert e1 < 3 * eps + 1e-11 and e2 < 3 * eps + 1e-11 and adj < 1e-12


def _c2_sky(torch, uvw_d, freqs, rng, nsrc=50, fov=4.0):
    """50 point sources at random positions in the field, fluxes log-uniform in
    [0.01, 1] Jy, plus complex Gaussian noise of sigma 0.1, per channel
    (computed on the GPU): V = sum_s F_s e^{-2 pi i f/c (u l_s + v m_s - w (n_s - 1))}."""
    half = np.deg2rad(fov) / 2
    ls = rng.uniform(-half, half, nsrc)
    ms = rng.uniform(-half, half, nsrc)
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), nsrc))
    nm1 = -(ls ** 2 + ms ** 2) / (np.sqrt(1 - ls ** 2 - ms ** 2) + 1)
    nrow = uvw_d.shape[0]
    vis = torch.zeros((nrow, len(freqs)), dtype=torch.complex128, device="cuda")
    u, v, w = uvw_d[:, 0], uvw_d[:, 1], uvw_d[:, 2]
    for c, f in enumerate(freqs):
        k = 2 * np.pi * f / C
        for s in range(nsrc):
            ph = -k * (u * ls[s] + v * ms[s] - w * nm1[s])
            vis[:, c] += flux[s] * torch.polar(torch.ones_like(ph), ph)
    noise = rng.standard_normal((nrow, len(freqs), 2)) * 0.1
    vis += torch.complex(torch.from_numpy(noise[..., 0]), torch.from_numpy(noise[..., 1])).cuda()
    return vis


def test_c2_round_trip_fp64():
    import bench as B

    torch = _torch()
    lib = _lib()
    nant, rad, nsnap, hh, freqs, n, fov, eps = B.CONFIGS["c2"]
    freqs = np.array(freqs)
    uvw = B.make_uvw(nant, rad, nsnap, hh, seed=1)
    nrow, nchan = len(uvw), len(freqs)
    say(f"[c2] {nrow} rows x {nchan} channels, {n}^2, fov {fov}, eps {eps}, FP64")
    assert nrow == 2080768
    rng = np.random.default_rng(2)
    d_uvw = torch.from_numpy(uvw).cuda()
    d_vis = _c2_sky(torch, d_uvw, freqs, rng)
    d_img = torch.zeros((nchan, n, n), dtype=torch.complex128, device="cuda")
    d_out = torch.zeros_like(d_vis)
    assert lib.curig_ms2dirty_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), d_vis.data_ptr(), None, None,
                                        d_img.data_ptr(), 1, eps, 2.0, -1, 0, None) == 0, lib.curig_last_error()
    assert lib.curig_dirty2ms_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), d_out.data_ptr(), None,
                                        None, d_img.data_ptr(), 1, eps, 2.0, -1, 0, None) == 0, lib.curig_last_error()
    torch.cuda.synchronize()
    pix = rng.choice(n * n, 2000, replace=False)
    rows = np.sort(rng.choice(nrow, 400, replace=False))
    bar = 3 * eps + 1e-11
    for c in (0, nchan - 1):
        vis_c = d_vis[:, c].contiguous()
        img_c = d_img[c].contiguous()
        truth_p = _gpu_pixels(torch, lib, d_uvw, vis_c, None, nrow, n, fov, freqs[c], pix, 0)
        got_p = img_c.reshape(-1).cpu().numpy()[pix]
        truth_r = _gpu_rows(torch, lib, d_uvw, img_c, None, nrow, n, fov, freqs[c], rows, 0)
        got_r = d_out[rows, c].cpu().numpy()
        e1, e2 = _err(got_p, truth_p), _err(got_r, truth_r)
        # the GPU direct sums against the CPU oracle on a subset
        cpu_p = O.dft_ms2dirty_pixels(uvw, vis_c.cpu().numpy(), None, n, n, fov, freqs[c], pix[:40], -1)
        cpu_r = O.dft_dirty2ms(uvw[rows[:10]], img_c.cpu().numpy(), None, n, n, fov, freqs[c], -1)
        e3, e4 = _err(truth_p[:40], cpu_p), _err(truth_r[:10], cpu_r)
        say(f"  channel {c}: ms2dirty {e1:.2e} (2000 px), dirty2ms {e2:.2e} (400 rows); "
            f"GPU direct sums vs CPU oracle {e3:.1e} / {e4:.1e}")
        assert e1 < bar and e2 < bar and e3 < 1e-10 and e4 < 1e-10
    # adjointness over all channels
    lhs = torch.vdot(d_vis.reshape(-1), d_out.reshape(-1)).item()
    rhs = torch.vdot(d_img.reshape(-1), d_img.reshape(-1)).item()
    adj = abs(lhs - rhs) / abs(lhs)
    say(f"  adjointness <V, A A^H V> vs |A^H V|^2: {adj:.1e}")
    assert adj < 1e-12


def test_c3_round_trip_fp32():
    import bench as B

    torch = _torch()
    lib = _lib()
    cfg = B.make_config("c3", "fp32", 1e-6)
    nrow, nchan = cfg["vis"].shape
    n, fov, freqs, eps = cfg["nx"], cfg["fov"], cfg["freqs"], cfg["eps"]
    say(f"[c3] {nrow} rows x {nchan} channels, {n}^2, fov {fov}, eps {eps}, FP32-mixed")
    assert nrow == 25067520
    d_uvw = torch.from_numpy(cfg["uvw"]).cuda()
    d_vis = torch.from_numpy(cfg["vis"]).cuda()
    d_wgt = torch.from_numpy(cfg["wgt"]).cuda()
    del cfg
    d_img = torch.zeros((n, n), dtype=torch.complex64, device="cuda")
    d_out = torch.zeros_like(d_vis)
    assert lib.curig_ms2dirty_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), d_vis.data_ptr(),
                                        d_wgt.data_ptr(), None, d_img.data_ptr(), 0, eps, 2.0, -1, 1, None) == 0
    assert lib.curig_dirty2ms_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), d_out.data_ptr(),
                                        d_wgt.data_ptr(), None, d_img.data_ptr(), 0, eps, 2.0, -1, 1, None) == 0
    torch.cuda.synchronize()
    rng = np.random.default_rng(7)
    pix = rng.choice(n * n, 10000, replace=False)
    rows = np.sort(rng.choice(nrow, 1000, replace=False))
    truth = sum(_gpu_pixels(torch, lib, d_uvw, d_vis[:, c].contiguous(), d_wgt[:, c].contiguous(), nrow, n, fov,
                            freqs[c], pix, 1) for c in range(nchan))
    e1 = _err(d_img.reshape(-1).cpu().numpy()[pix].astype(np.complex128), truth)
    tr = np.stack([_gpu_rows(torch, lib, d_uvw, d_img, d_wgt[:, c].contiguous(), nrow, n, fov, freqs[c], rows, 1)
                   for c in range(nchan)], axis=1)
    e2 = _err(d_out[rows].cpu().numpy().astype(np.complex128), tr)
    say(f"  ms2dirty {e1:.3e} (10k px), dirty2ms {e2:.3e} (1000 rows x {nchan} channels)")
    assert e1 < 1e-5 and e2 < 1e-5


def test_c4_batched_graph_vs_loop():
    import bench as B

    torch = _torch()
    lib = _lib()
    nchan, n, fov, eps = 64, 4096, 3.0, 1e-5
    freqs = np.linspace(1.0e9, 1.5e9, nchan)
    uvw = B.make_uvw(128, 5000.0, 256, 3.0, seed=1)
    nrow = len(uvw)
    rng = np.random.default_rng(11)
    frac = rng.uniform(0.0, 0.4, nchan)
    flags = (rng.random((nrow, nchan)) < frac[None, :]).astype(np.uint8)
    counts = nrow - flags.sum(axis=0)
    say(f"[c4] {nrow} rows x {nchan} channels, unflagged {counts.min()}..{counts.max()} per channel, {n}^2, eps {eps}")
    d_uvw = torch.from_numpy(uvw).cuda()
    d_vis = torch.complex(torch.randn((nrow, nchan), device="cuda"), torch.randn((nrow, nchan), device="cuda"))
    d_wgt = torch.rand((nrow, nchan), device="cuda") * 0.5 + 0.5
    d_flag = torch.from_numpy(flags).cuda()
    img_g = torch.zeros((nchan, n, n), dtype=torch.complex64, device="cuda")
    img_l = torch.zeros_like(img_g)
    v_g = torch.zeros_like(d_vis)
    v_l = torch.zeros_like(d_vis)

    def call(mode, img, vout):
        assert lib.curig_set_graph_mode(mode) == 0
        assert lib.curig_ms2dirty_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), d_vis.data_ptr(),
                                            d_wgt.data_ptr(), d_flag.data_ptr(), img.data_ptr(), 1, eps, 2.0, -1, 1,
                                            None) == 0, lib.curig_last_error()
        assert lib.curig_dirty2ms_multi_dev(nrow, nchan, n, n, fov, freqs, d_uvw.data_ptr(), vout.data_ptr(),
                                            d_wgt.data_ptr(), d_flag.data_ptr(), img.data_ptr(), 1, eps, 2.0, -1, 1,
                                            None) == 0, lib.curig_last_error()
        torch.cuda.synchronize()

    try:
        call(1, img_g, v_g)
        call(0, img_l, v_l)
    finally:
        lib.curig_set_graph_mode(1)
    di = float(torch.linalg.vector_norm(img_g - img_l) / torch.linalg.vector_norm(img_l))
    dv = float(torch.linalg.vector_norm(v_g - v_l) / torch.linalg.vector_norm(v_l))
    say(f"  graph vs loop: images {di:.1e}, visibilities {dv:.1e}")
    assert di < 1e-6 and dv < 1e-6
    pix = rng.choice(n * n, 2000, replace=False)
    bar = 3 * eps + 3e-6
    for c in (int(np.argmin(counts)), int(np.argmax(counts))):
        keep = d_flag[:, c] == 0
        vis_c = torch.where(keep, d_vis[:, c], torch.zeros_like(d_vis[:, c])).contiguous()
        truth = _gpu_pixels(torch, lib, d_uvw, vis_c, d_wgt[:, c].contiguous(), nrow, n, fov, freqs[c], pix, 1)
        e1 = _err(img_g[c].reshape(-1).cpu().numpy()[pix].astype(np.complex128), truth)
        rows = np.flatnonzero(keep.cpu().numpy())[::4000]
        tr = _gpu_rows(torch, lib, d_uvw, img_g[c].contiguous(), d_wgt[:, c].contiguous(), nrow, n, fov, freqs[c],
                       rows, 1)
        e2 = _err(v_g[rows, c].cpu().numpy().astype(np.complex128), tr)
        zero = bool(torch.all(v_g[:, c][~keep] == 0))
        say(f"  channel {c} ({int(keep.sum())} unflagged): ms2dirty {e1:.2e}, dirty2ms {e2:.2e} ({len(rows)} rows); "
            f"flagged samples zero: {zero}")
        assert e1 < bar and e2 < bar and zero


def test_c5_differentiation():
    import bench as B

    torch = _torch()
    from curig_amd.autograd import GridderGeometry, dirty2ms, ms2dirty

    n, fov, f0, eps, lam = 1024, 2.0, 1.4e9, 1e-6, 1e-3
    uvw = B.make_uvw(64, 3000.0, 512, 4.0, seed=5)
    nrow = len(uvw)
    say(f"[c5] {nrow} rows, {n}^2, fov {fov}, eps {eps}")
    assert nrow == 1032192
    rng = np.random.default_rng(21)
    wgt = rng.uniform(0.5, 1.0, nrow)
    geo = GridderGeometry(uvw, f0, n, n, fov, epsilon=eps, wgt=wgt.astype(np.float32))
    x = torch.randn((n, n), dtype=torch.complex64, device="cuda")
    y = torch.randn(nrow, dtype=torch.complex64, device="cuda")
    lhs = torch.vdot(dirty2ms(x, geo).to(torch.complex128), y.to(torch.complex128)).item()
    rhs = torch.vdot(x.reshape(-1).to(torch.complex128), ms2dirty(y, geo).reshape(-1).to(torch.complex128)).item()
    adj = abs(lhs - rhs) / abs(lhs)
    say(f"  adjointness {adj:.2e}")
    assert adj < 1e-6
    # loss ||W^1/2 (A x - V)||^2 + lam TV(x) and its autograd gradient
    V = torch.complex(torch.randn(nrow, device="cuda"), torch.randn(nrow, device="cuda"))
    sw = torch.tensor(np.sqrt(wgt), dtype=torch.float32, device="cuda")

    def tv(z, e=1e-8):
        dx = z[1:, :-1] - z[:-1, :-1]
        dy = z[:-1, 1:] - z[:-1, :-1]
        return (dx * dx + dy * dy + e).sqrt().sum()

    def loss_fn(z):
        r = (dirty2ms(z, geo) - V) * sw
        return (r.real.double() ** 2 + r.imag.double() ** 2).sum() + lam * tv(z.double())

    x0 = torch.randn((n, n), dtype=torch.float32, device="cuda") * 0.01
    xg = x0.clone().requires_grad_(True)
    loss_fn(xg).backward()
    grad = xg.grad.double().reshape(-1).cpu().numpy()
    # data term's gradient at sampled pixels from the FP64 direct sums:
    # d/dx_p sum_j w_j |A x - V|_j^2 = 2 Re sum_j w_j conj(a_jp) (A x - V)_j, a_jp the operator's entries
    with torch.no_grad():
        resid = ((dirty2ms(x0.to(torch.complex64), geo) - V) * sw * sw).to(torch.complex128)
    pix = rng.choice(n * n, 200, replace=False)
    r_np = resid.cpu().numpy()
    # sum_j wgt_j r_j e^{+i phi_jp} / n_p = the ms2dirty direct sum of (r, wgt) at p
    g_data = 2 * O.dft_ms2dirty_pixels(uvw, r_np, wgt, n, n, fov, f0, pix, -1).real
    with torch.no_grad():
        xt = x0.double().clone().requires_grad_(True)
    with torch.enable_grad():
        t = lam * tv(xt)
        t.backward()
    g_tv = xt.grad.reshape(-1).cpu().numpy()[pix]
    e_g = _err(grad[pix], g_data + g_tv)
    say(f"  gradient vs direct-sum adjoint on 200 pixels: {e_g:.2e}")
    assert e_g < 1e-4
    # central finite differences, h = 1e-3, 100 pixels
    h = 1e-3
    fpix = rng.choice(n * n, 100, replace=False)
    fd = np.empty(100)
    with torch.no_grad():
        for q, p in enumerate(fpix):
            e = torch.zeros(n * n, dtype=torch.float32, device="cuda")
            e[p] = h
            e = e.reshape(n, n)
            fd[q] = (loss_fn(x0 + e).item() - loss_fn(x0 - e).item()) / (2 * h)
    e_fd = _err(fd, grad[fpix])
    say(f"  finite differences (100 px, h 1e-3): {e_fd:.2e}")
    assert e_fd < 2e-3
