"""GPU: cached hipGraph replays that change only host-derived counts, and calls
on two streams sharing the library's per-device workspace.

* A multi-channel call is captured once and replayed while "the same call"
  repeats.  Changing the flag array alone (same buffers, same w extremes, so
  the same plan geometry) changes each channel's unflagged-row count, which
  the captured launches depend on (the scan sentinel, the accumulator fill and
  the finalize count): the replay must follow it.
* A call on a second stream must not overwrite buffers that the first
  stream's call is still reading."""
import sys

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lib():
    from curig_amd import _lib as L

    return L.load()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_follows_changed_flags(precision):
    torch = _torch()
    lib = _lib()
    say(f"[graph flags] precision {precision}")
    nrow, nchan, nx, ny, fov = 3000, 3, 128, 128, 3.0
    rng = np.random.default_rng(77 + precision)
    uvw = (rng.random((nrow, 3)) - 0.5) / (1e9 / C) * 60
    uvw[:, 2] *= 10
    freq = np.linspace(1.0e9, 1.4e9, nchan)
    cd = torch.complex128 if precision == 0 else torch.complex64
    rd = torch.float64 if precision == 0 else torch.float32
    # the rows holding the w extremes stay unflagged: every replay has the same plan
    keep = [int(np.argmin(uvw[:, 2])), int(np.argmax(uvw[:, 2]))]
    d_uvw = torch.from_numpy(uvw).cuda()
    wgt = rng.uniform(0.5, 1, (nrow, nchan))
    vis = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    d_wgt = torch.tensor(wgt, dtype=rd, device="cuda")
    d_vis = torch.tensor(vis, dtype=cd, device="cuda")
    d_flag = torch.zeros((nrow, nchan), dtype=torch.uint8, device="cuda")
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

    tol = 1e-12 if precision == 0 else 1e-6
    bar = 3e-6 if precision == 0 else 6e-6
    try:
        for rep, frac in enumerate([0.05, 0.45, 0.2, 0.0]):
            flags = (rng.random((nrow, nchan)) < frac).astype(np.uint8)
            flags[keep, :] = 0
            d_flag.copy_(torch.from_numpy(flags))
            say(f"  replay {rep}: flagged fraction {frac}, unflagged per channel {(flags == 0).sum(axis=0)}")
            call(1, img_g, v_g)
            torch.cuda.synchronize()
            call(0, img_l, v_l)
            torch.cuda.synchronize()
            assert O.l2error(img_l.cpu().numpy(), img_g.cpu().numpy()) < tol, rep
            assert O.l2error(v_l.cpu().numpy(), v_g.cpu().numpy()) < tol, rep
            c = rep % nchan
            ok = flags[:, c] == 0
            want = O.dft_ms2dirty(uvw[ok], vis[ok, c], wgt[ok, c], nx, ny, fov, freq[c], -1)
            assert O.l2error(want, img_g[c].cpu().numpy()) < bar, rep
            img_c = img_g[c].cpu().numpy().astype(np.complex128)
            sel = np.flatnonzero(ok)[:200]
            want_v = O.dft_dirty2ms(uvw[sel], img_c, wgt[sel, c], nx, ny, fov, freq[c], -1)
            assert O.l2error(want_v, v_g[sel, c].cpu().numpy()) < bar, rep
            assert np.all(v_g[:, c].cpu().numpy()[~ok] == 0), "flagged samples are written as 0"
    finally:
        lib.curig_set_graph_mode(1)


@pytest.mark.parametrize("precision", [0, 1])
def test_calls_on_two_streams(precision):
    """A long call queued on stream A, then a call on stream B without any
    synchronisation between them: both results equal the same calls run one
    after the other on one stream."""
    torch = _torch()
    lib = _lib()
    say(f"[two streams] precision {precision}")
    cd = torch.complex128 if precision == 0 else torch.complex64
    rng = np.random.default_rng(5 + precision)

    def problem(nrow, n, scale):
        uvw = (rng.random((nrow, 3)) - 0.5) / (1e9 / C) * scale
        uvw[:, 2] *= 5
        vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
        return torch.from_numpy(uvw).cuda(), torch.tensor(vis, dtype=cd, device="cuda"), n

    big = problem(400_000, 2048, 900)
    small = problem(5_000, 256, 100)

    def run(p, img, stream):
        d_uvw, d_vis, n = p
        assert lib.curig_ms2dirty_dev(d_uvw.shape[0], n, n, 2.0, 1e9, d_uvw.data_ptr(), d_vis.data_ptr(), None,
                                      img.data_ptr(), 1e-6, 2.0, -1, precision, stream) == 0, lib.curig_last_error()

    ref_big = torch.zeros((2048, 2048), dtype=cd, device="cuda")
    ref_small = torch.zeros((256, 256), dtype=cd, device="cuda")
    run(big, ref_big, None)
    run(small, ref_small, None)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(2):
        out_big = torch.zeros_like(ref_big)
        out_small = torch.zeros_like(ref_small)
        torch.cuda.synchronize()
        run(big, out_big, sa.cuda_stream)
        run(small, out_small, sb.cuda_stream)
        torch.cuda.synchronize()
        tol = 1e-12 if precision == 0 else 1e-6
        assert O.l2error(ref_big.cpu().numpy(), out_big.cpu().numpy()) < tol, rep
        assert O.l2error(ref_small.cpu().numpy(), out_small.cpu().numpy()) < tol, rep
