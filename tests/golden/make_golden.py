"""Writes the golden fixtures in tests/golden/ (inputs + FP64 direct-sum outputs).

The reference commits no fixtures (its tests draw random data at run time,
python/curagridder/test_wst.py:55), so these are generated here with an
independent numpy direct sum that follows the reference's own explicit
gridder/degridder conventions (python/curagridder/explicit_test.py:13-54:
pixel grid -ss/2 + arange(ss), nm1 = -(l^2+m^2)/(sqrt(1-l^2-m^2)+1), sign on w),
with the weight applied in both directions (src/astro/ra_exec.cu:98-105,
src/utils.cu:185-192).  Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

C = 299792458.0
HERE = os.path.dirname(os.path.abspath(__file__))


def numpy_dft(uvw, vis, wgt, image, nx, ny, fov, freq, sign):
    dx = fov * np.pi / 180 / nx
    dy = fov * np.pi / 180 / ny
    x, y = np.meshgrid((np.arange(nx) - nx // 2) * dx, (np.arange(ny) - ny // 2) * dy, indexing="ij")
    eps = x ** 2 + y ** 2
    nm1 = -eps / (np.sqrt(1.0 - eps) + 1.0)
    n = nm1 + 1
    w = np.ones(len(vis)) if wgt is None else wgt
    dirty = np.zeros((nx, ny), dtype=np.complex128)
    vis_out = np.zeros(len(vis), dtype=np.complex128)
    for r in range(uvw.shape[0]):
        phase = freq / C * (x * uvw[r, 0] + y * uvw[r, 1] + sign * uvw[r, 2] * nm1)
        e = np.exp(2j * np.pi * phase)
        dirty += w[r] * vis[r] * e
        vis_out[r] = w[r] * np.sum(image / n * np.conj(e))
    return dirty / n, vis_out


def make(name, seed, nrow, nx, ny, fov, freq, sign, use_wgt, scale):
    rng = np.random.default_rng(seed)
    uvw = (rng.random((nrow, 3)) - 0.5) / (freq / C) * scale * nx / 2
    uvw[:, 2] *= 4.0
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1.0, nrow) if use_wgt else None
    image = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    dirty, vis_out = numpy_dft(uvw, vis, wgt, image, nx, ny, fov, freq, sign)
    np.savez_compressed(os.path.join(HERE, name), uvw=uvw, vis=vis,
                        wgt=np.zeros(0) if wgt is None else wgt, image=image,
                        dirty=dirty, vis_out=vis_out, nx=nx, ny=ny, fov=fov,
                        freq=freq, sign=sign)


if __name__ == "__main__":
    make("small_nifty_w.npz", 42, 64, 32, 24, 8.0, 1.4e9, -1, True, 1.7)
    make("small_curig_sign.npz", 7, 48, 20, 30, 15.0, 1.0e9, 1, False, 0.9)
    make("odd_wide.npz", 3, 40, 17, 33, 30.0, 1.2e9, -1, True, 3.5)
