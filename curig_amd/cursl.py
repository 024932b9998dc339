"""NIFTY-style Python API: ms2dirty / dirty2ms (sign -1, sigma 2).

Mirrors python/curagridder/cursl.py:55-119 of the reference: fov = rad_pix_x *
nxdirty in degrees, sigma = 2, sign = -1, ms2dirty returns the real part of the
dirty image, dirty2ms takes a real image and returns [nrow, 1] complex128.
Extensions over the reference (which grids freq[0] only, cursl.py:80-84):
several channels (ms [nrow, nchan]) are gridded into one image, as NIFTY does;
rad_pix_y, nx, ny and do_wstacking keep the reference's meaning (ignored)."""
import numpy as np

from ._lib import check, load

__all__ = ["ms2dirty", "dirty2ms"]


def ms2dirty(uvw, freq, ms, wgt, nxdirty, nydirty, rad_pix_x, rad_pix_y, nx, ny, epsilon, do_wstacking, *args):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    nrow = uvw.shape[0]
    freq = np.atleast_1d(np.asarray(freq, dtype=np.float64))
    ms = np.asarray(ms, dtype=np.complex128).reshape(nrow, -1)
    sigma = 2.0
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    total = np.zeros((nxdirty, nydirty), dtype=np.complex128)
    for ch in range(ms.shape[1]):
        dirty = np.zeros((nxdirty, nydirty), dtype=np.complex128)
        vis = np.ascontiguousarray(ms[:, ch])
        if wgt is None:
            check(lib.ms2dirty_1(nrow, nxdirty, nydirty, fov, freq[ch], uvw, vis, dirty, epsilon, sigma, sign))
        else:
            w = np.ascontiguousarray(np.asarray(wgt, dtype=np.float64).reshape(nrow, -1)[:, ch])
            check(lib.ms2dirty_2(nrow, nxdirty, nydirty, fov, freq[ch], uvw, vis, w, dirty, epsilon, sigma, sign))
        total += dirty
    return total.real


def dirty2ms(uvw, freq, dirty, wgt, rad_pix_x, rad_pix_y, nx, ny, epsilon, do_wstacking, *args):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    nrow = uvw.shape[0]
    freq = np.atleast_1d(np.asarray(freq, dtype=np.float64))
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sigma = 2.0
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    nchan = len(freq)
    out = np.zeros((nrow, nchan), dtype=np.complex128)
    dirty1 = np.zeros(dirty.shape, dtype=np.complex128)
    dirty1.real = dirty
    for ch in range(nchan):
        ms = np.zeros(nrow, dtype=np.complex128)
        if wgt is None:
            check(lib.dirty2ms_1(nrow, nxdirty, nydirty, fov, freq[ch], uvw, ms, dirty1, epsilon, sigma, sign))
        else:
            w = np.ascontiguousarray(np.asarray(wgt, dtype=np.float64).reshape(nrow, -1)[:, ch])
            check(lib.dirty2ms_2(nrow, nxdirty, nydirty, fov, freq[ch], uvw, ms, w, dirty1, epsilon, sigma, sign))
        out[:, ch] = ms
    return out
