"""cuRIG-style Python API: vis2dirty / dirty2vis (sign +1).

Mirrors python/curagridder/curagridder.py:69-128 of the reference (vis2dirty
:69-98, dirty2vis :100-128): same names,
arguments and semantics (the caller provides the complex128 output array, which
is filled in place and returned).  Additionally accepts precision="fp32" for the
FP32-mixed form (complex64 data, float32 weights)."""
import numpy as np

from ._lib import check, load

__all__ = ["vis2dirty", "dirty2vis"]


def _host_call(to_image, uvw, freq, ms, wgt, dirty, fov, epsilon, sigma, sign, precision):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    nrow = uvw.shape[0]
    nx, ny = dirty.shape[0], dirty.shape[1]
    f0 = float(np.atleast_1d(freq)[0])
    if precision == "fp64":
        if wgt is None:
            fn = lib.ms2dirty_1 if to_image else lib.dirty2ms_1
            return check(fn(nrow, nx, ny, fov, f0, uvw, ms, dirty, epsilon, sigma, sign))
        fn = lib.ms2dirty_2 if to_image else lib.dirty2ms_2
        w = np.ascontiguousarray(wgt, dtype=np.float64)
        return check(fn(nrow, nx, ny, fov, f0, uvw, ms, w, dirty, epsilon, sigma, sign))
    w = None if wgt is None else np.ascontiguousarray(wgt, dtype=np.float32)
    fn = lib.ms2dirty_f32 if to_image else lib.dirty2ms_f32
    return check(fn(nrow, nx, ny, fov, f0, uvw.ctypes.data, ms.ctypes.data,
                    None if w is None else w.ctypes.data, dirty.ctypes.data, epsilon, sigma, sign))


def _check_out(a, name, dtype):
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous):
        raise TypeError(f"{name} must be a C-contiguous {np.dtype(dtype).name} array")


def vis2dirty(uvw, freq, ms, wgt, dirty, fov, epsilon=1e-6, sigma=1.25, precision="fp64"):
    """Image from visibilities; fills and returns `dirty` ([nx][ny] complex).

    uvw [nrow,3] metres, freq (freq[0] used), ms [nrow] complex, wgt [nrow] or None,
    fov degrees.  sign = +1 (python/curagridder/curagridder.py:87, 121)."""
    cd = np.complex128 if precision == "fp64" else np.complex64
    _check_out(dirty, "dirty", cd)
    ms = np.ascontiguousarray(ms, dtype=cd).reshape(-1)
    _host_call(True, uvw, freq, ms, wgt, dirty, fov, epsilon, sigma, 1, precision)
    return dirty


def dirty2vis(uvw, freq, ms, wgt, dirty, fov, epsilon=1e-6, sigma=1.25, precision="fp64"):
    """Visibilities from an image; fills and returns `ms` ([nrow] complex)."""
    cd = np.complex128 if precision == "fp64" else np.complex64
    _check_out(ms, "ms", cd)
    dirty = np.ascontiguousarray(dirty, dtype=cd)
    _host_call(False, uvw, freq, ms, wgt, dirty, fov, epsilon, sigma, 1, precision)
    return ms
