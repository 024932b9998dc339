"""ctypes access to the oracle's C direct-sum DFT (oracle/dft.c) for the tests.

The oracle is the checker; nothing here is used by the product path."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C")
_cp = np.ctypeslib.ndpointer(np.complex128, flags="C")
_ip = np.ctypeslib.ndpointer(np.int64, flags="C")


def _opt(arr):
    return None if arr is None else np.ascontiguousarray(arr, dtype=np.float64)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    lib = ctypes.cdll.LoadLibrary(LIB)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    wp = ctypes.c_void_p
    lib.oracle_threads.argtypes = [c_int]
    lib.oracle_threads.restype = c_int
    lib.oracle_dft_ms2dirty.argtypes = [c_int, c_int, c_int, c_double, c_double, _dp, _cp, wp, _cp, c_int, c_int]
    lib.oracle_dft_ms2dirty.restype = None
    lib.oracle_dft_dirty2ms.argtypes = [c_int, c_int, c_int, c_double, c_double, _dp, _cp, wp, _cp, c_int, c_int]
    lib.oracle_dft_dirty2ms.restype = None
    lib.oracle_dft_ms2dirty_pixels.argtypes = [c_int, c_int, c_int, c_double, c_double, _dp, _cp, wp, c_int, _ip,
                                               _cp, c_int, c_int]
    lib.oracle_dft_ms2dirty_pixels.restype = None
    lib.oracle_dft_nifty.argtypes = [c_int, c_int, c_int, c_int, c_int, c_double, c_double, _dp, _dp, _cp, wp, wp,
                                     c_int, _cp, c_int]
    lib.oracle_dft_nifty.restype = None
    _lib = lib
    return lib


def _wptr(w):
    return None if w is None else w.ctypes.data_as(ctypes.c_void_p)


def dft_ms2dirty(uvw, vis, wgt, nx, ny, fov, freq, sign=-1, nthreads=0):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128).ravel()
    w = _opt(wgt)
    w = None if w is None else w.ravel()
    out = np.zeros((nx, ny), dtype=np.complex128)
    lib.oracle_dft_ms2dirty(uvw.shape[0], nx, ny, fov, freq, uvw, vis, _wptr(w), out, sign, nthreads)
    return out


def dft_ms2dirty_pixels(uvw, vis, wgt, nx, ny, fov, freq, pix, sign=-1, nthreads=0):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128).ravel()
    w = _opt(wgt)
    w = None if w is None else w.ravel()
    pix = np.ascontiguousarray(pix, dtype=np.int64)
    out = np.zeros(pix.shape[0], dtype=np.complex128)
    lib.oracle_dft_ms2dirty_pixels(uvw.shape[0], nx, ny, fov, freq, uvw, vis, _wptr(w), pix.shape[0], pix, out,
                                   sign, nthreads)
    return out


def dft_dirty2ms(uvw, dirty, wgt, nx, ny, fov, freq, sign=-1, nthreads=0):
    lib = load()
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    dirty = np.ascontiguousarray(dirty, dtype=np.complex128).reshape(nx, ny)
    w = _opt(wgt)
    w = None if w is None else w.ravel()
    out = np.zeros(uvw.shape[0], dtype=np.complex128)
    lib.oracle_dft_dirty2ms(uvw.shape[0], nx, ny, fov, freq, uvw, out, _wptr(w), dirty, sign, nthreads)
    return out


def _nifty_args(uvw, freq, ms, wgt, mask):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    freq = np.ascontiguousarray(np.atleast_1d(freq), dtype=np.float64)
    nrow, nchan = uvw.shape[0], freq.shape[0]

    def per_sample(a, dtype):
        a = np.asarray(a)
        if a.ndim == 1 and a.shape[0] == nrow:
            a = a.reshape(nrow, 1)
        return np.ascontiguousarray(np.broadcast_to(a, (nrow, nchan)), dtype=dtype)

    ms = per_sample(ms, np.complex128)
    w = None if wgt is None else per_sample(wgt, np.float64)
    m = None if mask is None else per_sample(mask, np.uint8)
    return uvw, freq, ms, w, m


def dft_nifty_ms2dirty(uvw, freq, ms, wgt, mask, nx, ny, dx, dy, apply_w=True, nthreads=0):
    """NIFTY-form direct sum: complex image [nx][ny] summed over the channels."""
    uvw, freq, ms, w, m = _nifty_args(uvw, freq, ms, wgt, mask)
    out = np.zeros((nx, ny), dtype=np.complex128)
    load().oracle_dft_nifty(1, uvw.shape[0], freq.shape[0], nx, ny, dx, dy, freq, uvw, ms, _wptr(w), _wptr(m),
                            int(bool(apply_w)), out, nthreads)
    return out


def dft_nifty_dirty2ms(uvw, freq, dirty, wgt, mask, dx, dy, apply_w=True, nthreads=0):
    """Adjoint of dft_nifty_ms2dirty: ms [nrow][nchan]."""
    nrow, nchan = np.shape(uvw)[0], np.atleast_1d(freq).shape[0]
    uvw, freq, ms, w, m = _nifty_args(uvw, freq, np.zeros((nrow, nchan), complex), wgt, mask)
    dirty = np.ascontiguousarray(dirty, dtype=np.complex128)
    nx, ny = dirty.shape
    load().oracle_dft_nifty(0, nrow, nchan, nx, ny, dx, dy, freq, uvw, ms, _wptr(w), _wptr(m), int(bool(apply_w)),
                            dirty, nthreads)
    return ms


def l2error(a, b):
    """||a - b|| / ||a|| with a = truth (python/curagridder/test_wst.py:11-12)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(a) ** 2)))
