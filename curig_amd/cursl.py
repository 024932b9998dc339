"""NIFTY-style Python API: ms2dirty / dirty2ms (sign -1).

Same call signature and conventions as the reference's
python/curagridder/cursl.py:55-119 (ms2dirty returns the real part of the dirty
image; dirty2ms takes a real image and returns [nrow, nchan] visibilities), with
the arguments that file drops honoured, as the NIFTY gridder defines them:

* several channels: ms / wgt [nrow, nchan] with freq [nchan] (the reference
  grids freq[0] only, cursl.py:80-84); the image sums over the channels;
* rad_pix_y: the pixel size along y (the reference derives it from rad_pix_x);
* nx, ny: size of the oversampled grid (even, larger than the image; 0, 0 chooses
  an upsampling factor of 2); the kernel is chosen for min(nx/nxdirty, ny/nydirty);
* do_wstacking=False: the 2-D transform (w taken as 0, no division by n);
* mask (keyword, [nrow, nchan] uint8): samples with mask == 0 are skipped
  (dirty2ms writes 0 for them).

complex64 ms (ms2dirty) or float32 dirty (dirty2ms) select the FP32-mixed path.
Every call runs the HIP gridder through curig_nifty_ms2dirty / curig_nifty_dirty2ms."""
import numpy as np

from ._lib import check, load

__all__ = ["ms2dirty", "dirty2ms"]


def _channels(uvw, freq, nchan_hint=None):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    if uvw.ndim != 2 or uvw.shape[1] != 3:
        raise ValueError("uvw must have shape (nrow, 3)")
    freq = np.ascontiguousarray(np.atleast_1d(np.asarray(freq, dtype=np.float64)))
    if nchan_hint is not None and nchan_hint != freq.shape[0]:
        if freq.shape[0] < nchan_hint:
            raise ValueError(f"{freq.shape[0]} frequencies for {nchan_hint} channels")
        freq = np.ascontiguousarray(freq[:nchan_hint])
    return uvw, freq


def _per_sample(a, nrow, nchan, dtype, name):
    if a is None:
        return None
    a = np.asarray(a)
    if a.size == nrow and nchan == 1:
        a = a.reshape(nrow, 1)
    if a.shape != (nrow, nchan):
        raise ValueError(f"{name} must have shape ({nrow}, {nchan}), got {a.shape}")
    return np.ascontiguousarray(a, dtype=dtype)


def _grid(nx, ny):
    nx = int(nx or 0)
    ny = int(ny or 0)
    return (nx, ny) if nx > 0 and ny > 0 else (0, 0)


def ms2dirty(uvw, freq, ms, wgt, nxdirty, nydirty, rad_pix_x, rad_pix_y, nx, ny, epsilon, do_wstacking, *args,
             mask=None):
    """Dirty image (real, [nxdirty, nydirty]) of the visibilities ms [nrow, nchan]."""
    lib = load()
    ms = np.asarray(ms)
    nrow = np.shape(uvw)[0]
    nchan = 1 if ms.ndim == 1 else ms.shape[1]
    uvw, freq = _channels(uvw, freq, nchan)
    fp32 = ms.dtype == np.complex64
    prec = 1 if fp32 else 0
    ms = _per_sample(ms, nrow, nchan, np.complex64 if fp32 else np.complex128, "ms")
    wgt = _per_sample(wgt, nrow, nchan, np.float32 if fp32 else np.float64, "wgt")
    mask = _per_sample(mask, nrow, nchan, np.uint8, "mask")
    dirty = np.zeros((nxdirty, nydirty), dtype=np.complex64 if fp32 else np.complex128)
    nu, nv = _grid(nx, ny)
    check(lib.curig_nifty_ms2dirty(nrow, nchan, uvw, freq, ms.ctypes.data,
                                   None if wgt is None else wgt.ctypes.data,
                                   None if mask is None else mask.ctypes.data, nxdirty, nydirty, float(rad_pix_x),
                                   float(rad_pix_y), nu, nv, float(epsilon), int(bool(do_wstacking)),
                                   dirty.ctypes.data, prec))
    return dirty.real.copy()


def dirty2ms(uvw, freq, dirty, wgt, rad_pix_x, rad_pix_y, nx, ny, epsilon, do_wstacking, *args, mask=None):
    """Visibilities [nrow, nchan] (complex) of the real image dirty [nxdirty, nydirty]."""
    lib = load()
    dirty = np.asarray(dirty)
    uvw, freq = _channels(uvw, freq)
    nrow, nchan = uvw.shape[0], freq.shape[0]
    fp32 = dirty.dtype == np.float32
    prec = 1 if fp32 else 0
    cdt = np.complex64 if fp32 else np.complex128
    nxdirty, nydirty = dirty.shape
    dirty1 = np.zeros(dirty.shape, dtype=cdt)
    dirty1.real = dirty
    wgt = _per_sample(wgt, nrow, nchan, np.float32 if fp32 else np.float64, "wgt")
    mask = _per_sample(mask, nrow, nchan, np.uint8, "mask")
    ms = np.zeros((nrow, nchan), dtype=cdt)
    nu, nv = _grid(nx, ny)
    check(lib.curig_nifty_dirty2ms(nrow, nchan, uvw, freq, ms.ctypes.data, None if wgt is None else wgt.ctypes.data,
                                   None if mask is None else mask.ctypes.data, nxdirty, nydirty, float(rad_pix_x),
                                   float(rad_pix_y), nu, nv, float(epsilon), int(bool(do_wstacking)),
                                   dirty1.ctypes.data, prec))
    return ms
