"""Differentiable gridder for PyTorch-ROCm: dirty2ms / ms2dirty as autograd ops.

With A the dirty2ms operator (image -> visibilities,
V_j = wgt_j sum_pix x_pix / n_pix e^{-i phi_j,pix}), ms2dirty is exactly A^H
(the discrete adjoint, same grid, kernel and w planes), so

    Dirty2Ms.backward(g) = ms2dirty(g)        (real part for a real image)
    Ms2Dirty.backward(g) = dirty2ms(g)

Both run the HIP kernels through the device-pointer entry points on torch's
current stream; complex128 / float64 tensors select the FP64 path, complex64 /
float32 the FP32-mixed path.  Conventions as in the C API (NIFTY sign -1 by
default; fov in degrees, image [nx, ny] x-major)."""
import ctypes

import numpy as np
import torch

from ._lib import check, load

__all__ = ["dirty2ms", "ms2dirty", "Dirty2Ms", "Ms2Dirty", "GridderGeometry"]


class GridderGeometry:
    """Fixed part of the operator: uvw (metres, [nrow, 3]), one frequency,
    optional weights, image size, fov (degrees), epsilon, sigma, sign."""

    def __init__(self, uvw, freq, nx, ny, fov, epsilon=1e-6, wgt=None, sigma=2.0, sign=-1, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.uvw = torch.as_tensor(np.ascontiguousarray(uvw, dtype=np.float64) if not torch.is_tensor(uvw) else uvw,
                                   dtype=torch.float64, device=dev).contiguous()
        if self.uvw.dim() != 2 or self.uvw.shape[1] != 3:
            raise ValueError("uvw must be [nrow, 3]")
        self.freq = float(freq)
        self.nx, self.ny = int(nx), int(ny)
        self.fov = float(fov)
        self.epsilon = float(epsilon)
        self.sigma = float(sigma)
        self.sign = int(sign)
        self.wgt = None if wgt is None else torch.as_tensor(wgt, device=dev).reshape(-1)
        self.device = dev

    @property
    def nrow(self):
        return self.uvw.shape[0]

    def _wgt(self, real_dtype):
        return None if self.wgt is None else self.wgt.to(real_dtype).contiguous()


def _precision(cdtype):
    if cdtype == torch.complex128:
        return 0
    if cdtype == torch.complex64:
        return 1
    raise TypeError(f"expected complex64 or complex128, got {cdtype}")


def _real_of(cdtype):
    return torch.float64 if cdtype == torch.complex128 else torch.float32


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _d2m(geo, image_c):
    """image (complex, [nx, ny]) -> visibilities (complex, [nrow])."""
    lib = load()
    image_c = image_c.contiguous()
    vis = torch.empty(geo.nrow, dtype=image_c.dtype, device=image_c.device)
    w = geo._wgt(_real_of(image_c.dtype))
    stream = torch.cuda.current_stream(image_c.device).cuda_stream
    check(lib.curig_dirty2ms_dev(geo.nrow, geo.nx, geo.ny, geo.fov, geo.freq, _ptr(geo.uvw), _ptr(vis), _ptr(w),
                                 _ptr(image_c), geo.epsilon, geo.sigma, geo.sign, _precision(image_c.dtype),
                                 ctypes.c_void_p(stream)))
    return vis


def _m2d(geo, vis_c):
    """visibilities (complex, [nrow]) -> image (complex, [nx, ny])."""
    lib = load()
    vis_c = vis_c.contiguous()
    img = torch.empty((geo.nx, geo.ny), dtype=vis_c.dtype, device=vis_c.device)
    w = geo._wgt(_real_of(vis_c.dtype))
    stream = torch.cuda.current_stream(vis_c.device).cuda_stream
    check(lib.curig_ms2dirty_dev(geo.nrow, geo.nx, geo.ny, geo.fov, geo.freq, _ptr(geo.uvw), _ptr(vis_c), _ptr(w),
                                 _ptr(img), geo.epsilon, geo.sigma, geo.sign, _precision(vis_c.dtype),
                                 ctypes.c_void_p(stream)))
    return img


def _as_complex(x):
    if x.is_complex():
        return x
    return torch.complex(x, torch.zeros_like(x))


class Dirty2Ms(torch.autograd.Function):
    """V = A x.  x: [nx, ny] real or complex image."""

    @staticmethod
    def forward(ctx, image, geo):
        ctx.geo = geo
        ctx.real_input = not image.is_complex()
        return _d2m(geo, _as_complex(image))

    @staticmethod
    def backward(ctx, grad_vis):
        g = _m2d(ctx.geo, grad_vis.contiguous())
        return (g.real.contiguous() if ctx.real_input else g), None


class Ms2Dirty(torch.autograd.Function):
    """d = A^H V.  V: [nrow] complex visibilities; returns the complex image."""

    @staticmethod
    def forward(ctx, vis, geo):
        ctx.geo = geo
        return _m2d(geo, vis)

    @staticmethod
    def backward(ctx, grad_img):
        return _d2m(ctx.geo, _as_complex(grad_img).contiguous()), None


def dirty2ms(image, geo):
    return Dirty2Ms.apply(image, geo)


def ms2dirty(vis, geo):
    return Ms2Dirty.apply(vis, geo)
