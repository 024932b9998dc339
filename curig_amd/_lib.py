"""ctypes binding of libcurafft.so (the HIP gridder).

The four reference symbols carry the exact argtypes of the reference's binding
(python/curagridder/curagridder.py:47-67).  Loading fails loudly when the
library has not been built: there is no CPU fallback."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CURIG_LIBRARY selects another build of the library (same exports)
LIB_PATH = os.environ.get("CURIG_LIBRARY") or os.path.join(HERE, "libcurafft.so")

c_int, c_double, c_long, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p
_f64 = np.ctypeslib.ndpointer(np.double, flags="C")
_c128 = np.ctypeslib.ndpointer(np.complex128, flags="C")

# every symbol include/cugridder.h and include/curig_ext.h declare for C callers
EXPORTS = [
    "ms2dirty_1", "ms2dirty_2", "dirty2ms_1", "dirty2ms_2",
    "ms2dirty_f32", "dirty2ms_f32", "curig_ms2dirty_dev", "curig_dirty2ms_dev", "curig_ms2dirty_dev_noalloc",
    "curig_ms2dirty_multi_dev", "curig_dirty2ms_multi_dev", "curig_explicit_ms2dirty_dev",
    "curig_explicit_dirty2ms_dev",
    "curig_nifty_ms2dirty", "curig_nifty_dirty2ms", "curig_nufft_dev", "curig_nufft3_dev",
    "curig_last_error", "curig_plan_info", "curig_nifty_plan_info", "curig_kernel_fseries", "curig_gauss_legendre",
    "curig_set_tuning", "curig_set_kernel_eval", "curig_kernel_horner", "curig_set_fft_path", "curig_set_graph_mode", "curig_enable_stats", "curig_get_stats", "curig_release",
]
# C++-linkage plan APIs (include/cugridder.h, include/cufft_nufft.h), exported with mangled names
CXX_EXPORTS = ["gridder_setting", "gridder_execution", "gridder_destroy", "cunufft_setting", "cunufft_exec",
               "curafft_free"]

_lib = None


class CurigError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"curig error {code}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libcurafft.so not built ({LIB_PATH}); run python -m curig_amd.build")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7 and
    # the dynamic loader binds every later user of that soname to the first copy
    # loaded, so torch (when installed) is imported first and this library shares
    # its runtime; device pointers then pass freely between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.cdll.LoadLibrary(LIB_PATH)
    ref_args = [c_int, c_int, c_int, c_double, c_double, _f64, _c128, _c128, c_double, c_double, c_int]
    ref_args_w = [c_int, c_int, c_int, c_double, c_double, _f64, _c128, _f64, _c128, c_double, c_double, c_int]
    for name, args in (("ms2dirty_1", ref_args), ("dirty2ms_1", ref_args),
                       ("ms2dirty_2", ref_args_w), ("dirty2ms_2", ref_args_w)):
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = c_int
    host32 = [c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
              c_int]
    for name in ("ms2dirty_f32", "dirty2ms_f32"):
        f = getattr(lib, name)
        f.argtypes = host32
        f.restype = c_int
    dev = [c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
           c_int, c_int, c_void_p]
    for name in ("curig_ms2dirty_dev", "curig_dirty2ms_dev", "curig_ms2dirty_dev_noalloc"):
        f = getattr(lib, name)
        f.argtypes = dev
        f.restype = c_int
    multi = [c_int, c_int, c_int, c_int, c_double, _f64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
             c_double, c_double, c_int, c_int, c_void_p]
    for name in ("curig_ms2dirty_multi_dev", "curig_dirty2ms_multi_dev"):
        f = getattr(lib, name)
        f.argtypes = multi
        f.restype = c_int
    lib.curig_explicit_ms2dirty_dev.argtypes = [c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p,
                                                c_void_p, c_int, c_long, c_void_p, c_void_p, c_int, c_void_p]
    lib.curig_explicit_ms2dirty_dev.restype = c_int
    lib.curig_explicit_dirty2ms_dev.argtypes = [c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p,
                                                c_void_p, c_int, c_long, c_void_p, c_void_p, c_int, c_void_p]
    lib.curig_explicit_dirty2ms_dev.restype = c_int
    nifty = [c_int, c_int, _f64, _f64, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_double, c_int, c_int,
             c_double, c_int, c_void_p, c_int]
    for name in ("curig_nifty_ms2dirty", "curig_nifty_dirty2ms"):
        f = getattr(lib, name)
        f.argtypes = nifty
        f.restype = c_int
    lib.curig_nifty_plan_info.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_int, c_int,
                                          c_int, c_double, c_double, c_int,
                                          np.ctypeslib.ndpointer(np.int32, flags="C"), _f64]
    lib.curig_nifty_plan_info.restype = c_int
    lib.curig_nufft_dev.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int, c_double, c_double, c_int, c_void_p]
    lib.curig_nufft_dev.restype = c_int
    lib.curig_nufft3_dev.argtypes = [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double,
                                     c_double, c_int, c_void_p]
    lib.curig_nufft3_dev.restype = c_int
    lib.curig_last_error.argtypes = []
    lib.curig_last_error.restype = ctypes.c_char_p
    lib.curig_plan_info.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_int, c_double,
                                    c_double, c_int, np.ctypeslib.ndpointer(np.int32, flags="C"), _f64]
    lib.curig_plan_info.restype = c_int
    lib.curig_kernel_fseries.argtypes = [c_double, c_double, c_int, _f64]
    lib.curig_kernel_fseries.restype = c_int
    lib.curig_gauss_legendre.argtypes = [c_int, _f64, _f64]
    lib.curig_gauss_legendre.restype = c_int
    lib.curig_set_tuning.argtypes = [c_int, c_int, c_int, c_int]
    lib.curig_set_tuning.restype = c_int
    lib.curig_set_kernel_eval.argtypes = [c_int]
    lib.curig_set_kernel_eval.restype = c_int
    lib.curig_kernel_horner.argtypes = [c_double, c_double, c_int, np.ctypeslib.ndpointer(np.int32, flags="C"),
                                        _f64, c_int]
    lib.curig_kernel_horner.restype = c_int
    lib.curig_set_fft_path.argtypes = [c_int]
    lib.curig_set_fft_path.restype = c_int
    lib.curig_set_graph_mode.argtypes = [c_int]
    lib.curig_set_graph_mode.restype = c_int
    lib.curig_enable_stats.argtypes = [c_int]
    lib.curig_enable_stats.restype = c_int
    lib.curig_get_stats.argtypes = [np.ctypeslib.ndpointer(np.int32, flags="C"), _f64]
    lib.curig_get_stats.restype = c_int
    lib.curig_release.argtypes = []
    lib.curig_release.restype = None
    _lib = lib
    return lib


def check(code):
    """Raise CurigError for a nonzero return code."""
    if code != 0:
        raise CurigError(code, load().curig_last_error().decode(errors="replace"))
    return code


def plan_info(nrow_valid, nx, ny, fov, freq, eps, sigma, sign, wmin, wmax, precision=0):
    ints = np.zeros(10, dtype=np.int32)
    reals = np.zeros(8, dtype=np.float64)
    check(load().curig_plan_info(nrow_valid, nx, ny, fov, freq, eps, sigma, sign, wmin, wmax, precision, ints, reals))
    keys_i = ["kw", "nf1", "nf2", "nf3", "planes_per_batch", "nbatch", "tile_x", "tile_y", "fft_lds", "rsplit"]
    keys_r = ["beta", "c", "gamma", "h", "i_center", "o_center", "pixelsize_x", "pixelsize_y"]
    d = {k: int(v) for k, v in zip(keys_i, ints)}
    d.update({k: float(v) for k, v in zip(keys_r, reals)})
    return d


def nifty_plan_info(nrow_valid, nx, ny, pixsize_x, pixsize_y, freq, eps, nu, nv, do_wstacking, wmin, wmax,
                    precision=0):
    ints = np.zeros(10, dtype=np.int32)
    reals = np.zeros(8, dtype=np.float64)
    check(load().curig_nifty_plan_info(nrow_valid, nx, ny, pixsize_x, pixsize_y, freq, eps, nu, nv,
                                       int(bool(do_wstacking)), wmin, wmax, precision, ints, reals))
    keys_i = ["kw", "nf1", "nf2", "nf3", "planes_per_batch", "nbatch", "tile_x", "tile_y", "fft_lds", "rsplit"]
    keys_r = ["beta", "c", "gamma", "h", "i_center", "o_center", "pixelsize_x", "pixelsize_y"]
    d = {k: int(v) for k, v in zip(keys_i, ints)}
    d.update({k: float(v) for k, v in zip(keys_r, reals)})
    return d


def stats():
    ints = np.zeros(10, dtype=np.int32)
    reals = np.zeros(8, dtype=np.float64)
    load().curig_get_stats(ints, reals)
    keys_i = ["kw", "nf1", "nf2", "nf3", "planes_per_batch", "nbatch", "tile_x", "tile_y", "fft_lds", "rsplit"]
    keys_r = ["nvis", "t_prep", "t_grid", "t_fft", "t_plane", "t_final", "t_total", "timed"]
    d = {k: int(v) for k, v in zip(keys_i, ints)}
    d.update({k: float(v) for k, v in zip(keys_r, reals)})
    return d
