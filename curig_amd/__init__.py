"""MI355X-native ES-kernel improved-w-stacking gridder (ms2dirty / dirty2ms).

Drop-in for CuRIG's path: libcurafft.so exports the reference's C ABI
(ms2dirty_1/2, dirty2ms_1/2) and this package mirrors its Python binding
(package `curig`: vis2dirty / dirty2vis and the NIFTY-style ms2dirty /
dirty2ms)."""
from ._lib import CurigError, load, plan_info, stats  # noqa: F401
from .curagridder import dirty2vis, vis2dirty  # noqa: F401
from .cursl import dirty2ms, ms2dirty  # noqa: F401
