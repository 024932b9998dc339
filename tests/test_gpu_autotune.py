"""The autotuner (curig_amd/autotune.py) on a small problem: it measures its
candidates, applies one, caches it, and results under the tuned setting still
match the FP64 direct sum."""
import json
import os

import numpy as np
import pytest

from tests import oracle_helpers as O

pytestmark = pytest.mark.gpu
C = 299792458.0


def test_autotune_applies_and_caches(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from curig_amd import _lib, autotune

    print("[autotune] 256^2, 40000 rows", flush=True)
    rng = np.random.default_rng(5)
    n, f0, fov, eps = 256, 1.4e9, 2.0, 1e-6
    uvw = (rng.random((40000, 3)) - 0.5) / (f0 / C) * n * 0.8
    uvw[:, 2] *= 10
    cache = str(tmp_path / "tune.json")
    try:
        best = autotune.tune(uvw, f0, n, n, fov, eps, precision="fp32", cache=cache, repeat=1, max_candidates=6)
        assert {"tile_y", "planes_per_bin", "proxy_ms"} <= set(best)
        stored = json.load(open(cache))
        assert len(stored) == 1 and list(stored.values())[0]["tile_y"] == best["tile_y"]
        again = autotune.tune(uvw, f0, n, n, fov, eps, precision="fp32", cache=cache)
        assert again == list(stored.values())[0]
        # the tuned setting is in force: a call still matches the direct sum
        vis = (rng.standard_normal(len(uvw)) + 1j * rng.standard_normal(len(uvw))).astype(np.complex64)
        dirty = np.zeros((n, n), dtype=np.complex64)
        _lib.check(_lib.load().ms2dirty_f32(len(uvw), n, n, fov, f0, uvw.ctypes.data, vis.ctypes.data, None,
                                            dirty.ctypes.data, eps, 2.0, -1))
        pix = rng.choice(n * n, 200, replace=False)
        want = O.dft_ms2dirty_pixels(uvw, vis.astype(np.complex128), None, n, n, fov, f0, pix, -1)
        got = dirty.reshape(-1)[pix].astype(np.complex128)
        assert O.l2error(want, got) < 3 * eps + 3e-6
    finally:
        autotune.reset()
