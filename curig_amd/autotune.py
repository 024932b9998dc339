"""Autotuner for the gridder's bin and batch sizes.

Times candidate settings (bin strip / tile height, w planes per bin layer,
w planes per FFT batch) on a proxy of the caller's problem and applies the
fastest with curig_set_tuning.  The proxy keeps what sets the per-plane cost
(grid size, kernel width, precision, visibilities per w plane) but spans only
~48 w planes: it keeps the rows whose w lies in a band around the median, so
rows per plane match the full problem.  Results are cached in a JSON file
keyed by (grid size, kernel width, precision, direction mix, density class).

    from curig_amd import autotune
    best = autotune.tune(uvw, freq, nx, ny, fov, eps, precision="fp32")
"""
import json
import math
import os
import time

import numpy as np

from . import _lib

CACHE_ENV = "CURIG_TUNING_CACHE"


def _cache_path(path=None):
    if path:
        return path
    if os.environ.get(CACHE_ENV):
        return os.environ[CACHE_ENV]
    return os.path.join(os.path.expanduser("~"), ".cache", "curig_amd", "tuning.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(setting):
    """Apply a tuning dict {tile_x, tile_y, planes_per_batch, planes_per_bin}."""
    _lib.check(_lib.load().curig_set_tuning(int(setting.get("tile_x", 0)), int(setting.get("tile_y", 0)),
                                            int(setting.get("planes_per_batch", 0)),
                                            int(setting.get("planes_per_bin", 0))))


def reset():
    apply({})


def _key(info, precision, nrow_per_plane):
    dens = int(round(math.log2(max(nrow_per_plane, 1.0))))
    return f"nf{info['nf1']}x{info['nf2']}_kw{info['kw']}_{precision}_fft{info['fft_lds']}_d{dens}"


def candidates(info):
    """Candidate settings for the plan's FFT path."""
    out = []
    if info["fft_lds"]:
        # default first (strips of 2 rows, 4-plane layers, budget-sized batches)
        out.append(dict(tile_x=0, tile_y=2, planes_per_batch=0, planes_per_bin=4))
        for ty in (1, 2, 4, 8):
            for pbin in (2, 4, 8, 16):
                if (ty, pbin) != (2, 4):
                    out.append(dict(tile_x=0, tile_y=ty, planes_per_batch=0, planes_per_bin=pbin))
    else:
        for tx, ty in ((32, 32), (64, 32), (64, 16), (32, 16)):
            for pbin in (4, 8):
                out.append(dict(tile_x=tx, tile_y=ty, planes_per_batch=0, planes_per_bin=pbin))
    return out


def _proxy_rows(uvw, target_planes, planes_full):
    """Rows whose w lies in a band around the median holding target/full of them."""
    w = uvw[:, 2]
    frac = min(1.0, target_planes / max(planes_full, 1))
    if frac >= 1.0:
        return np.arange(len(w))
    lo, hi = np.quantile(w, [0.5 - frac / 2, 0.5 + frac / 2])
    return np.nonzero((w >= lo) & (w <= hi))[0]


def tune(uvw, freq, nx, ny, fov, epsilon, precision="fp32", directions=("ms2dirty", "dirty2ms"), cache=None,
         repeat=2, target_planes=48, max_candidates=None, log=None):
    """Measure the candidates on a proxy problem and apply the fastest.

    Returns the chosen setting (dict, with the proxy's measured times)."""
    import torch

    lib = _lib.load()
    prec = 0 if precision == "fp64" else 1
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    info = _lib.plan_info(len(uvw), nx, ny, fov, freq, epsilon, 2.0, -1, float(uvw[:, 2].min()),
                          float(uvw[:, 2].max()), prec)
    rows = _proxy_rows(uvw, target_planes, info["nf3"])
    key = _key(info, precision, len(uvw) / max(info["nf3"], 1)) + "_" + "+".join(directions)
    path = _cache_path(cache)
    store = _load(path)
    if key in store:
        apply(store[key])
        return store[key]
    sub = torch.from_numpy(uvw[rows]).cuda()
    n = len(rows)
    cd = torch.complex128 if prec == 0 else torch.complex64
    rng = np.random.default_rng(0)
    vis = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), dtype=cd, device="cuda")
    img = torch.zeros((nx, ny), dtype=cd, device="cuda")
    vout = torch.zeros(n, dtype=cd, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        if "ms2dirty" in directions:
            _lib.check(lib.curig_ms2dirty_dev(n, nx, ny, fov, freq, sub.data_ptr(), vis.data_ptr(), None,
                                              img.data_ptr(), epsilon, 2.0, -1, prec, stream))
        if "dirty2ms" in directions:
            _lib.check(lib.curig_dirty2ms_dev(n, nx, ny, fov, freq, sub.data_ptr(), vout.data_ptr(), None,
                                              img.data_ptr(), epsilon, 2.0, -1, prec, stream))

    results = []
    cands = candidates(info)
    if max_candidates:
        cands = cands[:max_candidates]
    for cand in cands:
        apply(cand)
        run()  # warm (allocations, plans)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeat):
            run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / repeat
        results.append((dt, cand))
        if log:
            log(f"[autotune] {cand} {dt * 1e3:.2f} ms")
    best_t, best = min(results, key=lambda r: r[0])
    chosen = dict(best, proxy_ms=round(best_t * 1e3, 3), proxy_rows=int(n),
                  default_ms=round(results[0][0] * 1e3, 3))
    store[key] = chosen
    try:
        _save(path, store)
    except OSError:
        pass
    apply(chosen)
    return chosen
