"""Benchmark: Mvis/s of ms2dirty + dirty2ms on one MI355X, FP32-mixed vs FP64.

Default workload ("c3"): 256 antennas in a 20 km-radius disk x 768 snapshots over
+-4 h (25,067,520 rows) x 2 channels (1.0 and 1.4 GHz) = 50,135,040 visibilities,
complex64, weights U[0.5, 1]; one 8192 x 8192 image (channels summed), fov 6 deg,
epsilon 1e-5, FP32-mixed.  Inputs are resident in HBM when timing starts; a step
is one ms2dirty over all channels followed by one dirty2ms over all channels.

Also reported: the FP64 path on the same workload, the error of the FP32-mixed
image against the FP64 direct sum on 10,000 random pixels (direct sum on the GPU,
itself checked against the CPU oracle in the tests), the spread kernel's
achieved HBM bandwidth against the 8 TB/s peak, and the oracle's direct sum timed
on the host cores (cpu_baseline).

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c1|c4s] [--precision fp32|fp64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

C_LIGHT = 299792458.0
HBM_PEAK_GBS = 8000.0


# ---------------------------------------------------------------------------
# synthetic observations
# ---------------------------------------------------------------------------
def make_uvw(nant, radius_m, nsnap, ha_hours, dec_deg=-30.0, seed=1):
    """Antennas uniform in a disk (z ~ U[-20, 20] m), all i<j baselines, rotated
    for hour angles linspace(-ha, +ha, nsnap) at declination dec.  [nsnap*nbl, 3]."""
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.random(nant))
    th = 2 * np.pi * rng.random(nant)
    xyz = np.stack([r * np.cos(th), r * np.sin(th), rng.uniform(-20, 20, nant)], axis=1)
    i, j = np.triu_indices(nant, 1)
    b = xyz[j] - xyz[i]
    X, Y, Z = b[:, 0], b[:, 1], b[:, 2]
    dec = np.deg2rad(dec_deg)
    sd, cd = np.sin(dec), np.cos(dec)
    ha = np.deg2rad(np.linspace(-ha_hours, ha_hours, nsnap) * 15.0)
    out = np.empty((nsnap, len(X), 3))
    for s, h in enumerate(ha):
        sh, ch = np.sin(h), np.cos(h)
        out[s, :, 0] = sh * X + ch * Y
        out[s, :, 1] = -sd * ch * X + sd * sh * Y + cd * Z
        out[s, :, 2] = cd * ch * X - cd * sh * Y + sd * Z
    return out.reshape(-1, 3)


CONFIGS = {
    # name: (nant, radius m, nsnap, +-hours, freqs Hz, image, fov deg, eps)
    "c1": (32, 1000.0, 64, 2.0, [1.4e9], 256, 2.0, 1e-6),
    "c2": (128, 5000.0, 256, 3.0, list(np.linspace(1.0e9, 1.4e9, 4)), 2048, 4.0, 1e-6),
    "c3": (256, 20000.0, 768, 4.0, [1.0e9, 1.4e9], 8192, 6.0, 1e-5),
    "c4s": (128, 5000.0, 256, 3.0, list(np.linspace(1.0e9, 1.5e9, 8)), 4096, 3.0, 1e-5),
}


def make_config(name, precision, eps=None, seed=1):
    nant, rad, nsnap, hh, freqs, npix, fov, eps0 = CONFIGS[name]
    uvw = make_uvw(nant, rad, nsnap, hh, seed=seed)
    rng = np.random.default_rng(seed + 1)
    nrow, nchan = uvw.shape[0], len(freqs)
    cd = np.complex64 if precision == "fp32" else np.complex128
    rd = np.float32 if precision == "fp32" else np.float64
    vis = np.empty((nrow, nchan), dtype=cd)
    for c in range(nchan):
        vis[:, c].real = rng.standard_normal(nrow)
        vis[:, c].imag = rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1.0, (nrow, nchan)).astype(rd)
    return dict(name=name, uvw=uvw, vis=vis, wgt=wgt, freqs=np.array(freqs, dtype=np.float64), nx=npix, ny=npix,
                fov=fov, eps=eps if eps is not None else eps0, precision=precision)


# ---------------------------------------------------------------------------
# GPU runner
# ---------------------------------------------------------------------------
class Runner:
    def __init__(self, cfg):
        import torch

        from curig_amd import _lib

        self.torch = torch
        self.L = _lib
        self.lib = _lib.load()
        self.cfg = cfg
        self.prec = 0 if cfg["precision"] == "fp64" else 1
        dev = torch.device("cuda:0")
        self.d_uvw = torch.from_numpy(cfg["uvw"]).to(dev)
        self.d_vis = torch.from_numpy(cfg["vis"]).to(dev)
        self.d_wgt = torch.from_numpy(cfg["wgt"]).to(dev)
        ct = torch.complex64 if self.prec else torch.complex128
        self.d_img = torch.zeros((cfg["nx"], cfg["ny"]), dtype=ct, device=dev)
        self.d_vout = torch.zeros_like(self.d_vis)
        self.nrow, self.nchan = cfg["vis"].shape
        self.stream = torch.cuda.current_stream().cuda_stream

    def ms2dirty(self):
        c = self.cfg
        self.L.check(self.lib.curig_ms2dirty_multi_dev(
            self.nrow, self.nchan, c["nx"], c["ny"], c["fov"], c["freqs"], self.d_uvw.data_ptr(),
            self.d_vis.data_ptr(), self.d_wgt.data_ptr(), None, self.d_img.data_ptr(), 0, c["eps"], 2.0, -1,
            self.prec, self.stream))

    def dirty2ms(self):
        c = self.cfg
        self.L.check(self.lib.curig_dirty2ms_multi_dev(
            self.nrow, self.nchan, c["nx"], c["ny"], c["fov"], c["freqs"], self.d_uvw.data_ptr(),
            self.d_vout.data_ptr(), self.d_wgt.data_ptr(), None, self.d_img.data_ptr(), 0, c["eps"], 2.0, -1,
            self.prec, self.stream))

    def step(self, stats=None):
        self.ms2dirty()
        if stats is not None:
            stats["m2d"].append(self.L.stats())
        self.dirty2ms()
        if stats is not None:
            stats["d2m"].append(self.L.stats())

    def explicit_pixels(self, pix):
        """FP64 direct sum of the channel-summed image at pixels pix (GPU)."""
        torch = self.torch
        c = self.cfg
        d_pix = torch.from_numpy(np.ascontiguousarray(pix, dtype=np.int64)).cuda()
        total = torch.zeros(len(pix), dtype=torch.complex128, device="cuda")
        out = torch.zeros_like(total)
        for ch in range(self.nchan):
            vis_c = self.d_vis[:, ch].contiguous()
            wgt_c = self.d_wgt[:, ch].contiguous()
            self.L.check(self.lib.curig_explicit_ms2dirty_dev(
                self.nrow, c["nx"], c["ny"], c["fov"], float(c["freqs"][ch]), self.d_uvw.data_ptr(),
                vis_c.data_ptr(), wgt_c.data_ptr(), self.prec, len(pix), d_pix.data_ptr(), out.data_ptr(), -1,
                self.stream))
            total += out
        torch.cuda.synchronize()
        return total.cpu().numpy()


def spread_bytes(stats_list, nrow, nchan, precision):
    """Algorithmic bytes of the spread over one ms2dirty (all channels):
    24 R + (s_c + s_w) M + s_c G per channel, G = planes x nf1 x nf2."""
    sc, sw = (8, 4) if precision == "fp32" else (16, 8)
    return sum(24 * nrow + (sc + sw) * s["nvis"] + sc * s["nf3"] * s["nf1"] * s["nf2"] for s in stats_list)


def timed_steps(run, steps, warmup):
    torch = run.torch
    for _ in range(warmup):
        run.step()
    torch.cuda.synchronize()
    run.lib.curig_enable_stats(1)
    per_call = {"m2d": [], "d2m": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(per_call)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run.lib.curig_enable_stats(0)
    return (t1 - t0) / steps, per_call


def cpu_baseline(cfg, target_s=12.0):
    """The oracle's FP64 direct sum (C, OpenMP) on the host cores: terms/s on a
    bounded sample, converted to Mvis/s of a full-image ms2dirty + dirty2ms."""
    from tests import oracle_helpers as O

    lib = O.load()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = lib.oracle_threads(threads)
    uvw = cfg["uvw"]
    vis = cfg["vis"][:, 0].astype(np.complex128)
    nx, ny, fov, f0 = cfg["nx"], cfg["ny"], cfg["fov"], float(cfg["freqs"][0])
    rng = np.random.default_rng(3)
    npix = 64
    pix = rng.choice(nx * ny, npix, replace=False)
    nr = 2000
    while True:
        t0 = time.perf_counter()
        O.dft_ms2dirty_pixels(uvw[:nr], vis[:nr], None, nx, ny, fov, f0, pix, -1, threads)
        dt = time.perf_counter() - t0
        if dt > 0.5 or nr >= len(vis):
            break
        nr = min(len(vis), nr * 4)
    nr = int(min(len(vis), max(nr, nr * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    O.dft_ms2dirty_pixels(uvw[:nr], vis[:nr], None, nx, ny, fov, f0, pix, -1, threads)
    dt = time.perf_counter() - t0
    terms_per_s = nr * npix / dt
    mvis = terms_per_s / (2.0 * nx * ny) / 1e6
    return dict(value=mvis, unit="Mvis/s", cores=cores, kind="port",
                sample=f"oracle FP64 direct sum, {nr} rows x {npix} pixels of the {nx}x{ny} image in {dt:.1f} s "
                       f"({terms_per_s:.3g} terms/s); Mvis/s = terms/s / (2 x {nx * ny} pixels)",
                seconds=dt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--eps", type=float, default=None)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-err", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build has no multi-GPU "
              f"path: each GPU would take a disjoint channel set or w-plane range)", file=sys.stderr)
        sys.exit(2)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    cfg = make_config(args.config, args.precision, args.eps)
    M = int(cfg["vis"].shape[0] * cfg["vis"].shape[1])
    log(f"[bench] {args.config}: {cfg['vis'].shape[0]} rows x {len(cfg['freqs'])} channels = {M} vis, "
        f"{cfg['nx']}^2, fov {cfg['fov']}, eps {cfg['eps']}, {args.precision}")
    run = Runner(cfg)
    t_step, calls = timed_steps(run, args.steps, args.warmup)
    value = M / t_step / 1e6
    m2d = calls["m2d"][-1]
    log(f"[bench] {t_step * 1e3:.1f} ms/step -> {value:.3f} Mvis/s; last ms2dirty stats {m2d}")
    # spread roofline (dominant kernel of ms2dirty), from the per-stage hipEvent times of the timed steps
    t_spread = sum(s["t_grid"] for s in calls["m2d"]) / len(calls["m2d"]) * 1e-3
    stats_by_ch = []
    # per-channel plane counts: recompute from the host plan with the data's w range
    from curig_amd import _lib

    w = cfg["uvw"][:, 2]
    for ch in range(run.nchan):
        stats_by_ch.append(dict(_lib.plan_info(cfg["vis"].shape[0], cfg["nx"], cfg["ny"], cfg["fov"],
                                               float(cfg["freqs"][ch]), cfg["eps"], 2.0, -1, float(w.min()),
                                               float(w.max()), run.prec), nvis=cfg["vis"].shape[0]))
    B = spread_bytes(stats_by_ch, cfg["vis"].shape[0], run.nchan, args.precision)
    achieved = B / t_spread / 1e9
    result = {
        "metric": "Mvis/s ms2dirty+dirty2ms, FP32-mixed vs FP64 at 1e-5 err; % HBM roofline",
        "value": round(value, 4),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (random-array uvw, Gaussian visibilities, seeded)",
        "config": {"workload": f"{args.config}: {cfg['vis'].shape[0]} rows x {run.nchan} channels "
                               f"({', '.join(f'{f / 1e9:.2f}' for f in cfg['freqs'])} GHz) = {M} vis, "
                               f"{cfg['nx']}x{cfg['ny']} image (channels summed), fov {cfg['fov']} deg, "
                               f"eps {cfg['eps']}, {args.precision}",
                   "planes_per_channel": [s["nf3"] for s in stats_by_ch], "grid": stats_by_ch[0]["nf1"],
                   "kw": stats_by_ch[0]["kw"]},
        "roofline": {"bound": "hbm", "kernel": "spread_tile_kernel", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_ms2dirty": B, "t_spread_ms": round(t_spread * 1e3, 2)},
        "stages_ms": {"ms2dirty": {k: round(v, 2) for k, v in m2d.items() if k.startswith("t_")},
                      "dirty2ms": {k: round(v, 2) for k, v in calls["d2m"][-1].items() if k.startswith("t_")}},
    }
    if not args.no_err:
        rng = np.random.default_rng(7)
        pix = rng.choice(cfg["nx"] * cfg["ny"], 10000, replace=False)
        t0 = time.perf_counter()
        truth = run.explicit_pixels(pix)
        img = run.d_img.cpu().numpy().reshape(-1)[pix]
        err = float(np.linalg.norm(img - truth) / np.linalg.norm(truth))
        result["error_vs_fp64_direct_sum"] = {"pixels": 10000, "rel_l2": err,
                                             "seconds": round(time.perf_counter() - t0, 1)}
        log(f"[bench] error vs FP64 direct sum on 10k pixels: {err:.3e}")
    del run
    if args.precision == "fp32" and not args.no_fp64:
        cfg64 = make_config(args.config, "fp64", args.eps)
        run64 = Runner(cfg64)
        t64, _ = timed_steps(run64, 1, 0)
        v64 = M / t64 / 1e6
        result["fp64_path"] = {"value": round(v64, 4), "ms_per_step": round(t64 * 1e3, 1),
                               "fp32_speedup": round(value / v64, 3)}
        log(f"[bench] fp64 path: {v64:.3f} Mvis/s, speedup {value / v64:.2f}x")
        del run64
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cfg)
        log(f"[bench] cpu baseline {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
