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
    python bench.py --form batched    # 64 flagged channels, hipGraph vs per-channel loop, autotuned
    python bench.py --form diff       # 200 Adam iterations of regularised imaging through autograd
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
    # c3 runs at epsilon 1e-6: the ES kernel of width 6 that epsilon 1e-5 selects
    # reaches 1.15e-5 on this workload, width 7 stays below 1e-5 (both directions)
    "c3": (256, 20000.0, 768, 4.0, [1.0e9, 1.4e9], 8192, 6.0, 1e-6),
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

    def dirty2ms_from(self, img):
        c = self.cfg
        self.L.check(self.lib.curig_dirty2ms_multi_dev(
            self.nrow, self.nchan, c["nx"], c["ny"], c["fov"], c["freqs"], self.d_uvw.data_ptr(),
            self.d_vout.data_ptr(), self.d_wgt.data_ptr(), None, img.data_ptr(), 0, c["eps"], 2.0, -1,
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

    def explicit_rows(self, rows):
        """FP64 direct sum of dirty2ms of the current image at rows (GPU), per
        channel: [len(rows), nchan] complex128."""
        torch = self.torch
        c = self.cfg
        d_rows = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.int64)).cuda()
        out = torch.zeros((self.nchan, len(rows)), dtype=torch.complex128, device="cuda")
        for ch in range(self.nchan):
            wgt_c = self.d_wgt[:, ch].contiguous()
            self.L.check(self.lib.curig_explicit_dirty2ms_dev(
                self.nrow, c["nx"], c["ny"], c["fov"], float(c["freqs"][ch]), self.d_uvw.data_ptr(),
                self.d_img.data_ptr(), wgt_c.data_ptr(), self.prec, len(rows), d_rows.data_ptr(), out[ch].data_ptr(),
                -1, self.stream))
        torch.cuda.synchronize()
        return out.cpu().numpy().T

    def errors(self, pix, rows):
        """Relative L2 errors of the current image (10k pixels) and of the
        dirty2ms output (rows x channels) against the FP64 direct sums."""
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        truth = self.explicit_pixels(pix)
        img = self.d_img.cpu().numpy().reshape(-1)[pix].astype(np.complex128)
        vt = self.explicit_rows(rows)
        vg = self.d_vout[rows].cpu().numpy().astype(np.complex128)
        return rel(img, truth), rel(vg, vt), img


def kernel_bytes(plans, nrow, precision, nx, ny):
    """Algorithmic HBM bytes of the plane-loop kernels over one ms2dirty and one
    dirty2ms (all channels), each counted once per launch as the kernel's
    minimum traffic: every input element read once, every output written once.

    s_c = bytes of a complex plane value (8 FP32, 16 FP64); s_w = weight bytes;
    R rows (24 B of uvw each), M visibilities, per channel P planes, planes of
    nf2 x N1 row-transformed cells (N1 = ny image columns), N = nx ny pixels,
    B batches.
      gridding (spread + row FFT):    24 R + (s_c + s_w) M + s_c P nf2 N1
      column FFT + w-accumulation:    s_c P nf2 N1 + 16 N (2 B - 1)
      w-fill + inverse column FFT:    s_c N B + s_c P nf2 N1
      inverse row FFT + interpolation: s_c P nf2 N1 + (32 + 32) M
    (the last: one 32-byte sorted record and one FP64 accumulator read and
    written per visibility).  The tile-spread path (rocFFT planes) writes full
    planes: 24 R + (s_c + s_w) M + s_c P nf1 nf2.

    "grid_full" is the gridding figure with every grid cell of every plane
    written once, 24 R + (s_c + s_w) M + s_c P nf1 nf2, on either path: the
    roofline's algorithmic bytes for the gridding kernel.  The fused LDS-path
    kernel moves less (it writes only the N1 image columns of each
    row-transformed plane), so its "grid" figure is the smaller one."""
    sc, sw = (8, 4) if precision == "fp32" else (16, 8)
    npix = nx * ny
    out = {"grid": 0, "grid_full": 0, "col_fwd": 0, "col_inv": 0, "row_interp": 0, "launches": 0}
    for s in plans:
        M, P = s["nvis"], s["nf3"]
        out["grid_full"] += 24 * nrow + (sc + sw) * M + sc * P * s["nf1"] * s["nf2"]
        if s["fft_lds"]:
            plane = sc * P * s["nf2"] * ny
            out["grid"] += 24 * nrow + (sc + sw) * M + plane
            out["col_fwd"] += plane + 16 * npix * (2 * s["nbatch"] - 1)
            out["col_inv"] += sc * npix * s["nbatch"] + plane
            out["row_interp"] += plane + 64 * M
        else:
            out["grid"] += 24 * nrow + (sc + sw) * M + sc * P * s["nf1"] * s["nf2"]
        out["launches"] += s["nbatch"]
    return out


def timed_steps(run, steps, warmup, stats=True):
    """Seconds per step over `steps` steps after `warmup` untimed ones.  With
    stats, every call also records HIP events around its stages on the launch
    stream (read once the call's work is done) and the per-call records are
    returned; without, the steps run exactly as a caller's would."""
    torch = run.torch
    for _ in range(warmup):
        run.step()
    torch.cuda.synchronize()
    run.lib.curig_enable_stats(1 if stats else 0)
    per_call = {"m2d": [], "d2m": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step(per_call if stats else None)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run.lib.curig_enable_stats(0)
    return (t1 - t0) / steps, per_call


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
    scale = target_s / max(dt, 1e-3)
    nr_new = int(min(len(vis), max(nr, nr * scale)))
    # all rows in use: make up the rest of the sample with more pixels
    npix = int(min(nx * ny, max(npix, npix * scale * nr / nr_new)))
    nr = nr_new
    pix = rng.choice(nx * ny, npix, replace=False)
    t0 = time.perf_counter()
    O.dft_ms2dirty_pixels(uvw[:nr], vis[:nr], None, nx, ny, fov, f0, pix, -1, threads)
    dt = time.perf_counter() - t0
    terms_per_s = nr * npix / dt
    mvis = terms_per_s / (2.0 * nx * ny) / 1e6
    return dict(value=mvis, unit="Mvis/s", cores=cores, kind="port",
                sample=f"oracle FP64 direct sum, {nr} rows x {npix} pixels of the {nx}x{ny} image in {dt:.1f} s "
                       f"({terms_per_s:.3g} terms/s); Mvis/s = terms/s / (2 x {nx * ny} pixels)",
                seconds=dt, cpu_model=cpu_model(), nproc=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def cpu_c1(log):
    """The 32-antenna, 1 km configuration (31,744 rows, 256^2, fov 2 deg,
    1.4 GHz, epsilon 1e-6, FP64) whole: the oracle's direct sums of ms2dirty and
    dirty2ms on the host cores, and the GPU FP64 path on the same data in the
    same run (device-resident, one warm-up), with the GPU results' errors
    against the CPU sums over the whole image / all rows."""
    import torch

    from tests import oracle_helpers as O

    lib = O.load()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = lib.oracle_threads(threads)
    cfg = make_config("c1", "fp64")
    uvw, vis, wgt = cfg["uvw"], cfg["vis"][:, 0], cfg["wgt"][:, 0]
    nx, fov, f0 = cfg["nx"], cfg["fov"], float(cfg["freqs"][0])
    M = len(vis)
    t0 = time.perf_counter()
    d_cpu = O.dft_ms2dirty(uvw, vis, wgt, nx, nx, fov, f0, -1, threads)
    t_m2d = time.perf_counter() - t0
    rng = np.random.default_rng(4)
    img = rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx))
    t0 = time.perf_counter()
    v_cpu = O.dft_dirty2ms(uvw, img, wgt, nx, nx, fov, f0, -1, threads)
    t_d2m = time.perf_counter() - t0
    run = Runner(cfg)
    run.L.check(run.lib.curig_set_graph_mode(0))

    def step():
        run.ms2dirty()
        run.dirty2ms_from(run.d_img2)

    run.d_img2 = torch.from_numpy(img).cuda()
    step()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t0) / reps
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    e_img = rel(run.d_img.cpu().numpy(), d_cpu)
    e_vis = rel(run.d_vout[:, 0].cpu().numpy(), v_cpu)
    del run
    terms = 2.0 * M * nx * nx
    out = {"workload": f"c1: {M} rows, {nx}x{nx}, fov {fov} deg, {f0 / 1e9:.1f} GHz, FP64",
           "cpu": {"ms2dirty_s": round(t_m2d, 3), "dirty2ms_s": round(t_d2m, 3),
                   "mvis_per_s": round(M / (t_m2d + t_d2m) / 1e6, 6),
                   "terms_per_s": round(terms / (t_m2d + t_d2m), 1), "cores": cores, "cpu_model": cpu_model(),
                   "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
           "gpu_fp64": {"ms_per_step": round(t_gpu * 1e3, 3), "mvis_per_s": round(M / t_gpu / 1e6, 3),
                        "error_vs_cpu_direct_sum": {"ms2dirty_full_image": e_img, "dirty2ms_all_rows": e_vis}}}
    log(f"[bench] c1: CPU direct sums {t_m2d:.2f} + {t_d2m:.2f} s on {cores} threads ({cpu_model()}); "
        f"GPU FP64 {t_gpu * 1e3:.2f} ms/step, errors {e_img:.2e} / {e_vis:.2e}")
    return out


def api_step_seconds(cfg):
    """One ms2dirty + dirty2ms over all channels through the host-pointer C ABI
    (ms2dirty_f32 / dirty2ms_f32 or the reference's ms2dirty_2 / dirty2ms_2),
    PCIe copies and per-call setup included."""
    import ctypes

    from curig_amd import _lib

    lib = _lib.load()
    uvw = np.ascontiguousarray(cfg["uvw"])
    nrow, nchan = cfg["vis"].shape
    fp32 = cfg["precision"] == "fp32"
    cd = np.complex64 if fp32 else np.complex128
    vis = [np.ascontiguousarray(cfg["vis"][:, c]) for c in range(nchan)]
    wgt = [np.ascontiguousarray(cfg["wgt"][:, c]) for c in range(nchan)]
    img = np.zeros((cfg["nx"], cfg["ny"]), dtype=cd)
    vout = np.zeros(nrow, dtype=cd)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    t0 = time.perf_counter()
    for c in range(nchan):
        f = float(cfg["freqs"][c])
        if fp32:
            _lib.check(lib.ms2dirty_f32(nrow, cfg["nx"], cfg["ny"], cfg["fov"], f, ptr(uvw), ptr(vis[c]), ptr(wgt[c]),
                                        ptr(img), cfg["eps"], 2.0, -1))
            _lib.check(lib.dirty2ms_f32(nrow, cfg["nx"], cfg["ny"], cfg["fov"], f, ptr(uvw), ptr(vout), ptr(wgt[c]),
                                        ptr(img), cfg["eps"], 2.0, -1))
        else:
            _lib.check(lib.ms2dirty_2(nrow, cfg["nx"], cfg["ny"], cfg["fov"], f, uvw, vis[c], wgt[c], img, cfg["eps"],
                                      2.0, -1))
            _lib.check(lib.dirty2ms_2(nrow, cfg["nx"], cfg["ny"], cfg["fov"], f, uvw, vout, wgt[c], img, cfg["eps"],
                                      2.0, -1))
    return time.perf_counter() - t0


# ---------------------------------------------------------------------------
# batched form: 64 flagged channels, one image per channel
# ---------------------------------------------------------------------------
def form_nifty(args, log):
    """c2 through the NIFTY host form (curig_amd.cursl: 4 channels in one call,
    host arrays in and out, FP64, pixel size fov/2048 on both axes, a 10 % random
    mask) against the reference ABI's per-channel loop over the same unmasked
    samples (ms2dirty_2 / dirty2ms_2 per channel, masked samples given weight 0).
    Both include the host<->device copies."""
    from curig_amd import _lib, cursl

    lib = _lib.load()
    cfg = make_config("c2", "fp64", args.eps)
    uvw, vis, freqs = cfg["uvw"], cfg["vis"], cfg["freqs"]
    nrow, nchan, npix, eps = uvw.shape[0], len(freqs), cfg["nx"], cfg["eps"]
    rng = np.random.default_rng(7)
    mask = (rng.random((nrow, nchan)) > 0.1).astype(np.uint8)
    wgt = np.ascontiguousarray(cfg["wgt"])
    nvalid = int(mask.sum(dtype=np.int64))
    pix = cfg["fov"] * np.pi / 180 / npix
    img = rng.standard_normal((npix, npix))
    log(f"[bench] nifty: {nrow} rows x {nchan} channels, {nvalid} unmasked vis, {npix}^2, eps {eps}, fp64")

    def step_nifty():
        d = cursl.ms2dirty(uvw, freqs, vis, wgt, npix, npix, pix, pix, 0, 0, eps, True, mask=mask)
        v = cursl.dirty2ms(uvw, freqs, img, wgt, pix, pix, 0, 0, eps, True, mask=mask)
        return d, v

    wz = np.ascontiguousarray(wgt * mask)
    img_c = np.ascontiguousarray(img + 0j)

    def step_loop():
        d = np.zeros((npix, npix))
        v = np.zeros((nrow, nchan), complex)
        for c in range(nchan):
            out = np.zeros((npix, npix), complex)
            _lib.check(lib.ms2dirty_2(nrow, npix, npix, cfg["fov"], freqs[c], uvw, np.ascontiguousarray(vis[:, c]),
                                      np.ascontiguousarray(wz[:, c]), out, eps, 2.0, -1))
            d += out.real
            vc = np.zeros(nrow, complex)
            _lib.check(lib.dirty2ms_2(nrow, npix, npix, cfg["fov"], freqs[c], uvw, vc,
                                      np.ascontiguousarray(wz[:, c]), img_c, eps, 2.0, -1))
            v[:, c] = vc
        return d, v

    def timed(fn):
        for _ in range(args.warmup):
            r = fn()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        return (time.perf_counter() - t0) / args.steps, r

    t_n, (d_n, v_n) = timed(step_nifty)
    t_l, (d_l, v_l) = timed(step_loop)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    d_rel, v_rel = rel(d_n, d_l), rel(v_n, v_l)
    value = nvalid / t_n / 1e6
    log(f"[bench] nifty {t_n * 1e3:.1f} ms/step ({value:.2f} Mvis/s), reference-ABI channel loop "
        f"{t_l * 1e3:.1f} ms/step ({nvalid / t_l / 1e6:.2f} Mvis/s); rel diff image {d_rel:.2e} vis {v_rel:.2e}")
    return {
        "metric": "Mvis/s ms2dirty+dirty2ms, NIFTY host form (PCIe-inclusive)",
        "value": round(value, 4), "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_n * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (random-array uvw, Gaussian visibilities, 10% random mask, seeded)",
        "config": {"workload": f"c2 via NIFTY host form: {nrow} rows x {nchan} channels (1.0-1.4 GHz), "
                               f"{nvalid} unmasked vis, {npix}x{npix}, fov {cfg['fov']} deg, eps {eps}, fp64"},
        "reference_abi_loop": {"ms_per_step": round(t_l * 1e3, 2), "value": round(nvalid / t_l / 1e6, 4)},
        "nifty_vs_loop_rel": {"image": d_rel, "vis": v_rel},
    }


def form_nufft(args, log):
    """Generic 2-D NUFFT: one type 1 and one type 2 per step, 4096 x 4096 modes,
    20M points uniform in [-pi, pi)^2, epsilon 1e-5, FP32-mixed, inputs resident
    on the GPU (curig_nufft_dev).  Parity is covered by the tests."""
    import torch

    from curig_amd import _lib

    lib = _lib.load()
    N, M, eps = 4096, 20_000_000, args.eps or 1e-5
    rng = np.random.default_rng(5)
    dev = torch.device("cuda:0")
    u = torch.from_numpy(rng.uniform(-np.pi, np.pi, M)).to(dev)
    v = torch.from_numpy(rng.uniform(-np.pi, np.pi, M)).to(dev)
    c = torch.complex(torch.randn(M, device=dev), torch.randn(M, device=dev))
    fk = torch.zeros((N, N), dtype=torch.complex64, device=dev)
    c2 = torch.zeros_like(c)
    st = torch.cuda.current_stream().cuda_stream
    log(f"[bench] nufft: 2-D types 1 + 2, {N}x{N} modes, {M} points, eps {eps}, fp32")

    def step():
        _lib.check(lib.curig_nufft_dev(1, 2, N, N, 1, M, u.data_ptr(), v.data_ptr(), None, c.data_ptr(),
                                       fk.data_ptr(), 1, eps, 2.0, 1, st))
        _lib.check(lib.curig_nufft_dev(2, 2, N, N, 1, M, u.data_ptr(), v.data_ptr(), None, c2.data_ptr(),
                                       fk.data_ptr(), -1, eps, 2.0, 1, st))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    value = M / t / 1e6
    log(f"[bench] nufft {t * 1e3:.1f} ms/step ({value:.1f} Mpts/s, type 1 + type 2)")
    return {
        "metric": "Mpts/s generic 2-D NUFFT type 1 + type 2", "value": round(value, 3), "unit": "Mpts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (uniform points, Gaussian strengths, seeded)",
        "config": {"workload": f"2-D NUFFT {N}x{N} modes, {M} points in [-pi,pi)^2, eps {eps}, fp32"},
    }


def form_batched(args, log):
    """64 channels linspace(1.0, 1.5) GHz of the 128-antenna 5 km array (256
    snapshots over +-3 h), per-channel flag fraction U[0, 0.4] as i.i.d. row
    masks, complex64, one 4096^2 image per channel, fov 3 deg, epsilon 1e-5.
    Bin and batch sizes from the autotuner; the hipGraph-captured multi-channel
    sequence is timed against the per-channel launch loop and compared."""
    import tempfile

    import torch

    from curig_amd import _lib, autotune

    lib = _lib.load()
    nchan, npix, fov, eps = 64, 4096, 3.0, args.eps or 1e-5
    freqs = np.linspace(1.0e9, 1.5e9, nchan)
    uvw = make_uvw(128, 5000.0, 256, 3.0, seed=1)
    nrow = len(uvw)
    rng = np.random.default_rng(11)
    frac = rng.uniform(0.0, 0.4, nchan)
    flags = (rng.random((nrow, nchan)) < frac[None, :]).astype(np.uint8)
    nvalid = int(nrow * nchan - flags.sum(dtype=np.int64))
    vis = np.empty((nrow, nchan), dtype=np.complex64)
    vis.real = rng.standard_normal((nrow, nchan), dtype=np.float32)
    vis.imag = rng.standard_normal((nrow, nchan), dtype=np.float32)
    wgt = rng.uniform(0.5, 1.0, (nrow, nchan)).astype(np.float32)
    counts = nrow - flags.sum(axis=0)
    log(f"[bench] batched: {nrow} rows x {nchan} channels, {nvalid} unflagged vis "
        f"({counts.min()}..{counts.max()} per channel), {npix}^2 per channel, fov {fov}, eps {eps}")
    dev = torch.device("cuda:0")
    d_uvw = torch.from_numpy(uvw).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    d_wgt = torch.from_numpy(wgt).to(dev)
    d_flag = torch.from_numpy(flags).to(dev)
    d_img = torch.zeros((nchan, npix, npix), dtype=torch.complex64, device=dev)
    d_vout = torch.zeros_like(d_vis)
    stream = torch.cuda.current_stream().cuda_stream
    # autotune on the middle channel's unflagged rows
    t0 = time.perf_counter()
    cache = os.path.join(tempfile.mkdtemp(prefix="curig_tune_"), "tuning.json")
    mid = nchan // 2
    tuned = autotune.tune(uvw[flags[:, mid] == 0], float(freqs[mid]), npix, npix, fov, eps, precision="fp32",
                          cache=cache, log=log)
    t_tune = time.perf_counter() - t0
    log(f"[bench] autotuned in {t_tune:.1f} s: {tuned}")

    def step():
        _lib.check(lib.curig_ms2dirty_multi_dev(nrow, nchan, npix, npix, fov, freqs, d_uvw.data_ptr(),
                                                d_vis.data_ptr(), d_wgt.data_ptr(), d_flag.data_ptr(),
                                                d_img.data_ptr(), 1, eps, 2.0, -1, 1, stream))
        _lib.check(lib.curig_dirty2ms_multi_dev(nrow, nchan, npix, npix, fov, freqs, d_uvw.data_ptr(),
                                                d_vout.data_ptr(), d_wgt.data_ptr(), d_flag.data_ptr(),
                                                d_img.data_ptr(), 1, eps, 2.0, -1, 1, stream))

    def timed(mode):
        _lib.check(lib.curig_set_graph_mode(mode))
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_graph = timed(1)
    img_g, vis_g = d_img.clone(), d_vout.clone()
    t_loop = timed(0)
    _lib.check(lib.curig_set_graph_mode(1))
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    d_img_rel, d_vis_rel = rel(img_g, d_img), rel(vis_g, d_vout)
    value = nvalid / t_graph / 1e6
    log(f"[bench] graph {t_graph * 1e3:.1f} ms/step ({value:.2f} Mvis/s), loop {t_loop * 1e3:.1f} ms/step; "
        f"graph vs loop rel diff image {d_img_rel:.2e} vis {d_vis_rel:.2e}")
    return {
        "metric": "Mvis/s ms2dirty+dirty2ms, batched flagged channels, hipGraph",
        "value": round(value, 4), "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_graph * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (random-array uvw, Gaussian visibilities, i.i.d. row flags, seeded)",
        "config": {"workload": f"batched: {nrow} rows x {nchan} channels (1.0-1.5 GHz), flag fraction U[0,0.4], "
                               f"{nvalid} unflagged vis, {npix}x{npix} image per channel, fov {fov} deg, eps {eps}, "
                               f"fp32"},
        "per_channel_loop_ms": round(t_loop * 1e3, 2), "graph_speedup": round(t_loop / t_graph, 3),
        "graph_vs_loop_rel_diff": {"image": d_img_rel, "vis": d_vis_rel},
        "autotune": dict(tuned, seconds=round(t_tune, 1)),
    }


# ---------------------------------------------------------------------------
# differentiation form: 200 Adam iterations of regularised imaging
# ---------------------------------------------------------------------------
def form_diff(args, log):
    """x (1024^2, FP32) fitted to 1,032,192 complex64 visibilities of 30 point
    sources (64 antennas in a 3 km disk, 512 snapshots over +-4 h, 1.4 GHz,
    fov 2 deg, weights U[0.5, 1]) by minimising ||W^1/2 (dirty2ms(x) - V)||^2
    + 1e-3 TV(x) with Adam; backward = ms2dirty.  Reports the time of 200
    iterations, the adjointness <Ax, y> = <x, A^H y> and central finite
    differences (h = 1e-3) of the loss on 100 random pixels."""
    import torch

    from curig_amd.autograd import GridderGeometry, dirty2ms, ms2dirty

    n, fov, f0, eps, lam = 1024, 2.0, 1.4e9, args.eps or 1e-5, 1e-3
    uvw = make_uvw(64, 3000.0, 512, 4.0, seed=5)
    nrow = len(uvw)
    rng = np.random.default_rng(21)
    wgt = rng.uniform(0.5, 1.0, nrow)
    # 30 point sources on pixel centres: V = sum_s F_s / n_s e^{-2 pi i f/c (u l + v m - w (n - 1))}
    pix = rng.choice(n * n, 30, replace=False)
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), 30))
    d = fov * np.pi / 180 / n
    ls, ms = (pix // n - n // 2) * d, (pix % n - n // 2) * d
    r2 = ls ** 2 + ms ** 2
    nm1 = -r2 / (np.sqrt(1 - r2) + 1)
    k = 2 * np.pi * f0 / C_LIGHT
    V = np.zeros(nrow, dtype=np.complex128)
    for s in range(30):
        V += flux[s] / (nm1[s] + 1) * np.exp(-1j * k * (uvw[:, 0] * ls[s] + uvw[:, 1] * ms[s] - uvw[:, 2] * nm1[s]))
    V *= wgt  # the operator multiplies by the weight
    geo = GridderGeometry(uvw, f0, n, n, fov, epsilon=eps, wgt=wgt.astype(np.float32))
    dev = geo.device
    data = torch.tensor(V, dtype=torch.complex64, device=dev)
    sw = torch.tensor(np.sqrt(wgt), dtype=torch.float32, device=dev)

    def tv(x, e=1e-8):
        dx = x[1:, :-1] - x[:-1, :-1]
        dy = x[:-1, 1:] - x[:-1, :-1]
        return (dx * dx + dy * dy + e).sqrt().sum()

    def loss_fn(x):
        r = (dirty2ms(x, geo) - data) * sw
        return (r.real.double() ** 2 + r.imag.double() ** 2).sum() + lam * tv(x.double())

    # adjointness and finite differences
    xa = torch.randn((n, n), dtype=torch.complex64, device=dev)
    ya = torch.randn(nrow, dtype=torch.complex64, device=dev)
    lhs = torch.vdot(dirty2ms(xa, geo).to(torch.complex128), ya.to(torch.complex128)).item()
    rhs = torch.vdot(xa.reshape(-1).to(torch.complex128), ms2dirty(ya, geo).reshape(-1).to(torch.complex128)).item()
    adj = abs(lhs - rhs) / abs(lhs)
    x0 = torch.randn((n, n), dtype=torch.float32, device=dev) * 0.01
    xg = x0.clone().requires_grad_(True)
    loss_fn(xg).backward()
    grad = xg.grad.double().reshape(-1).cpu().numpy()
    h = 1e-3
    fpix = rng.choice(n * n, 100, replace=False)
    fd = np.empty(100)
    with torch.no_grad():
        for q, p in enumerate(fpix):
            e = torch.zeros(n * n, dtype=torch.float32, device=dev)
            e[p] = h
            e = e.reshape(n, n)
            fd[q] = (loss_fn(x0 + e).item() - loss_fn(x0 - e).item()) / (2 * h)
    fd_err = float(np.linalg.norm(fd - grad[fpix]) / np.linalg.norm(grad[fpix]))
    log(f"[bench] diff: adjointness {adj:.2e}, FD rel error {fd_err:.2e}")

    x = torch.zeros((n, n), dtype=torch.float32, device=dev, requires_grad=True)
    opt = torch.optim.Adam([x], lr=args.lr)
    iters = 200
    losses = []
    for it in range(args.warmup):
        opt.zero_grad()
        loss_fn(x).backward()
    with torch.no_grad():
        x.zero_()
    opt = torch.optim.Adam([x], lr=args.lr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(iters):
        opt.zero_grad(set_to_none=True)
        L = loss_fn(x)
        L.backward()
        opt.step()
        if it in (0, iters - 1):
            losses.append(L.item())
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    peak = [float(x.detach().reshape(-1)[int(p)].item()) for p in pix[np.argsort(-flux)[:3]]]
    log(f"[bench] diff: 200 Adam iterations in {t:.2f} s; loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    return {
        "metric": "seconds for 200 Adam iterations (dirty2ms forward, ms2dirty backward)",
        "value": round(t, 3), "unit": "s", "n_gpus": 1, "steps": iters, "warmup": args.warmup,
        "ms_per_step": round(t / iters * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (random-array uvw, 30 point sources, seeded)",
        "config": {"workload": f"diff: {nrow} vis (64 antennas, 3 km, 512 snapshots), {n}x{n} FP32 image, fov {fov} "
                               f"deg, 1.4 GHz, eps {eps}, lambda {lam}, Adam lr {args.lr}"},
        "mvis_per_s": round(2 * nrow * iters / t / 1e6, 3),
        "adjointness_rel": adj, "fd_rel_error_100px_h1e-3": fd_err,
        "loss_first_last": losses, "brightest_sources_recovered": peak,
        "brightest_sources_true": [float(v) for v in np.sort(flux)[::-1][:3]],
    }


def traffic_for(kernel, args):
    """HBM bytes per launch of `kernel` measured by the FETCH_SIZE / WRITE_SIZE
    counter runs recorded in profiles/ (calibrated against a 1 GiB copy), for
    the default workload; --traffic overrides; None when not measured."""
    if args.traffic is not None:
        return args.traffic
    path = os.path.join(ROOT, "profiles", f"{args.config}_{args.precision}_traffic.json")
    if args.eps is not None or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("bytes_per_launch", {}).get(kernel)


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
    ap.add_argument("--no-api", action="store_true", help="skip the host-pointer (PCIe-inclusive) timing")
    ap.add_argument("--no-eps4", action="store_true", help="skip the epsilon 1e-4 run of the same workload")
    ap.add_argument("--kereval", type=int, default=0,
                    help="kernel evaluation: 0 exp(sqrt()), 1 piecewise polynomial (Horner)")
    ap.add_argument("--form", default="precision", choices=["precision", "batched", "diff", "nifty", "nufft"])
    ap.add_argument("--lr", type=float, default=0.02, help="Adam step of the diff form")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel from a counter run (reported as given)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device copy first (counter calibration of FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build has no multi-GPU "
              f"path: each GPU would take a disjoint channel set or w-plane range)", file=sys.stderr)
        sys.exit(2)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.kereval:
        from curig_amd import _lib

        _lib.check(_lib.load().curig_set_kernel_eval(args.kereval))
        log(f"[bench] kernel evaluation: piecewise polynomial (Horner)")
    if args.form != "precision":
        result = {"batched": form_batched, "diff": form_diff, "nifty": form_nifty,
                  "nufft": form_nufft}[args.form](args, log)
        print(json.dumps(result), flush=True)
        return
    cfg = make_config(args.config, args.precision, args.eps)
    M = int(cfg["vis"].shape[0] * cfg["vis"].shape[1])
    log(f"[bench] {args.config}: {cfg['vis'].shape[0]} rows x {len(cfg['freqs'])} channels = {M} vis, "
        f"{cfg['nx']}^2, fov {cfg['fov']}, eps {cfg['eps']}, {args.precision}")
    if args.calib_copy:
        import torch

        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    run = Runner(cfg)
    # one channel per launch sequence (no graph capture), so the per-stage HIP
    # events of the timed steps split the gridding kernel from the FFT passes
    run.L.check(run.lib.curig_set_graph_mode(0))
    # the headline steps run without stage timing; a second timed region of as
    # many steps records the per-stage HIP events the kernel split comes from
    t_step, _ = timed_steps(run, args.steps, args.warmup, stats=False)
    value = M / t_step / 1e6
    log(f"[bench] {t_step * 1e3:.1f} ms/step -> {value:.3f} Mvis/s")
    t_kstep, calls = timed_steps(run, args.steps, 0, stats=True)
    m2d = calls["m2d"][-1]
    log(f"[bench] kernel-timing pass {t_kstep * 1e3:.1f} ms/step; last ms2dirty stats {m2d}")
    # per-kernel achieved bandwidth, from the per-stage HIP events the library
    # records on the launch stream during the timed steps
    stats_by_ch = []
    from curig_amd import _lib

    w = cfg["uvw"][:, 2]
    for ch in range(run.nchan):
        stats_by_ch.append(dict(_lib.plan_info(cfg["vis"].shape[0], cfg["nx"], cfg["ny"], cfg["fov"],
                                               float(cfg["freqs"][ch]), cfg["eps"], 2.0, -1, float(w.min()),
                                               float(w.max()), run.prec), nvis=cfg["vis"].shape[0]))
    kb = kernel_bytes(stats_by_ch, cfg["vis"].shape[0], args.precision, cfg["nx"], cfg["ny"])
    launches = kb["launches"]
    mean = lambda key, d: sum(c[key] for c in calls[d]) / len(calls[d]) * 1e-3
    lds = bool(stats_by_ch[0]["fft_lds"])
    split = stats_by_ch[0]["rsplit"] > 0  # rows split over several workgroups
    part = "_part" if split else ""
    # FP32 rows of 4096 run the plane-loop row kernel whole (one part)
    planes = split or (run.prec == 1 and stats_by_ch[0]["nf1"] == 4096)
    spread = ("spread_row_part_planes_kernel" if planes else "spread_row_fft_kernel") if lds else "spread_tile_kernel"
    kernels = {spread: (kb["grid"], mean("t_grid", "m2d"))}
    if lds:
        # columns of 16384 run the register-resident column kernels
        reg = "_reg" if stats_by_ch[0]["nf2"] == 16384 else ""
        kernels[f"col_fwd{reg}_kernel"] = (kb["col_fwd"], mean("t_fft", "m2d"))
        kernels[f"col_inv{reg}_kernel"] = (kb["col_inv"], mean("t_fft", "d2m"))
        # rows and row parts of 4096 run the plane-loop interpolation kernel
        nf1 = stats_by_ch[0]["nf1"]
        rs = stats_by_ch[0]["rsplit"]
        ipl = nf1 // max(rs, 1) == 4096
        kernels["row_ifft_interp_part_planes_kernel" if ipl else f"row_ifft_interp{part}_kernel"] = (
            kb["row_interp"], mean("t_grid", "d2m"))
    per_kernel = {k: {"bytes_per_launch": round(b / launches), "avg_launch_ms": round(t / launches * 1e3, 3),
                      "achieved_GBps": round(b / t / 1e9, 1), "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4)}
                  for k, (b, t) in kernels.items() if t > 0}
    # the roofline prices the gridding kernel with the full-plane figure (every
    # grid cell of every plane written once); the row interpolation, which
    # takes about as long, reads the same planes: its mirror figure is
    # reported beside it
    kname = spread
    B, t_dom = kb["grid_full"], kernels[kname][1]
    achieved = B / t_dom / 1e9
    interp = next((k for k in kernels if k.startswith("row_ifft_interp")), None)
    mirror = None
    if interp and kernels[interp][1] > 0:
        a_i = kb["grid_full"] / kernels[interp][1] / 1e9
        mirror = {"kernel": interp, "achieved": round(a_i, 1), "frac": round(a_i / HBM_PEAK_GBS, 4),
                  "avg_launch_ms": round(kernels[interp][1] / launches * 1e3, 3),
                  "bytes_model": "24 R + (s_c + s_w) M + s_c P nf1 nf2: every grid cell of every plane read once"}
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
                   "kw": stats_by_ch[0]["kw"], "kernel_eval": "horner" if args.kereval else "exp_sqrt"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_for(kname, args), "launches_per_call": launches,
                     "bytes_per_launch": round(B / launches), "avg_launch_ms": round(t_dom / launches * 1e3, 3),
                     "bytes_model": "24 R + (s_c + s_w) M + s_c P nf1 nf2: every grid cell of every plane written once",
                     "interp_full_plane": mirror, "kernels": per_kernel},
        "kernel_timing_pass_ms_per_step": round(t_kstep * 1e3, 2),
        "stages_ms": {"ms2dirty": {k: round(v, 2) for k, v in m2d.items() if k.startswith("t_")},
                      "dirty2ms": {k: round(v, 2) for k, v in calls["d2m"][-1].items() if k.startswith("t_")}},
    }
    rng = np.random.default_rng(7)
    pix = rng.choice(cfg["nx"] * cfg["ny"], 10000, replace=False)
    rows = np.sort(rng.choice(cfg["vis"].shape[0], 1000, replace=False))
    img32 = None
    if not args.no_err:
        t0 = time.perf_counter()
        e_img, e_vis, img32 = run.errors(pix, rows)
        result["error_vs_fp64_direct_sum"] = {"ms2dirty_10k_pixels": e_img,
                                             f"dirty2ms_{len(rows)}_rows_x_{run.nchan}_channels": e_vis,
                                             "seconds": round(time.perf_counter() - t0, 1)}
        log(f"[bench] error vs FP64 direct sum: image (10k pixels) {e_img:.3e}, visibilities "
            f"({len(rows)} rows x {run.nchan} channels) {e_vis:.3e}")
    if args.precision == "fp32" and not args.no_eps4:
        # the same workload at the wider epsilons (kernel widths 6 and 5): one step each
        for e in (1e-5, 1e-4):
            if e == cfg["eps"]:
                continue
            run.cfg = dict(cfg, eps=e)
            te, _ = timed_steps(run, 1, 0, stats=False)
            leg = {"epsilon": e, "value": round(M / te / 1e6, 4), "ms_per_step": round(te * 1e3, 1)}
            if not args.no_err:
                ei, ev, _ = run.errors(pix, rows)
                leg["error_vs_fp64_direct_sum"] = {"ms2dirty": ei, "dirty2ms": ev}
            result[f"epsilon_{e:.0e}"] = leg
            log(f"[bench] eps {e:.0e}: {leg}")
        run.cfg = cfg
    del run
    if not args.no_api:
        t_api = api_step_seconds(cfg)
        result["api_host_pointers"] = {"value": round(M / t_api / 1e6, 4), "unit": "Mvis/s",
                                       "ms_per_step": round(t_api * 1e3, 1),
                                       "note": "host buffers through the C ABI, one call per channel and direction, "
                                               "PCIe copies included"}
        log(f"[bench] API (host pointers): {M / t_api / 1e6:.3f} Mvis/s")
    if args.precision == "fp32" and not args.no_fp64:
        cfg64 = make_config(args.config, "fp64", args.eps)
        run64 = Runner(cfg64)
        run64.L.check(run64.lib.curig_set_graph_mode(0))
        t64, calls64 = timed_steps(run64, 1, 1)
        v64 = M / t64 / 1e6
        result["fp64_path"] = {"value": round(v64, 4), "ms_per_step": round(t64 * 1e3, 1),
                               "fp32_speedup": round(value / v64, 3), "warmup": 1, "steps": 1,
                               "fft_path": "lds" if calls64["m2d"][-1]["fft_lds"] else "rocfft"}
        if not args.no_err:
            e64i, e64v, img64 = run64.errors(pix, rows)
            result["fp64_path"]["error_vs_fp64_direct_sum"] = {"ms2dirty": e64i, "dirty2ms": e64v}
            if img32 is not None:
                result["fp64_path"]["fp32_vs_fp64_image"] = float(np.linalg.norm(img32 - img64) /
                                                                  np.linalg.norm(img64))
        log(f"[bench] fp64 path: {result['fp64_path']}")
        del run64
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(cfg)
        log(f"[bench] cpu baseline {result['cpu_baseline']}")
        result["c1_cpu_vs_gpu"] = cpu_c1(log)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
