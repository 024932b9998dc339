// Device-resident gridder core: one channel, one direction, all pointers on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "curig_internal.h"

namespace curig {

// Overrides of the fov / sigma defaults (the NIFTY gridder's arguments).
struct GridSpec {
  double dx = 0, dy = 0;  // pixel sizes in radians along x (u) and y (v); 0: fov / nx, fov / ny
  int nu = 0, nv = 0;     // grid size along u and v (even, > image size); 0: from sigma
  bool wstack = true;     // false: w taken as 0 and the image not divided by n
  bool radians = false;   // coordinates are the transform's phases per image index (u, v in
                          // radians; no pixel geometry, freq unused): the plain 2-D NUFFT
  // with radians and wstack: w in radians as well, and each pixel's n - 1 read from
  // zout (device, [nx][ny]) whose values span [zmin, zmax] (a 1-D type-3 transform)
  const double *zout = nullptr;
  double zmin = 0, zmax = 0;
};

struct Problem {
  int nrow = 0, nx = 0, ny = 0;
  double fov = 0, freq = 0, eps = 1e-6, sigma = 2.0;
  int sign = -1;
  bool dbl = true;              // FP64 path, or FP32-mixed (complex64 data, FP32 grid)
  const double *uvw = nullptr;  // [nrow][3] metres, FP64
  void *vis = nullptr;          // complex (double2 / float2); row j at vis[j * vis_stride]
  long vis_stride = 1;
  const void *wgt = nullptr;    // real (double / float) or nullptr; row j at wgt[j * wgt_stride]
  long wgt_stride = 1;
  const uint8_t *flag = nullptr;  // nonzero = row excluded; row j at flag[j * flag_stride]
  long flag_stride = 1;
  void *dirty = nullptr;        // complex [nx][ny]
  bool accumulate = false;      // ms2dirty adds into dirty instead of overwriting it
  GridSpec spec;                // explicit pixel sizes / grid size / w handling (defaults: from fov, sigma)
};

// Filled on request: plan sizes and per-stage device times (ms, hipEvent).
struct Stats {
  int kw = 0, nf1 = 0, nf2 = 0, P = 0, Pb = 0, nbatch = 0, T1 = 0, T2 = 0;
  int fft_lds = 0;             // 1: pruned LDS FFT path, 0: rocFFT
  int rsplit = 0;              // LDS path: parts per grid row in the row kernels (0: whole rows)
  long nvis = 0;               // unflagged rows gridded
  double beta = 0, gamma = 0, C = 0, D = 0;
  double t_prep = 0, t_grid = 0, t_fft = 0, t_plane = 0, t_final = 0, t_total = 0;
  // t_grid: spread (ms2dirty) or interp (dirty2ms); t_plane: w-accumulate / w-fill
  int timed = 0;               // 1 when the times above were measured
};

int ms2dirty_device(const Problem &pb, hipStream_t stream, const Tuning *tune, Stats *st);
int dirty2ms_device(const Problem &pb, hipStream_t stream, const Tuning *tune, Stats *st);

// ms2dirty_device without any device allocation: every buffer is carved from
// a fixed per-device scratch array of noalloc_scratch_bytes() (128 MiB), with
// as many w planes per batch as fit; rejects (kErrScratch) plans off the LDS
// FFT path and working sets beyond the scratch.  Calls share the scratch:
// they are serialised on the host and ordered across streams like the other
// entry points; a caller that races the library's queued work from its own
// kernels on the output races as usual.
int ms2dirty_device_noalloc(const Problem &pb, hipStream_t stream, const Tuning *tune);
size_t noalloc_scratch_bytes();

// Several problems (channels) sharing image size, epsilon and sigma, queued in
// order on one stream.  graph_mode != 0: the whole sequence is captured once into
// a hipGraph and replayed while the same call repeats (LDS FFT path only; other
// grids run the plain per-channel loop).  Stats: totals over the problems.
int multi_device(bool to_image, const Problem *pbs, int n, hipStream_t stream, const Tuning *tune, int graph_mode,
                 Stats *st);

// FP64 direct sum on the GPU (the reference's explicit_gridder, precomp.cu:242-275),
// at npix chosen pixels (pix = i*ny + k) or, with pix == nullptr, at every pixel.
int explicit_ms2dirty_device(const Problem &pb, long npix, const int64_t *pix, void *out,
                             hipStream_t stream);

// FP64 direct sum of the adjoint (dirty2ms) on the GPU at nsel chosen rows:
// out[q] (complex128) = wgt_j sum_pix dirty[pix] / n e^{-i ph(j, pix)}, j = rows[q]
// (the conjugate phase of explicit_ms2dirty_device; flags are not consulted).
int explicit_dirty2ms_device(const Problem &pb, long nsel, const int64_t *rows, void *out, hipStream_t stream);

// Host-only plan computation used by the tests (no device access).
int plan_geometry(int nrow_valid, int nx, int ny, double fov, double freq, double eps, double sigma,
                  int sign, double wmin_m, double wmax_m, bool dbl, const Tuning *tune, Geom *g,
                  KernelOpts *ko, const GridSpec *spec = nullptr);

void release_workspace();

}  // namespace curig
