// ES-kernel improved-w-stacking gridder for CDNA4 (gfx950): host orchestration.
//
// Pipeline (one channel):
//   ms2dirty  prep: w range -> plan; bin rows by (w-plane batch, grid tile of the
//             kernel's first cell); counting sort into 32/56-byte records
//             for each batch of Pb w planes:
//               power-of-two grids (LDS path): spread + row FFT -> column-tiled
//               planes; column FFT fused with the w-accumulation (FP64)
//               other grids: tile-owner spread into whole planes, rocFFT,
//               w-accumulation
//             epilogue: ES / w corrections, / n
//   dirty2ms  prologue: corrections, / n; for each batch: w-fill + inverse
//             column FFT, inverse row FFT fused with the interpolation into
//             FP64 per-row accumulators (other grids: w-fill, rocFFT,
//             interpolation); finalize: weight and e^{-i D w}, scatter to row order.
//
// Mathematically this is the reference's operator (src/astro/ra_exec.cu:141-436):
// same grid sizes, w-plane count and scaling (set_nhg_type3), kernel and
// correction factors; the full nf3 x nf2 x nf1 grid of the reference is never
// materialised (only Pb planes at a time), and the kernel is normalised to peak 1
// (its Fourier-series factors carry the same normalisation, so it cancels).
// The kernels live in kern_prep.hip, kern_plane.hip and kern_lds_*.hip; this
// file queues them through the launchers of kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdlib>

#include "curig_internal.h"
#include "fft.h"
#include "fft_core.h"
#include "gridder.h"
#include "kernels.h"

namespace curig {


// ==========================================================================
// host side
// ==========================================================================
namespace {

// Grow-only device buffers, one set per device.  Calls on one device are
// serialised on the host (g_ws_mu), and a call on another stream than the
// previous call's waits for that call's launches on the device (ws_order).
enum Slot {
  S_PART, S_CNT, S_HIST, S_OFF, S_BSUM, S_BINRANK, S_RECS, S_PERM, S_GRID, S_ACC, S_VACC, S_TAB, S_FFTW,
  S_H, S_TW, S_HC, S_TAPS,
  S_NSLOT
};
struct Workspace {
  void *p[S_NSLOT] = {};
  size_t n[S_NSLOT] = {};
  int tw_key[3] = {-1, -1, -1};  // (nf1, nf2, sizeof(T)) of the twiddles held in S_TW
  // the last call's stream and an event recorded there after its last launch:
  // a call on another stream waits for it before touching the buffers
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
  bool pending = false;
};
std::mutex g_ws_mu;
std::vector<Workspace> g_ws;

// Fixed device scratch of the entry point that allocates nothing
// (ms2dirty_device_noalloc): a file-scope array every device holds once the
// library runs there, carved into the call's buffers instead of the workspace.
// In host code `scratch` is only a stand-in: its device address comes from
// hipGetSymbolAddress, once per device (scratch_base).
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
};
Arena *g_arena = nullptr;        // set while a no-allocation call runs (g_ws_mu held)
std::vector<char *> g_scratch;   // per device: the scratch array's device address

int scratch_base(int dev, char **out) {
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1, nullptr);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    CURIG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
      set_error("no device address for the scratch array");
      return kErrHip;
    }
    g_scratch[dev] = (char *)ptr;
  }
  *out = g_scratch[dev];
  return kOk;
}

int ws_get(int dev, Slot s, size_t bytes, void **out) {
  if (g_arena) {
    // no-allocation call: consecutive 256-byte aligned pieces of the scratch
    const size_t a = ((bytes ? bytes : 16) + 255) & ~size_t(255);
    if (g_arena->used + a > g_arena->cap) {
      set_error("working set exceeds the %zu MiB fixed scratch (needs %zu more bytes at buffer %d)",
                g_arena->cap >> 20, g_arena->used + a - g_arena->cap, (int)s);
      return kErrScratch;
    }
    *out = g_arena->base + g_arena->used;
    g_arena->used += a;
    return kOk;
  }
  if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
  Workspace &w = g_ws[dev];
  if (bytes == 0) bytes = 16;
  if (w.n[s] < bytes) {
    if (w.p[s]) {
      CURIG_HIP(hipFree(w.p[s]));
      w.p[s] = nullptr;
      w.n[s] = 0;
    }
    hipError_t e = hipMalloc(&w.p[s], bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      return kErrAlloc;
    }
    w.n[s] = bytes;
    if (s == S_TW) w.tw_key[0] = w.tw_key[1] = w.tw_key[2] = -1;
  }
  *out = w.p[s];
  return kOk;
}

// Orders a call on stream st after the previous call on this device's
// workspace when that call ran on another stream (its kernels may still be
// reading the buffers this call is about to overwrite).
int ws_order(int dev, hipStream_t st) {
  if ((int)g_ws.size() <= dev) g_ws.resize(dev + 1);
  Workspace &w = g_ws[dev];
  if (w.pending && w.last != st) CURIG_HIP(hipStreamWaitEvent(st, w.done, 0));
  return kOk;
}

// Marks the end of a call's launches on stream st.
int ws_mark(int dev, hipStream_t st) {
  Workspace &w = g_ws[dev];
  if (!w.done) CURIG_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
  CURIG_HIP(hipEventRecord(w.done, st));
  w.last = st;
  w.pending = true;
  return kOk;
}



struct Timer {
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
  int mark() {
    if (!on) return -1;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    (void)hipEventRecord(e, s);
    ev.push_back(e);
    return (int)ev.size() - 1;
  }
  double ms(int a, int b) {
    if (!on || a < 0 || b < 0) return 0;
    float t = 0;
    (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return t;
  }
  // the time between marks a and b, added to *sum once the stream has been
  // synchronised (settle): no host wait inside the launch sequence
  struct Span {
    int a, b;
    double *sum;
  };
  std::vector<Span> spans;
  void span(int a, int b, double *sum) {
    if (on) spans.push_back({a, b, sum});
  }
  void settle() {
    for (const Span &p : spans) *p.sum += ms(p.a, p.b);
    spans.clear();
  }
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// Per-stage tracing, switched on by environment variables read once:
//   CURIG_TRACE=1   roctx ranges around each call and each of its stages
//                   (binning, every plane batch, epilogue), seen by
//                   rocprofv3 --marker-trace;
//   CURIG_TIMING=1  every call times its stages with HIP events (as the
//                   stats API does) and prints them to stderr.
struct TraceFlags {
  bool ranges = false, timing = false;
  TraceFlags() {
    const char *t = std::getenv("CURIG_TRACE");
    const char *m = std::getenv("CURIG_TIMING");
    ranges = t && *t && *t != '0';
    timing = m && *m && *m != '0';
  }
};
const TraceFlags &trace_flags() {
  static const TraceFlags f;
  return f;
}

// roctx range for the enclosing scope (no-op unless CURIG_TRACE is set)
struct Range {
  bool on;
  explicit Range(const char *name) : on(trace_flags().ranges) {
    if (on) roctxRangePushA(name);
  }
  ~Range() {
    if (on) roctxRangePop();
  }
};

// Pixel sizes (rad): dx = fov / nx pairs with u, dy = fov / ny with v
// (src/astro/cugridder.cu:48-49), unless the caller gives both.
void pixel_sizes(int nx, int ny, double fov, const GridSpec *spec, double &dx, double &dy) {
  if (spec && spec->radians) {  // no sky geometry: l = m = 0, n = 1
    dx = dy = 0.0;
    return;
  }
  if (spec && (spec->dx != 0 || spec->dy != 0)) {
    dx = spec->dx;
    dy = spec->dy;
    return;
  }
  dx = fov / 180.0 * M_PI / (double)nx;
  dy = fov / 180.0 * M_PI / (double)ny;
}

int check_args(const Problem &pb) {
  if (pb.nrow < 0) {
    set_error("nrow %d < 0", pb.nrow);
    return kErrNrow;
  }
  if (pb.nx <= 0 || pb.ny <= 0) {
    set_error("image size %d x %d invalid", pb.nx, pb.ny);
    return kErrSize;
  }
  if (pb.sign != 1 && pb.sign != -1) {
    set_error("sign must be +1 or -1 (got %d)", pb.sign);
    return kErrArg;
  }
  const bool own_pix = pb.spec.dx != 0 || pb.spec.dy != 0;
  if (own_pix && !(std::isfinite(pb.spec.dx) && pb.spec.dx > 0 && std::isfinite(pb.spec.dy) && pb.spec.dy > 0)) {
    set_error("pixel sizes (%g, %g) must be positive and finite", pb.spec.dx, pb.spec.dy);
    return kErrNonFinite;
  }
  if (!pb.spec.radians &&
      ((!own_pix && !(std::isfinite(pb.fov) && pb.fov > 0)) || !(std::isfinite(pb.freq) && pb.freq > 0))) {
    set_error("fov (%g) and freq (%g) must be positive and finite", pb.fov, pb.freq);
    return kErrNonFinite;
  }
  if (!pb.dirty || (pb.nrow > 0 && (!pb.uvw || !pb.vis))) {
    set_error("NULL data pointer");
    return kErrArg;
  }
  double dx, dy;
  pixel_sizes(pb.nx, pb.ny, pb.fov, &pb.spec, dx, dy);
  double lmax = 0.5 * dx * pb.nx, mmax = 0.5 * dy * pb.ny;
  if (!(lmax * lmax + mmax * mmax < 1.0)) {
    set_error("field of view %g x %g rad reaches beyond the horizon (l^2 + m^2 >= 1)", dx * pb.nx, dy * pb.ny);
    return kErrArg;
  }
  return kOk;
}

}  // namespace

// FP32 columns of 16384 run the register-resident column kernels as FP64 does
// (1: yes; 0: the LDS column kernels)
#ifndef CURIG_COLREG_FP32
#define CURIG_COLREG_FP32 1
#endif
inline bool col_reg_fp32() { return CURIG_COLREG_FP32 != 0; }

int plan_geometry(int nrow_valid, int nx, int ny, double fov, double freq, double eps, double sigma, int sign,
                  double wmin_m, double wmax_m, bool dbl, const Tuning *tune, Geom *gp, KernelOpts *kop,
                  const GridSpec *spec) {
  const GridSpec sp = spec ? *spec : GridSpec();
  const bool own_grid = sp.nu != 0 || sp.nv != 0;
  if (own_grid) {
    // a given grid size fixes the upsampling factor: the smaller of the two axes'
    if (sp.nu <= nx || sp.nv <= ny || (sp.nu & 1) || (sp.nv & 1)) {
      set_error("grid %d x %d must be even and larger than the image %d x %d", sp.nu, sp.nv, nx, ny);
      return kErrSize;
    }
    sigma = std::min((double)sp.nu / nx, (double)sp.nv / ny);
  }
  KernelOpts ko;
  int ier = setup_conv_opts(ko, eps, sigma);
  if (ier != kOk) return ier;
  Geom g{};
  g.N1 = ny;
  g.N2 = nx;
  g.kw = ko.kw;
  g.beta = ko.beta;
  g.c = ko.c;
  // src/astro/cugridder.cu:48-49 (N2 = nx pairs with u, N1 = ny with v)
  pixel_sizes(nx, ny, fov, &sp, g.dx, g.dy);
  double fc = freq / kSpeedOfLight;
  g.uscale = sp.radians ? 1.0 : fc * 2 * M_PI * g.dx;
  g.vscale = sp.radians ? 1.0 : fc * 2 * M_PI * g.dy;
  // 2-D form: every w is taken as 0 (w range 0, C = 0) and the 1 / n dropped
  g.divn = sp.wstack ? 1 : 0;
  g.wscale = sp.wstack ? (sp.radians ? 1.0 : fc * sign * 2 * M_PI) : 0.0;
  g.zout = sp.zout;
  // type-3 parameters (cugridder.cu:52-76)
  double l_min = -0.5 * g.dx * nx, m_min = -0.5 * g.dy * ny;
  double n_lm = std::sqrt(1.0 - l_min * l_min - m_min * m_min);
  double we0 = wmin_m * g.wscale, we1 = wmax_m * g.wscale;
  double i_max = nrow_valid > 0 ? std::max(we0, we1) : 0.0;
  double i_min = nrow_valid > 0 ? std::min(we0, we1) : 0.0;
  g.C = (i_max + i_min) / 2.0;
  double X = (i_max - i_min) / 2.0;
  double o_min = n_lm - 1;
  g.D = o_min / 2.0;
  double S = std::fabs(o_min / 2.0);
  if (sp.zout) {  // given outputs: centre and half-width of their range
    g.D = 0.5 * (sp.zmin + sp.zmax);
    S = 0.5 * (sp.zmax - sp.zmin);
  }
  set_nhg_type3(S, X, ko, g.P, g.h, g.gamma);
  g.kww = g.kw;
  if (!sp.wstack) {
    // 2-D form: one plane holding every row with weight 1 (w-DFT phase and
    // phihat3 are 1), so the transform is the plain 2-D NUFFT
    g.P = 1;
    g.kww = 1;
    g.h = 2 * M_PI;
    g.gamma = 1.0;
  }
  if (g.P > 10000000) {
    set_error("w range needs %d w planes (more than 1e7)", g.P);
    return kErrSize;
  }
  g.nf1 = own_grid ? sp.nv : get_num_cells(ny, ko);
  g.nf2 = own_grid ? sp.nu : get_num_cells(nx, ko);
  if (g.nf1 < 2 * ko.kw || g.nf2 < 2 * ko.kw) {
    set_error("grid %d x %d smaller than twice the kernel width %d", g.nf2, g.nf1, ko.kw);
    return kErrSize;
  }
  if ((double)g.nf1 * g.nf2 > 2.0e9) {
    set_error("grid %d x %d too large", g.nf1, g.nf2);
    return kErrSize;
  }
  Tuning tn = tune ? *tune : Tuning();
  // kernel evaluation: exp(sqrt()), or the piecewise polynomial (coefficients
  // uploaded with the other tables when the buffers are reserved)
  g.hdeg = -1;
  g.hcoef_d = nullptr;
  g.hcoef_f = nullptr;
  if (tn.kereval == 1) {
    std::vector<double> coef;
    g.hdeg = horner_fit(ko, horner_tol(eps, dbl), coef);
    if (g.hdeg < 0) {
      set_error("no polynomial of degree <= %d fits the kernel (kw %d) to %g", kMaxHornerDeg, ko.kw,
                horner_tol(eps, dbl));
      return kErrArg;
    }
  }
  g.Pbin = std::max(1, std::min(tn.Pbin > 0 ? tn.Pbin : 4, g.P));
  g.nlayer = (g.P + g.Pbin - 1) / g.Pbin;
  // pruned LDS FFT: power-of-two grids whose rows and columns fit one
  // workgroup's LDS, with the image at most half the grid on each axis
  g.fft_lds = 0;
  g.rpw = g.cpw = g.nct = g.tstride = g.nslot = g.slot_lg = g.lds_row_bytes = g.lds_col_bytes = g.rsplit =
      g.lds_rowp_bytes = g.colreg = g.cpw_inv = g.lds_coli_bytes = 0;
  auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
  const size_t cs = dbl ? 16 : 8;
  const size_t lds_max = tn.lds_bytes > 0 ? (size_t)tn.lds_bytes : size_t(160) << 10;
  const int max_threads = dbl ? 512 : 1024;
  if (tn.fft != 1 && g.kw <= 16 && pow2(g.nf1) && pow2(g.nf2) && g.nf1 >= 64 && g.nf2 >= 64 && g.nf1 <= 16384 &&
      g.nf2 <= 16384 && 2 * g.N1 <= g.nf1 && 2 * g.N2 <= g.nf2) {
    const int rpw = std::max(1, std::min(4096 / g.nf1, g.nf2));
    const int cpw = std::max(1, std::min(4096 / g.nf2, 16));
    // the row kernels' static LDS (bin walk table) comes on top of the row
    const size_t rb = (size_t)(fftc::padded_size(rpw * g.nf1) + fftc::twiddle_entries(g.nf1)) * cs;
    const size_t cb = (size_t)(fftc::padded_size(cpw * g.nf2) + fftc::twiddle_entries(g.nf2)) * cs;
    // a row part of a row split kRowSplit ways (256 threads at most)
    const size_t rbp = (size_t)(fftc::padded_size(g.nf1 / kRowSplit) + fftc::twiddle_entries(g.nf1)) * cs;
    const bool row_whole = rb + 2048 <= lds_max && rpw * g.nf1 / 16 <= max_threads;
    const bool row_split = rpw == 1 && g.nf1 >= 4096 && g.nf1 / kRowSplit <= 16 * 256 && rbp + 2048 <= lds_max / 2;
    const bool col_whole = cb <= lds_max && cpw * g.nf2 / 16 <= max_threads;
    // columns of 16384 held in registers (FP64: exchanged through LDS one
    // component at a time; FP32: the first and last stages in registers)
    const size_t cbr = colreg_lds_bytes(cs);
    const bool col_reg = g.nf2 == 16384 && cbr <= lds_max && (dbl || col_reg_fp32());
    if ((row_whole || row_split) && (col_whole || col_reg) && rpw * g.nf1 >= 16 * 64 && cpw * g.nf2 >= 16) {
      g.fft_lds = 1;
      g.rpw = rpw;
      g.colreg = col_reg ? 1 : 0;
      g.cpw = col_reg ? 1 : cpw;
      g.nslot = g.N1;
      g.lds_row_bytes = row_whole ? (int)rb : 0;
      g.lds_col_bytes = (int)(col_reg ? cbr : cb);
      // the dirty2ms column pass writes H: FP32 workgroups of fewer than
      // kTile columns leave each 32-byte chunk to several workgroups, which
      // reach a plane at different times, and the L2 then writes the chunk's
      // lines back more than once (2.4x the bytes at 4096^2); with whole tiles
      // a workgroup writes whole chunks (1.0x)
      g.cpw_inv = g.cpw;
      g.lds_coli_bytes = g.lds_col_bytes;
      for (int c = kTile; !dbl && !col_reg && cpw < c && c >= 2; c /= 2) {
        // whole tiles, else half tiles (16-byte pieces of a chunk)
        const size_t cbi = (size_t)(fftc::padded_size(c * g.nf2) + fftc::twiddle_entries(g.nf2)) * cs;
        if (cbi <= lds_max && c * g.nf2 / 16 <= max_threads) {
          g.cpw_inv = c;
          g.lds_coli_bytes = (int)cbi;
          break;
        }
      }
      // rows too long for two workgroups per CU run split kRowSplit ways
      if (!row_whole || (rpw == 1 && rb + 2048 > lds_max / 2 && g.nf1 >= 4096)) {
        g.rsplit = kRowSplit;
        g.nslot = g.nf1 / 2;
        g.slot_lg = 31 - __builtin_clz(g.nf1 / (2 * kRowSplit));
        g.lds_rowp_bytes = (int)rbp;
      }
      g.nct = (g.nslot + kTile - 1) / kTile;
      g.tstride = g.nf2 * kTile + (int)(kTilePadBytes / cs);
    }
  }
  // planes per batch from the memory budget: full planes (rocFFT path) or, on
  // the LDS path, the compact plane plus the one or two column-pass buffers
  size_t per_plane = (size_t)g.nf1 * g.nf2 * cs;
  if (g.fft_lds) per_plane = (size_t)g.nct * g.tstride * cs;
  int Pb = tn.Pb;
  if (Pb <= 0) {
    Pb = (int)std::max<size_t>(1, tn.grid_budget_bytes / per_plane);
    Pb = std::min(Pb, 32);
  }
  g.Pb = std::max(1, std::min(Pb, g.P));
  g.nbatch = (g.P + g.Pb - 1) / g.Pb;
  // bins: full-width row strips for the LDS path (one workgroup owns grid rows),
  // T1 x T2 tiles for the tile-owner spread
  if (g.fft_lds) {
    g.T1 = g.nf1;
    g.T2 = std::min(tn.T2 > 0 ? tn.T2 : 2, g.nf2);
  } else {
    g.T1 = std::min(tn.T1 > 0 ? tn.T1 : (dbl ? 32 : 64), g.nf1);
    g.T2 = std::min(tn.T2 > 0 ? tn.T2 : 32, g.nf2);
  }
  g.ntx = (g.nf1 + g.T1 - 1) / g.T1;
  g.nty = (g.nf2 + g.T2 - 1) / g.T2;
  if (g.fft_lds) {
    // a row workgroup walks at most kMaxWalk bins: coarsen strips, then layers
    while (walk_bins_bound(g.kww, g.kw, g.Pbin, g.nlayer, g.rpw, g.T2, g.nty) > kMaxWalk) {
      if (g.T2 < g.nf2) {
        g.T2 = std::min(2 * g.T2, g.nf2);
        g.nty = (g.nf2 + g.T2 - 1) / g.T2;
      } else {
        g.Pbin = std::min(2 * g.Pbin, g.P);
        g.nlayer = (g.P + g.Pbin - 1) / g.Pbin;
      }
    }
  }
  g.ntiles = g.ntx * g.nty;
  if ((double)g.nlayer * g.ntiles > 1.5e9) {
    set_error("too many bins (%d layers x %d tiles)", g.nlayer, g.ntiles);
    return kErrSize;
  }
  double qx[kMaxQuadNodes], qg[kMaxQuadNodes];
  g.nquad = kernel_quadrature(ko, qx, qg);
  *gp = g;
  if (kop) *kop = ko;
  return kOk;
}

namespace {

int device_lds_bytes() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 65536;
    cached = v;
  }
  return cached;
}


// Phase 1 (one wait): w range and unflagged count of every problem, then their geometries.
int plan_all(const Problem *pbs, int n, hipStream_t st, const Tuning *tune, int dev, std::vector<ChanPlan> &cps) {
  std::vector<int> nbm(n), base(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    nbm[i] = pbs[i].nrow > 0 ? std::min(nblocks(pbs[i].nrow), 1024) : 0;
    base[i + 1] = base[i] + nbm[i];
  }
  void *vp;
  CURIG_TRY(ws_get(dev, S_PART, sizeof(double) * 2 * std::max(base[n], 1), &vp));
  double *part = (double *)vp;
  CURIG_TRY(ws_get(dev, S_CNT, sizeof(int) * std::max(base[n], 1), &vp));
  int *cnt = (int *)vp;
  for (int i = 0; i < n; ++i)
    if (nbm[i] > 0)
      launch_w_minmax(nbm[i], pbs[i].uvw, pbs[i].nrow, pbs[i].flag, pbs[i].flag_stride, part + 2 * base[i],
                      cnt + base[i], st);
  CURIG_HIP(hipGetLastError());
  std::vector<double> hp(2 * std::max(base[n], 1));
  std::vector<int> hc(std::max(base[n], 1));
  if (base[n] > 0) {
    CURIG_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * base[n], hipMemcpyDeviceToHost, st));
    CURIG_HIP(hipMemcpyAsync(hc.data(), cnt, sizeof(int) * base[n], hipMemcpyDeviceToHost, st));
  }
  CURIG_HIP(hipStreamSynchronize(st));
  Tuning tn = tune ? *tune : Tuning();
  if (tn.lds_bytes <= 0) tn.lds_bytes = device_lds_bytes();
  cps.assign(n, ChanPlan());
  for (int i = 0; i < n; ++i) {
    double wmin = INFINITY, wmax = -INFINITY;
    long nvalid = 0;
    bool bad = false;
    for (int b = base[i]; b < base[i + 1]; ++b) {
      nvalid += hc[b];
      if (hc[b] == 0) continue;
      if (std::isnan(hp[2 * b]) || std::isnan(hp[2 * b + 1])) bad = true;
      wmin = std::min(wmin, hp[2 * b]);
      wmax = std::max(wmax, hp[2 * b + 1]);
    }
    if (bad) {
      set_error("non-finite uvw coordinates");
      return kErrNonFinite;
    }
    if (nvalid == 0) wmin = wmax = 0;
    const Problem &pb = pbs[i];
    CURIG_TRY(plan_geometry((int)nvalid, pb.nx, pb.ny, pb.fov, pb.freq, pb.eps, pb.sigma, pb.sign, wmin, wmax,
                            pb.dbl, &tn, &cps[i].g, &cps[i].ko, &pb.spec));
    cps[i].nvalid = nvalid;
  }
  return kOk;
}

// Two-level twiddle tables (fftc::TwoLevel: hi | lo) for n = nf1 and n = nf2,
// FP64 values rounded to T, uploaded when the sizes change.
template <typename T>
int twiddles(int dev, const Geom &g, hipStream_t st, const cpx<T> *&tw1, const cpx<T> *&tw2) {
  void *vp;
  const int n1 = fftc::twiddle_entries(g.nf1), n2 = fftc::twiddle_entries(g.nf2);
  const size_t bytes = sizeof(cpx<T>) * (size_t)(n1 + n2);
  CURIG_TRY(ws_get(dev, S_TW, bytes, &vp));
  cpx<T> *t = (cpx<T> *)vp;
  int *key = g_ws[dev].tw_key;
  // the scratch of a no-allocation call holds whatever the last call left there
  if (g_arena || key[0] != g.nf1 || key[1] != g.nf2 || key[2] != (int)sizeof(T)) {
    std::vector<cpx<T>> h(n1 + n2);
    auto fill = [](cpx<T> *o, int n) {
      const int lb = fftc::twiddle_split(n), nhi = n >> lb, nlo = 1 << lb;
      for (int j = 0; j < nhi; ++j) {
        const double a = 2 * M_PI * ((double)j * nlo) / n;
        o[j] = fftc::mk((T)std::cos(a), (T)std::sin(a));
      }
      for (int j = 0; j < nlo; ++j) o[nhi + j] = fftc::mk((T)std::cos(2 * M_PI * j / n), (T)std::sin(2 * M_PI * j / n));
    };
    fill(h.data(), g.nf1);
    fill(h.data() + n1, g.nf2);
    CURIG_HIP(hipMemcpyAsync(t, h.data(), bytes, hipMemcpyHostToDevice, st));
    CURIG_HIP(hipStreamSynchronize(st));
    if (!g_arena) {
      key[0] = g.nf1;
      key[1] = g.nf2;
      key[2] = (int)sizeof(T);
    }
  }
  tw1 = t;
  tw2 = t + n1;
  return kOk;
}

// FFT for a batch of nb planes, at most kMaxFftElems elements per rocFFT call
// (the reference caps cuFFT batches the same way, MAX_CUFFT_ELEM, cugridder.cu:230-246;
// a single call over 2^32 elements returns wrong planes).
constexpr size_t kMaxFftElems = size_t(1) << 28;

template <typename T>
int batch_fft(int dev, const Geom &g, int nb, int sign, void *grid, hipStream_t st) {
  const size_t plane = (size_t)g.nf1 * g.nf2;
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nb, kMaxFftElems / plane));
  for (int k0 = 0; k0 < nb; k0 += chunk) {
    int n = std::min(chunk, nb - k0);
    FftPlan fp;
    CURIG_TRY(fft_plan(dev, g.nf1, g.nf2, n, sizeof(T) == 8, sign, &fp));
    void *work = nullptr;
    if (fp.work_bytes) CURIG_TRY(ws_get(dev, S_FFTW, fp.work_bytes, &work));
    CURIG_TRY(fft_exec(fp, (v2_t<T> *)grid + (size_t)k0 * plane, work, st));
  }
  return kOk;
}

// Phase 2 (one wait): allocate the working set for all problems, upload the
// correction tables and twiddles, create FFT plans and kernel attributes, so
// that phase 3 queues kernels only.
template <typename T>
int reserve(const std::vector<ChanPlan> &cps, const Problem *pbs, bool to_image, int dev, hipStream_t st, Bufs &b) {
  const Geom &g0 = cps[0].g;
  size_t nbins = 1, nscan = 1, nrow = 1, grid = 1, h = 1, npix = 1, nvalid = 1;
  for (size_t i = 0; i < cps.size(); ++i) {
    const Geom &g = cps[i].g;
    if (g.kw != g0.kw || g.nf1 != g0.nf1 || g.nf2 != g0.nf2 || g.N1 != g0.N1 || g.N2 != g0.N2) {
      set_error("channels of one call must share image size, epsilon and sigma");
      return kErrArg;
    }
    size_t nb = (size_t)g.nlayer * g.ntiles + 1;
    nbins = std::max(nbins, nb);
    nscan = std::max(nscan, (nb + kScanTile - 1) / kScanTile + 1);
    nrow = std::max(nrow, (size_t)std::max(pbs[i].nrow, 1));
    nvalid = std::max(nvalid, (size_t)std::max<long>(cps[i].nvalid, 1));
    const size_t plane = (size_t)g.nf1 * g.nf2, pix = (size_t)g.N1 * g.N2;
    if (!g.fft_lds) grid = std::max(grid, plane * g.Pb);
    if (g.fft_lds) h = std::max(h, (size_t)g.nct * g.tstride * g.Pb);
    npix = std::max(npix, pix);
  }
  void *vp;
  CURIG_TRY(ws_get(dev, S_HIST, sizeof(int) * nbins, &vp));
  b.hist = (int *)vp;
  CURIG_TRY(ws_get(dev, S_OFF, sizeof(int) * nbins, &vp));
  b.off = (int *)vp;
  CURIG_TRY(ws_get(dev, S_BSUM, sizeof(int) * nscan, &vp));
  b.bsum = (int *)vp;
  CURIG_TRY(ws_get(dev, S_BINRANK, sizeof(int2) * nrow, &vp));
  b.binrank = (int2 *)vp;
  CURIG_TRY(ws_get(dev, S_RECS, sizeof(Rec<T>) * nrow, &vp));
  b.recs = vp;
  CURIG_TRY(ws_get(dev, S_PERM, sizeof(int) * nrow, &vp));
  b.perm = (int *)vp;
  CURIG_TRY(ws_get(dev, S_GRID, sizeof(v2_t<T>) * grid, &vp));
  b.grid = vp;
  CURIG_TRY(ws_get(dev, S_ACC, sizeof(double2) * npix, &vp));
  b.acc = (double2 *)vp;
  if (!to_image) {
    CURIG_TRY(ws_get(dev, S_VACC, sizeof(double2) * nvalid, &vp));
    b.vacc = (double2 *)vp;
  }
  if (g0.hdeg >= 0) {
    // same fit as the plan's (same options and tolerance), FP64 then FP32 copies
    std::vector<double> coef;
    const int deg = horner_fit(cps[0].ko, horner_tol(pbs[0].eps, pbs[0].dbl), coef);
    if (deg != g0.hdeg) {
      set_error("kernel polynomial degree changed between plan and upload");
      return kErrArg;
    }
    // device layout [deg + 1][kw]: one power of every tap contiguous
    const size_t nc = coef.size();
    const int kw = cps[0].ko.kw, np1 = deg + 1;
    std::vector<unsigned char> host(nc * (sizeof(double) + sizeof(float)));
    double *cd = (double *)host.data();
    float *cf = (float *)(host.data() + nc * sizeof(double));
    for (int i = 0; i < kw; ++i)
      for (int k = 0; k < np1; ++k) {
        cd[(size_t)k * kw + i] = coef[(size_t)i * np1 + k];
        cf[(size_t)k * kw + i] = (float)coef[(size_t)i * np1 + k];
      }
    CURIG_TRY(ws_get(dev, S_HC, host.size(), &vp));
    CURIG_HIP(hipMemcpyAsync(vp, host.data(), host.size(), hipMemcpyHostToDevice, st));
    CURIG_HIP(hipStreamSynchronize(st));
    b.hc_d = (const double *)vp;
    b.hc_f = (const float *)((const char *)vp + nc * sizeof(double));
  }
  if (g0.fft_lds) {
    CURIG_TRY(ws_get(dev, S_H, sizeof(cpx<T>) * h, &vp));
    b.H = vp;
    CURIG_TRY(ws_get(dev, S_TAPS, sizeof(T) * 3 * (size_t)tap_width(g0.kw) * nrow, &vp));
    b.taps = vp;
    const cpx<T> *t1, *t2;
    CURIG_TRY(twiddles<T>(dev, g0, st, t1, t2));
    b.tw1 = t1;
    b.tw2 = t2;
    CURIG_TRY(allow_spread_rows<T>(g0));
    CURIG_TRY(allow_row_interp<T>(g0));
    CURIG_TRY(allow_cols<T>(g0));
  } else {
    // create every FFT plan the batches will use (full and tail batches, split
    // into chunks of at most kMaxFftElems elements) and size the work buffer
    for (const ChanPlan &cp : cps) {
      const Geom &g = cp.g;
      const size_t plane = (size_t)g.nf1 * g.nf2;
      const int chunk = (int)std::max<size_t>(1, std::min<size_t>(g.Pb, kMaxFftElems / plane));
      const int tail = g.P % g.Pb;
      const int sizes[4] = {std::min(chunk, g.Pb), g.Pb % chunk, std::min(chunk, tail), tail % chunk};
      for (int k = 0; k < 4; ++k) {
        if (sizes[k] <= 0) continue;
        FftPlan fp;
        CURIG_TRY(fft_plan(dev, g.nf1, g.nf2, sizes[k], sizeof(T) == 8, to_image ? +1 : -1, &fp));
        if (fp.work_bytes) CURIG_TRY(ws_get(dev, S_FFTW, fp.work_bytes, &vp));
      }
    }
  }
  // correction tables (shared by all channels: same kernel and grid sizes)
  const KernelOpts &ko = cps[0].ko;
  size_t n1 = g0.nf1 / 2 + 1, n2 = g0.nf2 / 2 + 1;
  std::vector<double> ht(n1 + n2 + 2 * kMaxQuadNodes);
  fseries_grid(ko, g0.nf1, ht.data());
  fseries_grid(ko, g0.nf2, ht.data() + n1);
  kernel_quadrature(ko, ht.data() + n1 + n2, ht.data() + n1 + n2 + g0.nquad);
  CURIG_TRY(ws_get(dev, S_TAB, sizeof(double) * ht.size(), &vp));
  b.tabs = (double *)vp;
  CURIG_HIP(hipMemcpyAsync(b.tabs, ht.data(), sizeof(double) * ht.size(), hipMemcpyHostToDevice, st));
  CURIG_HIP(hipStreamSynchronize(st));
  b.dev = dev;
  return kOk;
}


void fill_stats(Stats *sts, const ChanPlan &cp) {
  const Geom &g = cp.g;
  sts->kw = g.kw;
  sts->nf1 = g.nf1;
  sts->nf2 = g.nf2;
  sts->P = g.P;
  sts->Pb = g.Pb;
  sts->nbatch = g.nbatch;
  sts->T1 = g.T1;
  sts->T2 = g.T2;
  sts->fft_lds = g.fft_lds;
  sts->rsplit = g.rsplit;
  sts->nvis = cp.nvalid;
  sts->beta = g.beta;
  sts->gamma = g.gamma;
  sts->C = g.C;
  sts->D = g.D;
}

// Phase 3: queue one ms2dirty problem (no waits and no allocation; with stats,
// HIP events between the launches and one wait at the end).
template <typename T>
int enqueue_ms2dirty(const Problem &pb, const ChanPlan &cp, const Bufs &b, hipStream_t st, Stats *sts) {
  const Geom &g = cp.g;
  Timer tm;
  tm.on = sts != nullptr;
  tm.s = st;
  int e0 = tm.mark();
  {
    Range r("curig ms2dirty binning");
    enqueue_binning<T>(pb, cp, b, st, true);
  }
  CURIG_HIP(hipGetLastError());
  int e1 = tm.mark();
  const Rec<T> *recs = (const Rec<T> *)b.recs;
  v2_t<T> *grid = (v2_t<T> *)b.grid;
  const double *fk1 = b.tabs, *fk2 = b.tabs + (g.nf1 / 2 + 1), *q = fk2 + (g.nf2 / 2 + 1);
  double t_spread = 0, t_fft = 0, t_wacc = 0;
  for (int bi = 0; bi < g.nbatch; ++bi) {
    const int p0 = bi * g.Pb, nb = std::min(g.Pb, g.P - p0);
    Range r("curig ms2dirty plane batch");
    const int a = tm.mark();
    if (g.fft_lds) {
      // spread + row FFT counts as spread (t_grid); the column FFT fused with
      // the w-accumulation as FFT (t_fft)
      lds_spread_rows<T>(recs, (const T *)b.taps, b.off, g, p0, nb, (cpx<T> *)b.H, (const cpx<T> *)b.tw1, st);
      CURIG_HIP(hipGetLastError());
      const int bm = tm.mark();
      lds_cols_forward<T>(g, p0, nb, (const cpx<T> *)b.H, b.acc, (const cpx<T> *)b.tw2, bi == 0 ? 1 : 0, st);
      CURIG_HIP(hipGetLastError());
      const int c = tm.mark();
      tm.span(a, bm, &t_spread);
      tm.span(bm, c, &t_fft);
      continue;
    }
    spread_dispatch<T>(recs, b.off, grid, g, p0, nb, st);
    const int bm = tm.mark();
    CURIG_TRY(batch_fft<T>(b.dev, g, nb, +1, grid, st));
    CURIG_HIP(hipGetLastError());
    const int c = tm.mark();
    launch_wacc<T>(grid, b.acc, g, p0, nb, bi == 0 ? 1 : 0, st);
    CURIG_HIP(hipGetLastError());
    const int d = tm.mark();
    tm.span(a, bm, &t_spread);
    tm.span(bm, c, &t_fft);
    tm.span(c, d, &t_wacc);
  }
  int e2 = tm.mark();
  Range r("curig ms2dirty epilogue");
  launch_epilogue(pb.dbl, b.acc, pb.dirty, g, fk1, fk2, q, pb.accumulate ? 1 : 0, g.fft_lds, st);
  CURIG_HIP(hipGetLastError());
  int e3 = tm.mark();
  if (sts) {
    CURIG_HIP(hipStreamSynchronize(st));
    tm.settle();
    fill_stats(sts, cp);
    sts->t_prep = tm.ms(e0, e1);
    sts->t_grid = t_spread;
    sts->t_fft = t_fft;
    sts->t_plane = t_wacc;
    sts->t_final = tm.ms(e2, e3);
    sts->t_total = tm.ms(e0, e3);
    sts->timed = 1;
  }
  return kOk;
}

template <typename T>
int enqueue_dirty2ms(const Problem &pb, const ChanPlan &cp, const Bufs &b, hipStream_t st, Stats *sts) {
  const Geom &g = cp.g;
  Timer tm;
  tm.on = sts != nullptr;
  tm.s = st;
  int e0 = tm.mark();
  Range rp("curig dirty2ms binning and prologue");
  enqueue_binning<T>(pb, cp, b, st, false);
  // every output row is written: flagged rows get 0
  launch_zero_rows(pb.dbl, pb.vis, pb.vis_stride, pb.nrow, st);
  const Rec<T> *recs = (const Rec<T> *)b.recs;
  v2_t<T> *grid = (v2_t<T> *)b.grid;
  double2 *gimg = b.acc;
  const double *fk1 = b.tabs, *fk2 = b.tabs + (g.nf1 / 2 + 1), *q = fk2 + (g.nf2 / 2 + 1);
  if (g.fft_lds)
    launch_prologue_cols<T>(pb.dbl, pb.dirty, (cpx<T> *)gimg, g, fk1, fk2, q, st);
  else
    launch_prologue(pb.dbl, pb.dirty, gimg, g, fk1, fk2, q, st);
  launch_fill_zero_c128(b.vacc, cp.nvalid, st);
  CURIG_HIP(hipGetLastError());
  int e1 = tm.mark();
  double t_interp = 0, t_fft = 0, t_fill = 0;
  for (int bi = 0; bi < g.nbatch; ++bi) {
    const int p0 = bi * g.Pb, nb = std::min(g.Pb, g.P - p0);
    Range r("curig dirty2ms plane batch");
    const int a = tm.mark();
    if (g.fft_lds) {
      // w-fill + column iFFT counts as FFT (t_fft); the row iFFT fused with
      // the interpolation as interpolation (t_grid)
      lds_cols_backward<T>(g, p0, nb, (const cpx<T> *)gimg, (cpx<T> *)b.H, (const cpx<T> *)b.tw2, st);
      CURIG_HIP(hipGetLastError());
      const int c = tm.mark();
      if (cp.nvalid > 0)
        lds_row_interp<T>(recs, (const T *)b.taps, b.off, (const cpx<T> *)b.H, b.vacc, (const cpx<T> *)b.tw1, g, p0,
                          nb, st);
      CURIG_HIP(hipGetLastError());
      const int d = tm.mark();
      tm.span(a, c, &t_fft);
      tm.span(c, d, &t_interp);
      continue;
    }
    launch_wfill<T>(gimg, grid, g, p0, nb, st);
    CURIG_HIP(hipGetLastError());
    const int bm = tm.mark();
    CURIG_TRY(batch_fft<T>(b.dev, g, nb, -1, grid, st));
    CURIG_HIP(hipGetLastError());
    const int c = tm.mark();
    interp_dispatch<T>(recs, b.off, grid, b.vacc, g, cp.nvalid, p0, nb, st);
    CURIG_HIP(hipGetLastError());
    const int d = tm.mark();
    tm.span(a, bm, &t_fill);
    tm.span(bm, c, &t_fft);
    tm.span(c, d, &t_interp);
  }
  int e2 = tm.mark();
  Range rf("curig dirty2ms finalize");
  launch_finalize(pb.dbl, b.perm, b.vacc, (int)cp.nvalid, pb.uvw, pb.wgt, pb.wgt_stride, pb.vis, pb.vis_stride, g,
                  st);
  CURIG_HIP(hipGetLastError());
  int e3 = tm.mark();
  if (sts) {
    CURIG_HIP(hipStreamSynchronize(st));
    tm.settle();
    fill_stats(sts, cp);
    sts->t_prep = tm.ms(e0, e1);
    sts->t_grid = t_interp;
    sts->t_fft = t_fft;
    sts->t_plane = t_fill;
    sts->t_final = tm.ms(e2, e3);
    sts->t_total = tm.ms(e0, e3);
    sts->timed = 1;
  }
  return kOk;
}

// Graph of a whole multi-channel call, kept while the same call repeats.
struct GraphCache {
  uint64_t key = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
};
std::vector<GraphCache> g_graphs;

uint64_t fnv(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

// Field-by-field hashing (no struct padding bytes enter the key).
struct KeyHash {
  uint64_t h = 1469598103934665603ull;
  template <typename V> KeyHash &operator<<(const V &v) {
    static_assert(std::is_arithmetic<V>::value || std::is_pointer<V>::value, "scalar fields only");
    h = fnv(h, &v, sizeof(V));
    return *this;
  }
};

// Every host-side value a captured multi-channel sequence bakes into its
// launches: the problems' sizes, options and pointers, each channel's plan
// (geometry and unflagged-row count: the scan sentinel, the accumulator
// zeroing and the finalize count) and the working-set pointers.
uint64_t graph_key(bool to_image, const Problem *pbs, int n, const std::vector<ChanPlan> &cps, const Bufs &b) {
  KeyHash k;
  k << to_image << n;
  for (int i = 0; i < n; ++i) {
    const Problem &p = pbs[i];
    k << p.nrow << p.nx << p.ny << p.fov << p.freq << p.eps << p.sigma << p.sign << p.dbl << p.uvw << p.vis
      << p.vis_stride << p.wgt << p.wgt_stride << p.flag << p.flag_stride << p.dirty << p.accumulate;
    const GridSpec &s = p.spec;
    k << s.dx << s.dy << s.nu << s.nv << s.wstack << s.radians << s.zout << s.zmin << s.zmax;
  }
  for (const ChanPlan &cp : cps) {
    const Geom &g = cp.g;
    k << cp.nvalid << g.N1 << g.N2 << g.nf1 << g.nf2 << g.P << g.kw << g.kww << g.hdeg << g.hcoef_d << g.hcoef_f
      << g.beta << g.c << g.dx << g.dy << g.divn << g.zout << g.C << g.D << g.gamma << g.h << g.uscale << g.vscale
      << g.wscale << g.T1 << g.T2 << g.ntx << g.nty << g.ntiles << g.Pb << g.nbatch << g.Pbin << g.nlayer
      << g.nquad << g.fft_lds << g.rpw << g.cpw << g.cpw_inv << g.lds_coli_bytes << g.nct << g.tstride << g.nslot << g.slot_lg << g.lds_row_bytes
      << g.lds_col_bytes << g.rsplit << g.lds_rowp_bytes << g.colreg;
  }
  k << b.dev << b.tabs << b.hist << b.off << b.bsum << b.perm << b.binrank << b.recs << b.taps << b.grid << b.H << b.acc
    << b.vacc << b.tw1 << b.tw2 << b.hc_d << b.hc_f;
  return k.h;
}

void drop_graph(GraphCache &gc) {
  if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
  if (gc.graph) (void)hipGraphDestroy(gc.graph);
  gc.exec = nullptr;
  gc.graph = nullptr;
  gc.key = 0;
}

// One call's plan, reservation and launches (workspace lock held, ordered after
// the previous call).
template <typename T>
int run_locked(bool to_image, const Problem *pbs, int n, hipStream_t st, const Tuning *tune, int graph_mode,
               Stats *sts, int dev) {
  std::vector<ChanPlan> cps;
  CURIG_TRY(plan_all(pbs, n, st, tune, dev, cps));
  Bufs b;
  if (g_arena) {
    // no-allocation call: the LDS FFT path only (rocFFT plans allocate), and
    // as many planes per batch as the scratch holds
    for (const ChanPlan &cp : cps)
      if (!cp.g.fft_lds) {
        set_error("the no-allocation entry needs the LDS FFT path (power-of-two grid of at most 16384 per axis, "
                  "image at most half the grid, kernel width <= 16); grid %d x %d, kernel width %d",
                  cp.g.nf2, cp.g.nf1, cp.g.kw);
        return kErrScratch;
      }
    const size_t used0 = g_arena->used;
    int ier = kErrScratch;
    for (int Pb = cps[0].g.Pb; Pb >= 1; Pb = Pb > 1 ? Pb / 2 : 0) {
      for (ChanPlan &cp : cps) {
        cp.g.Pb = std::min(Pb, cp.g.P);
        cp.g.nbatch = (cp.g.P + cp.g.Pb - 1) / cp.g.Pb;
      }
      g_arena->used = used0;
      ier = reserve<T>(cps, pbs, to_image, dev, st, b);
      if (ier != kErrScratch) break;
    }
    CURIG_TRY(ier);
  } else {
    CURIG_TRY((reserve<T>(cps, pbs, to_image, dev, st, b)));
  }
  for (ChanPlan &cp : cps) {
    cp.g.hcoef_d = b.hc_d;
    cp.g.hcoef_f = b.hc_f;
  }
  // graph mode needs every kernel of the sequence to be a plain launch (LDS FFT path)
  bool use_graph = graph_mode != 0;
  for (const ChanPlan &cp : cps) use_graph = use_graph && cp.g.fft_lds;
  if (!use_graph) {
    Stats total, one;
    for (int i = 0; i < n; ++i) {
      CURIG_TRY(to_image ? enqueue_ms2dirty<T>(pbs[i], cps[i], b, st, sts ? &one : nullptr)
                         : enqueue_dirty2ms<T>(pbs[i], cps[i], b, st, sts ? &one : nullptr));
      if (sts) {
        if (i == 0) total = one;
        else {
          total.nvis += one.nvis;
          total.P += one.P;
          total.t_prep += one.t_prep;
          total.t_grid += one.t_grid;
          total.t_fft += one.t_fft;
          total.t_plane += one.t_plane;
          total.t_final += one.t_final;
          total.t_total += one.t_total;
        }
      }
    }
    if (sts) *sts = total;
    return kOk;
  }
  if ((int)g_graphs.size() <= dev) g_graphs.resize(dev + 1);
  GraphCache &gc = g_graphs[dev];
  const uint64_t key = graph_key(to_image, pbs, n, cps, b);
  Timer tm;
  tm.on = sts != nullptr;
  tm.s = st;
  int e0 = tm.mark();
  if (key != gc.key || !gc.exec) {
    drop_graph(gc);
    if (!gc.cap) CURIG_HIP(hipStreamCreateWithFlags(&gc.cap, hipStreamNonBlocking));
    CURIG_HIP(hipStreamBeginCapture(gc.cap, hipStreamCaptureModeThreadLocal));
    int ier = kOk;
    for (int i = 0; i < n && ier == kOk; ++i)
      ier = to_image ? enqueue_ms2dirty<T>(pbs[i], cps[i], b, gc.cap, nullptr)
                     : enqueue_dirty2ms<T>(pbs[i], cps[i], b, gc.cap, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(gc.cap, &graph);
    if (ier != kOk) {
      if (graph) (void)hipGraphDestroy(graph);
      return ier;
    }
    CURIG_HIP(e);
    gc.graph = graph;
    CURIG_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
    gc.key = key;
  }
  CURIG_HIP(hipGraphLaunch(gc.exec, st));
  int e1 = tm.mark();
  if (sts) {
    CURIG_HIP(hipStreamSynchronize(st));
    fill_stats(sts, cps.back());
    sts->nvis = 0;
    sts->P = 0;
    for (const ChanPlan &cp : cps) {
      sts->nvis += cp.nvalid;
      sts->P += cp.g.P;
    }
    sts->t_total = tm.ms(e0, e1);
    sts->timed = 1;
  }
  return kOk;
}

template <typename T>
int run_problems(bool to_image, const Problem *pbs, int n, hipStream_t st, const Tuning *tune, int graph_mode,
                 Stats *sts) {
  for (int i = 0; i < n; ++i) CURIG_TRY(check_args(pbs[i]));
  int dev = 0;
  CURIG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  CURIG_TRY(ws_order(dev, st));
  Range r(to_image ? "curig ms2dirty" : "curig dirty2ms");
  Stats own;
  const bool print = trace_flags().timing;
  if (print && !sts) sts = &own;
  const int ier = run_locked<T>(to_image, pbs, n, st, tune, graph_mode, sts, dev);
  if (print && ier == kOk)
    fprintf(stderr,
            "[curig] %s %s: %d channel(s), %ld vis, grid %d x %d x %d (%d batches), prep %.3f ms, %s %.3f ms, "
            "fft %.3f ms, %s %.3f ms, final %.3f ms, total %.3f ms\n",
            to_image ? "ms2dirty" : "dirty2ms", sizeof(T) == 8 ? "fp64" : "fp32", n, sts->nvis, sts->nf2, sts->nf1,
            sts->P, sts->nbatch, sts->t_prep, to_image ? "spread" : "interp", sts->t_grid, sts->t_fft,
            to_image ? "w-acc" : "w-fill", sts->t_plane, sts->t_final, sts->t_total);
  // marked on every path: a failed call may have queued launches too
  const int ierm = ws_mark(dev, st);
  return ier != kOk ? ier : ierm;
}

}  // namespace

int ms2dirty_device(const Problem &pb, hipStream_t st, const Tuning *tune, Stats *sts) {
  return pb.dbl ? run_problems<double>(true, &pb, 1, st, tune, 0, sts)
                : run_problems<float>(true, &pb, 1, st, tune, 0, sts);
}

int dirty2ms_device(const Problem &pb, hipStream_t st, const Tuning *tune, Stats *sts) {
  return pb.dbl ? run_problems<double>(false, &pb, 1, st, tune, 0, sts)
                : run_problems<float>(false, &pb, 1, st, tune, 0, sts);
}

int ms2dirty_device_noalloc(const Problem &pb, hipStream_t st, const Tuning *tune) {
  CURIG_TRY(check_args(pb));
  int dev = 0;
  CURIG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  CURIG_TRY(ws_order(dev, st));
  char *base = nullptr;
  CURIG_TRY(scratch_base(dev, &base));
  Arena ar;
  ar.base = base;
  ar.cap = kScratchBytes;
  g_arena = &ar;
  const int ier = pb.dbl ? run_locked<double>(true, &pb, 1, st, tune, 0, nullptr, dev)
                         : run_locked<float>(true, &pb, 1, st, tune, 0, nullptr, dev);
  g_arena = nullptr;
  const int ierm = ws_mark(dev, st);
  return ier != kOk ? ier : ierm;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int multi_device(bool to_image, const Problem *pbs, int n, hipStream_t st, const Tuning *tune, int graph_mode,
                 Stats *sts) {
  if (n <= 0) return kOk;
  return pbs[0].dbl ? run_problems<double>(to_image, pbs, n, st, tune, graph_mode, sts)
                    : run_problems<float>(to_image, pbs, n, st, tune, graph_mode, sts);
}

int explicit_ms2dirty_device(const Problem &pb, long npix, const int64_t *pix, void *out, hipStream_t st) {
  CURIG_TRY(check_args(pb));
  Geom g{};
  g.N1 = pb.ny;
  g.N2 = pb.nx;
  pixel_sizes(pb.nx, pb.ny, pb.fov, &pb.spec, g.dx, g.dy);
  g.divn = pb.spec.wstack ? 1 : 0;
  if (pix == nullptr) npix = (long)pb.nx * pb.ny;
  if (npix <= 0) return kOk;
  launch_explicit_ms2dirty(pb.dbl, pb, g, 2 * M_PI * pb.freq / kSpeedOfLight, pix, npix, (double2 *)out, st);
  CURIG_HIP(hipGetLastError());
  return kOk;
}

int explicit_dirty2ms_device(const Problem &pb, long nsel, const int64_t *rows, void *out, hipStream_t st) {
  CURIG_TRY(check_args(pb));
  if (nsel < 0 || (nsel > 0 && !rows)) {
    set_error("row selection: %ld rows, pointer %p", nsel, (const void *)rows);
    return kErrArg;
  }
  Geom g{};
  g.N1 = pb.ny;
  g.N2 = pb.nx;
  pixel_sizes(pb.nx, pb.ny, pb.fov, &pb.spec, g.dx, g.dy);
  g.divn = pb.spec.wstack ? 1 : 0;
  launch_explicit_dirty2ms(pb.dbl, pb, g, 2 * M_PI * pb.freq / kSpeedOfLight, rows, nsel, (double2 *)out, st);
  CURIG_HIP(hipGetLastError());
  return kOk;
}

void release_workspace() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (size_t d = 0; d < g_graphs.size(); ++d) {
    drop_graph(g_graphs[d]);
    if (g_graphs[d].cap) (void)hipStreamDestroy(g_graphs[d].cap);
  }
  g_graphs.clear();
  for (size_t d = 0; d < g_ws.size(); ++d) {
    (void)hipSetDevice((int)d);
    if (g_ws[d].pending) (void)hipEventSynchronize(g_ws[d].done);
    if (g_ws[d].done) (void)hipEventDestroy(g_ws[d].done);
    for (int s = 0; s < S_NSLOT; ++s)
      if (g_ws[d].p[s]) (void)hipFree(g_ws[d].p[s]);
  }
  g_ws.clear();
  (void)hipSetDevice(cur);
  fft_release_all();
}

}  // namespace curig
