// Host-visible interface of the gridder's kernel sources: the record types,
// one call's plan and working set, and the launch functions each kernel
// source defines (kern_prep.hip, kern_plane.hip, kern_lds_*.hip).  The host orchestration (gridder.hip) queues kernels
// only through these.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "curig_internal.h"
#include "fft_core.h"
#include "gridder.h"

namespace curig {

using fftc::cpx;

template <typename T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <typename T> using v2_t = typename V2<T>::type;

// Sorted visibility record: first grid cell of the kernel support on each axis
// (already reduced modulo the axis length), offset of that first cell from the
// exact position (kernel argument of the first tap), and the weighted value.
template <typename T> struct Rec;
template <> struct __align__(16) Rec<float> {
  int ix, iy, iw;
  float dx, dy, dw, re, im;
};
template <> struct __align__(8) Rec<double> {
  int ix, iy, iw, pad;
  double dx, dy, dw, re, im;
};

// One problem's plan: geometry, kernel options and unflagged-row count.
struct ChanPlan {
  Geom g;
  KernelOpts ko;
  long nvalid = 0;
};

// Device pointers of one call's working set (slots of the per-device workspace).
struct Bufs {
  int dev = 0;
  double *tabs = nullptr;  // fk1 [nf1/2+1] | fk2 [nf2/2+1] | quadrature nodes | weights
  int *hist = nullptr, *off = nullptr, *bsum = nullptr, *perm = nullptr;
  int2 *binrank = nullptr;
  void *recs = nullptr, *grid = nullptr, *H = nullptr;
  void *taps = nullptr;  // LDS path: each record's kernel taps, [nrec][3][tap_width] of the precision
  double2 *acc = nullptr, *vacc = nullptr;
  const void *tw1 = nullptr, *tw2 = nullptr;
  const double *hc_d = nullptr;  // piecewise-polynomial kernel coefficients, FP64 | FP32
  const float *hc_f = nullptr;
};

// Taps stored per record and axis on the LDS path: the kernel width rounded up
// to the row kernels' unrolled width (8 or 16).
inline int tap_width(int kw) { return kw <= 8 ? 8 : 16; }

// Dynamic LDS of the column kernels for columns of 16384 held in registers:
// the exchange buffer (FP64: one component; FP32: complex values), padded and
// rounded up to 16 bytes, then the column twiddles (cs bytes per complex).
inline size_t colreg_lds_bytes(size_t cs) {
  return (size_t)(fftc::padded_size(16384) + 1) * 8 + (size_t)fftc::twiddle_entries(16384) * cs;
}

inline int nblocks(long n, int bs = 256, long cap = 1 << 20) {
  long b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

// Bins a row workgroup of the LDS path walks at most (the plan coarsens bins to stay within it).
constexpr int kMaxWalk = 128;
// Column-tile width of the LDS path's intermediate planes.
constexpr int kTile = 1 << kTileLog2;

// Number of tiles walk_tiles(a, len, n, t, ntile) visits.
__host__ __device__ inline int tiles_count(int a, int len, int n, int t, int ntile) {
  if (len > n) len = n;
  int tile = a / t;
  int covered = ((tile + 1) * t < n ? (tile + 1) * t : n) - a;
  int visited = 1;
  while (covered < len && visited < ntile) {
    tile = (tile + 1 == ntile) ? 0 : tile + 1;
    covered += ((tile + 1) * t < n ? (tile + 1) * t : n) - tile * t;
    ++visited;
  }
  return visited;
}

// Upper bound of the bins a row workgroup walks (any p, row0).
__host__ __device__ inline int walk_bins_bound(int kww, int kw, int Pbin, int nlayer, int nrows, int T2, int nty) {
  int nl = (kww - 1) / Pbin + 2, ns = (nrows + kw - 2) / T2 + 2;
  if (nl > nlayer) nl = nlayer;
  if (ns > nty) ns = nty;
  return nl * ns;
}

// Exclusive scan of the bin counts: tiles of kScanTile ints per block.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// ---- kern_prep.hip: w range, binning, fills --------------------------------
void launch_w_minmax(int nblk, const double *uvw, int nrow, const uint8_t *flag, long fs, double *part, int *cnt,
                     hipStream_t st);
// count, scan and scatter of one problem's rows into bin order (records and permutation)
template <typename T>
void enqueue_binning(const Problem &pb, const ChanPlan &cp, const Bufs &b, hipStream_t st, bool with_value);
void launch_fill_zero_c128(double2 *p, long n, hipStream_t st);
void launch_zero_rows(bool dbl, void *vis, long stride, int nrow, hipStream_t st);

// ---- kern_plane.hip: full-plane path (rocFFT grids), corrections, direct sums
template <typename T>
void spread_dispatch(const Rec<T> *recs, const int *off, v2_t<T> *grid, const Geom &g, int p0, int nb, hipStream_t st);
template <typename T>
void interp_dispatch(const Rec<T> *recs, const int *off, const v2_t<T> *grid, double2 *vacc, const Geom &g,
                     long nvalid, int p0, int nb, hipStream_t st);
template <typename T>
void launch_wacc(const v2_t<T> *grid, double2 *acc, const Geom &g, int p0, int nb, int first, hipStream_t st);
template <typename T>
void launch_wfill(const double2 *gimg, v2_t<T> *grid, const Geom &g, int p0, int nb, hipStream_t st);
// out (precision of dbl_out) = corrected acc; colmajor: acc is [N1][N2]
void launch_epilogue(bool dbl_out, const double2 *acc, void *out, const Geom &g, const double *fk1, const double *fk2,
                     const double *q, int accumulate, int colmajor, hipStream_t st);
void launch_prologue(bool dbl_in, const void *dirty, double2 *gimg, const Geom &g, const double *fk1,
                     const double *fk2, const double *q, hipStream_t st);
template <typename T>
void launch_prologue_cols(bool dbl_in, const void *dirty, cpx<T> *gT, const Geom &g, const double *fk1,
                          const double *fk2, const double *q, hipStream_t st);
void launch_finalize(bool dbl, const int *perm, const double2 *vacc, int nrec, const double *uvw, const void *wgt,
                     long ws, void *vis, long vs, const Geom &g, hipStream_t st);
void launch_explicit_ms2dirty(bool dbl, const Problem &pb, const Geom &g, double fc2pi, const int64_t *pix, long npix,
                              double2 *out, hipStream_t st);
void launch_explicit_dirty2ms(bool dbl, const Problem &pb, const Geom &g, double fc2pi, const int64_t *rows,
                              long nsel, double2 *out, hipStream_t st);

// ---- kern_lds_*.hip: pruned LDS FFT path (kernel width <= 16) ---------------
// ms2dirty, one batch: spread + row FFT -> H (column-tiled)
template <typename T>
void lds_spread_rows(const Rec<T> *recs, const T *taps, const int *off, const Geom &g, int p0, int nb, cpx<T> *H,
                     const cpx<T> *tw1, hipStream_t st);
// ms2dirty, one batch: column FFT + w-accumulation into accT
template <typename T>
void lds_cols_forward(const Geom &g, int p0, int nb, const cpx<T> *H, double2 *accT, const cpx<T> *tw2, int first,
                      hipStream_t st);
// dirty2ms, one batch: w-fill + inverse column FFT -> H (column-tiled)
template <typename T>
void lds_cols_backward(const Geom &g, int p0, int nb, const cpx<T> *gT, cpx<T> *H, const cpx<T> *tw2, hipStream_t st);
// dirty2ms, one batch: inverse row FFT fused with the interpolation
template <typename T>
void lds_row_interp(const Rec<T> *recs, const T *taps, const int *off, const cpx<T> *H, double2 *vacc,
                    const cpx<T> *tw1, const Geom &g, int p0, int nb, hipStream_t st);
// raise the dynamic-LDS limit of the LDS-path kernels a plan uses
template <typename T> int allow_spread_rows(const Geom &g);
template <typename T> int allow_row_interp(const Geom &g);
template <typename T> int allow_cols(const Geom &g);

}  // namespace curig
