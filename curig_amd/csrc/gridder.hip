// ES-kernel improved-w-stacking gridder for CDNA4 (gfx950).
//
// Pipeline (one channel):
//   ms2dirty  prep: w range -> plan; bin rows by (w-plane batch, grid tile of the
//             kernel's first cell); counting sort into 32/56-byte records
//             for each batch of Pb w planes:
//               spread   tile-owner gather into LDS, every grid cell written once
//               rocFFT   batched 2-D, sign +1
//               w-accum  acc[pix] += sum_p plane_p[cell(pix)] e^{+i z'(p - P/2)} (FP64)
//             epilogue: ES / w corrections, / n
//   dirty2ms  prologue: corrections, / n; for each batch: w-fill (zero-padded planes
//             with the w phase), rocFFT sign -1, interpolation into FP64 per-row
//             accumulators; finalize: weight and e^{-i D w}, scatter to row order.
//
// Mathematically this is the reference's operator (src/astro/ra_exec.cu:141-436):
// same grid sizes, w-plane count and scaling (set_nhg_type3), kernel and
// correction factors; the full nf3 x nf2 x nf1 grid of the reference is never
// materialised (only Pb planes at a time), and the kernel is normalised to peak 1
// (its Fourier-series factors carry the same normalisation, so it cancels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "curig_internal.h"
#include "fft.h"
#include "fft_core.h"
#include "trig.h"
#include "gridder.h"

namespace curig {

#define M_1_2PI_D 0.159154943091895336  // include/utils.h:61
#define M_2PI_D 6.28318530717958648     // include/utils.h:62

// --------------------------------------------------------------------------
// types
// --------------------------------------------------------------------------
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

// --------------------------------------------------------------------------
// device helpers
// --------------------------------------------------------------------------
// Normalised ES kernel exp(beta (sqrt(1 - c d^2) - 1)), 0 for |d| >= kw/2.  The
// exponent is formed as -beta c d^2 / (sqrt(1 - c d^2) + 1): no cancellation, so
// the FP32 value keeps ~1e-7 relative accuracy near the peak (beta ~ 10-30 would
// amplify the rounding of sqrt(.) - 1 to ~1e-6).
__device__ __forceinline__ float es_eval(float d, float beta, float c, float hw) {
  const float t = c * d * d;
  return (fabsf(d) < hw && t < 1.0f) ? expf(-beta * t / (sqrtf(1.0f - t) + 1.0f)) : 0.0f;
}
__device__ __forceinline__ double es_eval(double d, double beta, double c, double hw) {
  const double t = c * d * d;
  return (fabs(d) < hw && t < 1.0) ? exp(-beta * t / (sqrt(1.0 - t) + 1.0)) : 0.0;
}

// Kernel value at tap i of a record whose first tap lies at offset d (d + i in
// [-kw/2, kw/2)): exp(sqrt()) evaluation, or Horner's rule on the piecewise
// polynomial of interval i at z = 2 (d + kw/2) - 1 (hdeg >= 0; coefficients
// [hdeg + 1][kw], highest power first, so one power of all taps is contiguous).
template <typename T> struct Ker {
  T beta, c, hw;
  const T *coef;
  int deg, kw;
  bool flatw;  // 2-D form: the single plane's weight is 1
  // weight of plane tap k (first tap at offset d)
  __device__ __forceinline__ T w(T d, int k) const { return flatw ? (T)1 : (*this)(d, k); }
  __device__ __forceinline__ T operator()(T d, int i) const {
    if (deg >= 0) {
      const T z = (T)2 * (d + hw) - (T)1;
      T r = coef[i];
      for (int k = 1; k <= deg; ++k) r = r * z + coef[k * kw + i];
      return r;
    }
    return es_eval(d + (T)i, beta, c, hw);
  }
  // all kw taps (zero beyond kw): the polynomial's powers outermost, so each
  // step is KWM independent multiply-adds
  template <int KWM> __device__ __forceinline__ void taps(T d, T *o) const {
    if (deg >= 0) {
      const T z = (T)2 * (d + hw) - (T)1;
#pragma unroll
      for (int i = 0; i < KWM; ++i) o[i] = (i < kw) ? coef[i] : (T)0;
      for (int k = 1; k <= deg; ++k) {
        const T *q = coef + k * kw;
#pragma unroll
        for (int i = 0; i < KWM; ++i)
          if (i < kw) o[i] = o[i] * z + q[i];
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < KWM; ++i) o[i] = (i < kw) ? es_eval(d + (T)i, beta, c, hw) : (T)0;
  }
};
template <typename T> __device__ __forceinline__ Ker<T> make_ker(const Geom &g) {
  Ker<T> k;
  k.beta = (T)g.beta;
  k.c = (T)g.c;
  k.hw = (T)(0.5 * g.kw);
  k.deg = g.hdeg;
  k.kw = g.kw;
  k.flatw = g.kww == 1;
  if constexpr (sizeof(T) == 8)
    k.coef = (const T *)g.hcoef_d;
  else
    k.coef = (const T *)g.hcoef_f;
  return k;
}

__device__ __forceinline__ int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// x in radians, folded into [-pi, pi) and rescaled to grid units in [0, nf):
// SHIFT_RESCALE(x, nf, 1) (include/conv.h:21) after an exact 2pi fold.
__device__ __forceinline__ double rescale_fold(double x, int nf) {
  double xf = x - floor(x * M_1_2PI_D + 0.5) * M_2PI_D;
  return (xf * M_1_2PI_D + 0.5) * nf;
}

// n - 1 without cancellation, and n, for pixel (row, col)
__device__ __forceinline__ void pixel_n(int row, int col, const Geom &g, double &nm1, double &n) {
  if (g.zout) {  // outputs of a type-3 transform: n - 1 given per pixel
    nm1 = g.zout[(long)row * g.N1 + col];
    n = 1.0;
    return;
  }
  double l = (double)(row - g.N2 / 2) * g.dx;
  double m = (double)(col - g.N1 / 2) * g.dy;
  double r2 = l * l + m * m;
  n = sqrt(1.0 - r2);
  nm1 = -r2 / (n + 1.0);
}

struct Coords {
  int ix, iy, iw;
  double dx, dy, dw, we;
};

// Grid coordinates of row j (src/astro/precomp.cu:48-72 scaling, cuft.cu:53-75
// type-3 pre-stage, conv.cu:176-198 start cell and first kernel argument).
__device__ __forceinline__ Coords row_coords(const double *__restrict__ uvw, long j, const Geom &g) {
  double u = uvw[3 * j], v = uvw[3 * j + 1], w = uvw[3 * j + 2];
  Coords c;
  double xu = u * g.uscale;  // grid slow axis (nf2)
  double yv = v * g.vscale;  // grid fast axis (nf1)
  c.we = w * g.wscale;
  double wp = (c.we - g.C) / g.gamma;
  double tx = rescale_fold(yv, g.nf1);
  double ty = rescale_fold(xu, g.nf2);
  double tw = (wp * M_1_2PI_D + 0.5) * g.P;
  double hk = 0.5 * g.kw;
  double sx = ceil(tx - hk), sy = ceil(ty - hk), sw = ceil(tw - hk);
  c.dx = sx - tx;
  c.dy = sy - ty;
  c.dw = sw - tw;
  c.ix = imod((int)sx, g.nf1);
  c.iy = imod((int)sy, g.nf2);
  c.iw = imod((int)sw, g.P);
  if (g.kww == 1) {  // 2-D form: the single plane, kernel argument 0 (weight 1)
    c.iw = 0;
    c.dw = 0.0;
  }
  return c;
}

__device__ __forceinline__ int row_bin(const Coords &c, const Geom &g) {
  int tile = (c.iy / g.T2) * g.ntx + (c.ix / g.T1);
  return (c.iw / g.Pbin) * g.ntiles + tile;
}

template <typename T> __device__ __forceinline__ double2 load_c(const void *p, long j);
template <> __device__ __forceinline__ double2 load_c<double>(const void *p, long j) {
  return ((const double2 *)p)[j];
}
template <> __device__ __forceinline__ double2 load_c<float>(const void *p, long j) {
  float2 v = ((const float2 *)p)[j];
  return make_double2(v.x, v.y);
}
template <typename T> __device__ __forceinline__ double load_r(const void *p, long j) {
  return (double)((const T *)p)[j];
}
template <typename T> __device__ __forceinline__ void store_c(void *p, long j, double re, double im);
template <> __device__ __forceinline__ void store_c<double>(void *p, long j, double re, double im) {
  ((double2 *)p)[j] = make_double2(re, im);
}
template <> __device__ __forceinline__ void store_c<float>(void *p, long j, double re, double im) {
  ((float2 *)p)[j] = make_float2((float)re, (float)im);
}
template <typename T> __device__ __forceinline__ void add_c(void *p, long j, double re, double im);
template <> __device__ __forceinline__ void add_c<double>(void *p, long j, double re, double im) {
  double2 o = ((double2 *)p)[j];
  ((double2 *)p)[j] = make_double2(o.x + re, o.y + im);
}
template <> __device__ __forceinline__ void add_c<float>(void *p, long j, double re, double im) {
  float2 o = ((float2 *)p)[j];
  ((float2 *)p)[j] = make_float2((float)(o.x + re), (float)(o.y + im));
}

__device__ __forceinline__ bool row_skipped(const uint8_t *flag, long fs, long j) {
  return flag != nullptr && flag[j * fs] != 0;
}

// --------------------------------------------------------------------------
// prep kernels
// --------------------------------------------------------------------------
// Per-block min / max of w (metres) over unflagged rows.
__global__ void __launch_bounds__(256) w_minmax_kernel(const double *__restrict__ uvw, int nrow,
                                                        const uint8_t *__restrict__ flag, long fs,
                                                        double *__restrict__ part, int *__restrict__ cnt) {
  __shared__ double smin[256], smax[256];
  __shared__ int scnt[256];
  double lo = INFINITY, hi = -INFINITY;
  int n = 0;
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nrow; j += (long)gridDim.x * blockDim.x) {
    if (row_skipped(flag, fs, j)) continue;
    double w = uvw[3 * j + 2];
    // NaN propagates to the host check through the sum of |u|+|v|+|w|
    double a = uvw[3 * j] + uvw[3 * j + 1];
    if (!(fabs(a) < INFINITY) || !(fabs(w) < INFINITY)) {
      lo = NAN;
      hi = NAN;
    } else if (lo == lo) {
      lo = fmin(lo, w);
      hi = fmax(hi, w);
    }
    ++n;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  scnt[threadIdx.x] = n;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double a = smin[threadIdx.x], b = smin[threadIdx.x + s];
      smin[threadIdx.x] = (a != a || b != b) ? NAN : fmin(a, b);
      a = smax[threadIdx.x];
      b = smax[threadIdx.x + s];
      smax[threadIdx.x] = (a != a || b != b) ? NAN : fmax(a, b);
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
    cnt[blockIdx.x] = scnt[0];
  }
}

__global__ void fill_int_kernel(int *__restrict__ p, long n, int v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void fill_zero_c128_kernel(double2 *__restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

template <typename T>
__global__ void zero_rows_kernel(void *__restrict__ vis, long stride, int nrow) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nrow; j += (long)gridDim.x * blockDim.x)
    store_c<T>(vis, j * stride, 0.0, 0.0);
}

// Count rows per bin; each row keeps its rank inside the bin (GM-sort idea).
__global__ void __launch_bounds__(256) bin_count_kernel(const double *__restrict__ uvw, int nrow,
                                                         const uint8_t *__restrict__ flag, long fs, Geom g,
                                                         int *__restrict__ hist, int2 *__restrict__ binrank) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nrow; j += (long)gridDim.x * blockDim.x) {
    if (row_skipped(flag, fs, j)) {
      binrank[j] = make_int2(-1, 0);
      continue;
    }
    Coords c = row_coords(uvw, j, g);
    int b = row_bin(c, g);
    int r = atomicAdd(&hist[b], 1);
    binrank[j] = make_int2(b, r);
  }
}

// Exclusive scan of n ints: three passes (block sums, scan of sums, apply).
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__(kScanBlock) scan_reduce_kernel(const int *__restrict__ in, long n,
                                                                  int *__restrict__ bsum) {
  __shared__ int s[kScanBlock];
  long base = (long)blockIdx.x * kScanTile;
  int acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    long i = base + (long)k * kScanBlock + threadIdx.x;
    if (i < n) acc += in[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = kScanBlock / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

// single block: exclusive scan of nb block sums in place
__global__ void __launch_bounds__(1024) scan_sums_kernel(int *__restrict__ bsum, int nb) {
  __shared__ int s[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < nb ? bsum[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
}

// out[i] = bsum[block] + exclusive scan within the tile; out[n] = total
__global__ void __launch_bounds__(kScanBlock) scan_apply_kernel(const int *__restrict__ in, long n,
                                                                 const int *__restrict__ bsum,
                                                                 int *__restrict__ out, int total) {
  __shared__ int s[kScanBlock];
  long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
  int loc[kScanItems];
  int sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    long i = base + k;
    loc[k] = i < n ? in[i] : 0;
    sum += loc[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int run = bsum[blockIdx.x] + s[threadIdx.x] - sum;
  for (int k = 0; k < kScanItems; ++k) {
    long i = base + k;
    if (i < n) out[i] = run;
    run += loc[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = total;
}

// Scatter rows into bin order; the value carries the weight and e^{+i D w}
// (ms2dirty pre-stages, src/astro/ra_exec.cu:98-110).
template <typename T>
__global__ void __launch_bounds__(256) bin_scatter_kernel(const double *__restrict__ uvw, int nrow,
                                                           const void *__restrict__ vis, long vs,
                                                           const void *__restrict__ wgt, long ws, Geom g,
                                                           const int2 *__restrict__ binrank,
                                                           const int *__restrict__ off, Rec<T> *__restrict__ recs,
                                                           int *__restrict__ perm, int with_value) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < nrow; j += (long)gridDim.x * blockDim.x) {
    int2 br = binrank[j];
    if (br.x < 0) continue;
    Coords c = row_coords(uvw, j, g);
    int dst = off[br.x] + br.y;
    Rec<T> r;
    r.ix = c.ix;
    r.iy = c.iy;
    r.iw = c.iw;
    r.dx = (T)c.dx;
    r.dy = (T)c.dy;
    r.dw = (T)c.dw;
    if (with_value) {
      double2 v = load_c<T>(vis, j * vs);
      double wt = wgt ? load_r<T>(wgt, j * ws) : 1.0;
      double s, co;
      sincos(g.D * c.we, &s, &co);
      r.re = (T)(wt * (v.x * co - v.y * s));
      r.im = (T)(wt * (v.x * s + v.y * co));
    } else {
      r.re = 0;
      r.im = 0;
    }
    recs[dst] = r;
    perm[dst] = (int)j;
  }
}

// --------------------------------------------------------------------------
// spread: one block owns one T1 x T2 tile of one w plane and writes every cell
// of it exactly once (no global atomics, no memset of the grid).
// --------------------------------------------------------------------------
// Walk the cyclic interval [a, a+len) of an axis of length n cut into tiles of
// size t (last tile may be partial), calling f(tile) once per tile touched.
template <typename F>
__device__ __forceinline__ void walk_tiles(int a, int len, int n, int t, int ntile, F f) {
  if (len > n) len = n;
  int tile = a / t;
  int end = min((tile + 1) * t, n);
  int covered = end - a;
  f(tile);
  int visited = 1;
  while (covered < len && visited < ntile) {
    tile = (tile + 1 == ntile) ? 0 : tile + 1;
    covered += min((tile + 1) * t, n) - tile * t;
    f(tile);
    ++visited;
  }
}

template <typename T, int KWM>
__global__ void __launch_bounds__(256) spread_tile_kernel(const Rec<T> *__restrict__ recs,
                                                          const int *__restrict__ off,
                                                          v2_t<T> *__restrict__ grid, Geom g, int p0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *acc = reinterpret_cast<T *>(smem_raw);
  const int tile = blockIdx.x;
  const int pl = blockIdx.y;
  const int p = p0 + pl;
  const int tx = tile % g.ntx, ty = tile / g.ntx;
  const int x0 = tx * g.T1, y0 = ty * g.T2;
  const int wx = min(g.T1, g.nf1 - x0), wy = min(g.T2, g.nf2 - y0);
  const int ncell = g.T1 * g.T2;
  for (int i = threadIdx.x; i < 2 * ncell; i += blockDim.x) acc[i] = (T)0;
  __syncthreads();

  const int kw = g.kw, kww = g.kww;
  const Ker<T> K = make_ker<T>(g);
  // source rows: first plane in [p-kww+1, p], first cell in [x0-kw+1, x0+wx-1] x [y0-kw+1, y0+wy-1]
  walk_tiles(imod(p - kww + 1, g.P), kww, g.P, g.Pbin, g.nlayer, [&](int sb) {
    walk_tiles(imod(y0 - kw + 1, g.nf2), wy + kw - 1, g.nf2, g.T2, g.nty, [&](int sty) {
      walk_tiles(imod(x0 - kw + 1, g.nf1), wx + kw - 1, g.nf1, g.T1, g.ntx, [&](int stx) {
        const int bin = sb * g.ntiles + sty * g.ntx + stx;
        const int r0 = off[bin], r1 = off[bin + 1];
        for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
          const Rec<T> rc = recs[r];
          int k = p - rc.iw;
          if (k < 0) k += g.P;
          if (k >= kww) continue;
          int rx = rc.ix - x0;
          if (rx < 0) rx += g.nf1;
          int ry = rc.iy - y0;
          if (ry < 0) ry += g.nf2;
          // quick reject: footprint misses the tile on either axis
          // (a footprint that wraps past the axis end always covers cell 0 of the tile row)
          const bool hitx = (rx < wx) || (rx + kw - 1 >= g.nf1);
          const bool hity = (ry < wy) || (ry + kw - 1 >= g.nf2);
          if (!hitx || !hity) continue;
          const T vw = K.w(rc.dw, k);
          const T vre = rc.re * vw, vim = rc.im * vw;
          T kx[KWM];
          K.template taps<KWM>(rc.dx, kx);
#pragma unroll
          for (int jj = 0; jj < KWM; ++jj) {
            if (jj >= kw) break;
            int cy = ry + jj;
            if (cy >= g.nf2) cy -= g.nf2;
            if (cy >= wy) continue;
            const T ky = K(rc.dy, jj);
            const T are = vre * ky, aim = vim * ky;
            T *rowp = acc + 2 * (cy * g.T1);
#pragma unroll
            for (int i = 0; i < KWM; ++i) {
              if (i >= kw) break;
              int cx = rx + i;
              if (cx >= g.nf1) cx -= g.nf1;
              if (cx >= wx) continue;
              atomicAdd(rowp + 2 * cx, are * kx[i]);
              atomicAdd(rowp + 2 * cx + 1, aim * kx[i]);
            }
          }
        }
      });
    });
  });
  __syncthreads();
  v2_t<T> *out = grid + ((size_t)pl * g.nf2 + y0) * g.nf1 + x0;
  for (int i = threadIdx.x; i < wy * g.T1; i += blockDim.x) {
    int cy = i / g.T1, cx = i - cy * g.T1;
    if (cx < wx) {
      v2_t<T> v;
      v.x = acc[2 * i];
      v.y = acc[2 * i + 1];
      out[(size_t)cy * g.nf1 + cx] = v;
    }
  }
}

// --------------------------------------------------------------------------
// w-accumulate (ms2dirty): acc[pix] += sum_{p in batch} F_p[cell(pix)] e^{+i z'(p - P/2)}
// src/fourier/cuft.cu:520-572 restated per batch, FP64 phase and sum.
// --------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) wacc_kernel(const v2_t<T> *__restrict__ grid, double2 *__restrict__ acc,
                                                    Geom g, int p0, int nb, int first) {
  const long npix = (long)g.N1 * g.N2;
  const size_t plane = (size_t)g.nf1 * g.nf2;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < npix; idx += (long)gridDim.x * blockDim.x) {
    int row = (int)(idx / g.N1), col = (int)(idx - (long)row * g.N1);
    int w1 = col >= g.N1 / 2 ? col - g.N1 / 2 : g.nf1 + col - g.N1 / 2;
    int w2 = row >= g.N2 / 2 ? row - g.N2 / 2 : g.nf2 + row - g.N2 / 2;
    double nm1, n;
    pixel_n(row, col, g, nm1, n);
    double zp = (nm1 - g.D) * g.gamma * g.h;
    double s, c, ss, cs;
    sincos(zp * (double)(p0 - g.P / 2), &s, &c);
    sincos(zp, &ss, &cs);
    double ar = 0.0, ai = 0.0;
    const v2_t<T> *src = grid + (size_t)w2 * g.nf1 + w1;
    for (int k = 0; k < nb; ++k) {
      v2_t<T> f = src[(size_t)k * plane];
      double fr = (double)f.x, fi = (double)f.y;
      ar += fr * c - fi * s;
      ai += fr * s + fi * c;
      double cn = c * cs - s * ss;
      s = s * cs + c * ss;
      c = cn;
    }
    if (first) {
      acc[idx] = make_double2(ar, ai);
    } else {
      double2 o = acc[idx];
      acc[idx] = make_double2(o.x + ar, o.y + ai);
    }
  }
}

// phihat3(z') = sum_n g_n cos(x_n z')  (src/fourier/deconv.cu:45-58)
__device__ __forceinline__ double phihat3(double zp, const double *__restrict__ q, int nq) {
  double s = 0.0;
  for (int i = 0; i < nq; ++i) s += q[nq + i] * cos(q[i] * zp);
  return s;
}

// Corrections for pixel (row, col): 1 / (phihat1 phihat2 phihat3 n) and the
// w-term phase C (z - D)  (deconv.cu:236-276, 371-399; ra_exec.cu:22-37).  In
// the 2-D form (divn == 0: every w is 0, so C == 0; one plane, kww == 1) the
// 1 / n and phihat3 are left out.
__device__ __forceinline__ void pixel_correction(int row, int col, const Geom &g, const double *__restrict__ fk1,
                                                 const double *__restrict__ fk2, const double *__restrict__ q,
                                                 double &scale, double &phase) {
  double nm1, n;
  pixel_n(row, col, g, nm1, n);
  double zp = (nm1 - g.D) * g.gamma * g.h;
  double ker = fk1[abs(col - g.N1 / 2)] * fk2[abs(row - g.N2 / 2)] * (g.kww == 1 ? 1.0 : phihat3(zp, q, g.nquad));
  scale = 1.0 / (g.divn ? ker * n : ker);
  phase = (nm1 - g.D) * g.C;
}

// acc row-major [N2][N1], or column-major [N1][N2] (LDS path) when colmajor
template <typename TO>
__global__ void __launch_bounds__(256) epilogue_kernel(const double2 *__restrict__ acc, void *__restrict__ out,
                                                        Geom g, const double *__restrict__ fk1,
                                                        const double *__restrict__ fk2,
                                                        const double *__restrict__ q, int accumulate, int colmajor) {
  const long npix = (long)g.N1 * g.N2;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < npix; idx += (long)gridDim.x * blockDim.x) {
    int row = (int)(idx / g.N1), col = (int)(idx - (long)row * g.N1);
    double scale, phase, s, c;
    pixel_correction(row, col, g, fk1, fk2, q, scale, phase);
    sincos(phase, &s, &c);
    double2 a = acc[colmajor ? (long)col * g.N2 + row : idx];
    double re = (a.x * c - a.y * s) * scale, im = (a.x * s + a.y * c) * scale;
    if (accumulate)
      add_c<TO>(out, idx, re, im);
    else
      store_c<TO>(out, idx, re, im);
  }
}

// dirty2ms prologue: g[pix] = dirty / n e^{-i C (z - D)} / (phihat1 phihat2 phihat3)
template <typename TI>
__global__ void __launch_bounds__(256) prologue_kernel(const void *__restrict__ dirty, double2 *__restrict__ gimg,
                                                        Geom g, const double *__restrict__ fk1,
                                                        const double *__restrict__ fk2,
                                                        const double *__restrict__ q) {
  const long npix = (long)g.N1 * g.N2;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < npix; idx += (long)gridDim.x * blockDim.x) {
    int row = (int)(idx / g.N1), col = (int)(idx - (long)row * g.N1);
    double scale, phase, s, c;
    pixel_correction(row, col, g, fk1, fk2, q, scale, phase);
    sincos(-phase, &s, &c);
    double2 a = load_c<TI>(dirty, idx);
    gimg[idx] = make_double2((a.x * c - a.y * s) * scale, (a.x * s + a.y * c) * scale);
  }
}

// w-fill (dirty2ms): plane p = zero-padded image x e^{-i z'(p - P/2)} (cuft.cu:575-612)
template <typename T>
__global__ void __launch_bounds__(256) wfill_kernel(const double2 *__restrict__ gimg, v2_t<T> *__restrict__ grid,
                                                     Geom g, int p0, int nb) {
  const long ncell = (long)g.nf1 * g.nf2;
  const int h1 = g.N1 / 2, h2 = g.N2 / 2;
  for (long cell = blockIdx.x * (long)blockDim.x + threadIdx.x; cell < ncell; cell += (long)gridDim.x * blockDim.x) {
    int k2 = (int)(cell / g.nf1), k1 = (int)(cell - (long)k2 * g.nf1);
    int col = k1 < g.N1 - h1 ? k1 + h1 : (k1 >= g.nf1 - h1 ? k1 - g.nf1 + h1 : -1);
    int row = k2 < g.N2 - h2 ? k2 + h2 : (k2 >= g.nf2 - h2 ? k2 - g.nf2 + h2 : -1);
    v2_t<T> *dst = grid + cell;
    if (col < 0 || row < 0) {
      v2_t<T> z;
      z.x = 0;
      z.y = 0;
      for (int k = 0; k < nb; ++k) dst[(size_t)k * ncell] = z;
      continue;
    }
    double2 v = gimg[(long)row * g.N1 + col];
    double nm1, n;
    pixel_n(row, col, g, nm1, n);
    double zp = (nm1 - g.D) * g.gamma * g.h;
    double s, c, ss, cs;
    sincos(-zp * (double)(p0 - g.P / 2), &s, &c);
    sincos(-zp, &ss, &cs);
    for (int k = 0; k < nb; ++k) {
      v2_t<T> o;
      o.x = (T)(v.x * c - v.y * s);
      o.y = (T)(v.x * s + v.y * c);
      dst[(size_t)k * ncell] = o;
      double cn = c * cs - s * ss;
      s = s * cs + c * ss;
      c = cn;
    }
  }
}

// Records whose first plane lies in the cyclic interval [lo, lo + len) (len <= P):
// up to two ranges of bin layers, read from the bin offsets on the device.
__device__ __forceinline__ void layer_ranges(const int *__restrict__ off, const Geom &g, int lo, int len, int &a0,
                                             int &a1, int &b0, int &b1) {
  if (len > g.P) len = g.P;
  const int first = lo / g.Pbin;
  const int last_plane = lo + len - 1;
  a0 = off[(long)first * g.ntiles];
  b0 = b1 = 0;
  if (last_plane < g.P) {
    a1 = off[(long)(last_plane / g.Pbin + 1) * g.ntiles];
  } else {
    a1 = off[(long)g.nlayer * g.ntiles];
    int lastl = (last_plane - g.P) / g.Pbin;
    if (lastl >= first) lastl = first - 1;  // the whole circle is covered
    if (lastl >= 0) {
      b0 = off[0];
      b1 = off[(long)(lastl + 1) * g.ntiles];
    }
  }
}

// interpolation (dirty2ms): per record, sum over the kernel support inside the
// batch's planes; FP32 (or FP64) multiply-accumulate per plane, FP64 across planes.
// src/fourier/interp.cu:143-228.
template <typename T, int KWM>
__global__ void __launch_bounds__(256) interp_kernel(const Rec<T> *__restrict__ recs,
                                                      const int *__restrict__ off,
                                                      const v2_t<T> *__restrict__ grid,
                                                      double2 *__restrict__ vacc, Geom g, int p0, int nb) {
  int ra0, ra1, rb0, rb1;
  layer_ranges(off, g, imod(p0 - g.kww + 1, g.P), nb + g.kww - 1, ra0, ra1, rb0, rb1);
  const long n0 = ra1 - ra0, ntot = n0 + (rb1 - rb0);
  const size_t plane = (size_t)g.nf1 * g.nf2;
  const int kw = g.kw;
  const Ker<T> K = make_ker<T>(g);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < ntot; t += (long)gridDim.x * blockDim.x) {
    const int r = t < n0 ? ra0 + (int)t : rb0 + (int)(t - n0);
    const Rec<T> rc = recs[r];
    T kx[KWM], ky[KWM];
    bool any = false;
    for (int k = 0; k < g.kww; ++k) {
      int q = rc.iw + k;
      if (q >= g.P) q -= g.P;
      if (q >= p0 && q < p0 + nb) any = true;
    }
    if (!any) continue;
    K.template taps<KWM>(rc.dx, kx);
    K.template taps<KWM>(rc.dy, ky);
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < g.kww; ++k) {
      int q = rc.iw + k;
      if (q >= g.P) q -= g.P;
      if (q < p0 || q >= p0 + nb) continue;
      const T kz = K.w(rc.dw, k);
      const v2_t<T> *pb = grid + (size_t)(q - p0) * plane;
      T pr = 0, pi = 0;
#pragma unroll
      for (int jj = 0; jj < KWM; ++jj) {
        if (jj < kw) {
          int cy = rc.iy + jj;
          if (cy >= g.nf2) cy -= g.nf2;
          const v2_t<T> *rowp = pb + (size_t)cy * g.nf1;
          T rr = 0, ri = 0;
#pragma unroll
          for (int i = 0; i < KWM; ++i) {
            if (i < kw) {
              int cx = rc.ix + i;
              if (cx >= g.nf1) cx -= g.nf1;
              const v2_t<T> f = rowp[cx];
              rr += f.x * kx[i];
              ri += f.y * kx[i];
            }
          }
          pr += rr * ky[jj];
          pi += ri * ky[jj];
        }
      }
      sr += (double)(pr * kz);
      si += (double)(pi * kz);
    }
    const double2 o = vacc[r];
    vacc[r] = make_double2(o.x + sr, o.y + si);
  }
}

// finalize (dirty2ms): vis_j = acc * wgt_j * e^{-i D w_j}  (ra_exec.cu:56-89)
template <typename T>
__global__ void __launch_bounds__(256) finalize_kernel(const int *__restrict__ perm, const double2 *__restrict__ vacc,
                                                        int nrec, const double *__restrict__ uvw,
                                                        const void *__restrict__ wgt, long ws,
                                                        void *__restrict__ vis, long vs, Geom g) {
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < nrec; s += (long)gridDim.x * blockDim.x) {
    long j = perm[s];
    double we = uvw[3 * j + 2] * g.wscale;
    double sn, cs;
    sincos(-g.D * we, &sn, &cs);
    double wt = wgt ? load_r<T>(wgt, j * ws) : 1.0;
    double2 a = vacc[s];
    store_c<T>(vis, j * vs, wt * (a.x * cs - a.y * sn), wt * (a.x * sn + a.y * cs));
  }
}

// FP64 direct sum at selected pixels (explicit_gridder, precomp.cu:242-275), with
// the weight applied and flagged rows skipped.  One block per pixel.
template <typename TI>
__global__ void __launch_bounds__(256) explicit_kernel(const double *__restrict__ uvw, int nrow,
                                                        const void *__restrict__ vis, long vs,
                                                        const void *__restrict__ wgt, long ws,
                                                        const uint8_t *__restrict__ flag, long fs, Geom g,
                                                        double fc2pi, int sign, const int64_t *__restrict__ pix,
                                                        long npix, double2 *__restrict__ out) {
  __shared__ double sre[256], sim[256];
  for (long q = blockIdx.x; q < npix; q += gridDim.x) {
    long p = pix ? pix[q] : q;
    int i = (int)(p / g.N1), k = (int)(p - (long)i * g.N1);
    double nm1, n;
    pixel_n(i, k, g, nm1, n);
    if (!g.divn) nm1 = 0.0, n = 1.0;  // 2-D form: w ignored, no 1 / n
    double l = (double)(i - g.N2 / 2) * g.dx, m = (double)(k - g.N1 / 2) * g.dy;
    double re = 0.0, im = 0.0;
    for (long j = threadIdx.x; j < nrow; j += blockDim.x) {
      if (row_skipped(flag, fs, j)) continue;
      double ph = fc2pi * (l * uvw[3 * j] + m * uvw[3 * j + 1] + sign * uvw[3 * j + 2] * nm1);
      double s, c;
      sincos(ph, &s, &c);
      double2 v = load_c<TI>(vis, j * vs);
      double wt = wgt ? load_r<TI>(wgt, j * ws) : 1.0;
      re += wt * (v.x * c - v.y * s);
      im += wt * (v.x * s + v.y * c);
    }
    sre[threadIdx.x] = re;
    sim[threadIdx.x] = im;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) {
        sre[threadIdx.x] += sre[threadIdx.x + st];
        sim[threadIdx.x] += sim[threadIdx.x + st];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = make_double2(sre[0] / n, sim[0] / n);
    __syncthreads();
  }
}


// --------------------------------------------------------------------------
// pruned LDS FFT path (power-of-two grids)
//
// ms2dirty per batch of planes: (1) one workgroup per grid row of a plane
// spreads the row's visibilities into LDS, runs the row FFT (length nf1) and
// writes the N1 image columns to the column-tiled buffer H; (2) one workgroup
// per image column runs the column FFTs (length nf2) of all the batch's planes
// and accumulates the N2 image rows with their w phases in FP64.  dirty2ms runs
// the mirror image: (1) per image column, w-fill + inverse column FFT of every
// plane -> H; (2) per grid row, inverse row FFT (the row expanded to nf1 cells)
// and interpolation.  Each plane crosses HBM twice (written once, read once) at
// half the grid size.
// --------------------------------------------------------------------------
using fftc::cpx;

// Record ranges of one row workgroup (plane p, grid rows [row0, row0 + nrows)):
// the bins of every layer holding a first plane in [p - kw + 1, p] and of every
// strip overlapping rows [row0 - kw + 1, row0 + nrows - 1] (both cyclic), with
// exclusive prefix counts, so that the workgroup walks them as one flat index
// range and issues all its record loads at once.
constexpr int kMaxWalk = 128;
struct WalkTable {
  int nb, total;
  int start[kMaxWalk], pre[kMaxWalk];
};

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

// Called by the 64 lanes of wave 0; the table is read after a barrier.
__device__ __forceinline__ void walk_table_build(WalkTable &wt, const int *__restrict__ off, const Geom &g, int p,
                                                 int row0, int nrows) {
  const int lane = threadIdx.x;
  const int al = imod(p - g.kww + 1, g.P), as = imod(row0 - g.kw + 1, g.nf2);
  const int nl = tiles_count(al, g.kww, g.P, g.Pbin, g.nlayer);
  const int ns = tiles_count(as, nrows + g.kw - 1, g.nf2, g.T2, g.nty);
  int nb = nl * ns;
  if (nb > kMaxWalk) nb = kMaxWalk;  // the plan keeps nb within kMaxWalk
  const int l0 = al / g.Pbin, s0 = as / g.T2;
  int carry = 0;
  for (int base = 0; base < nb; base += 64) {
    const int k = base + lane;
    int st = 0, len = 0;
    if (k < nb) {
      const int j = k / ns, i = k - j * ns;
      int layer = l0 + j, strip = s0 + i;
      if (layer >= g.nlayer) layer -= g.nlayer;
      if (strip >= g.nty) strip -= g.nty;
      const int b = layer * g.ntiles + strip;
      st = off[b];
      len = off[b + 1] - st;
    }
    int x = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (k < nb) {
      wt.start[k] = st;
      wt.pre[k] = carry + x - len;
    }
    carry += __shfl(x, 63, 64);
  }
  if (lane == 0) {
    wt.nb = nb;
    wt.total = carry;
  }
}

// Record index of flat position t (0 <= t < total): the last range starting at or before t.
__device__ __forceinline__ int walk_record(const WalkTable &wt, int t) {
  int lo = 0, hi = wt.nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wt.pre[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return wt.start[lo] + (t - wt.pre[lo]);
}

// Copies a two-level twiddle table (hi | lo, fftc::twiddle_entries(n) entries)
// from global memory to LDS at dst; the caller's next barrier publishes it.
template <typename T>
__device__ __forceinline__ fftc::TwoLevel<T, const cpx<T> *> lds_twiddles(cpx<T> *dst,
                                                                         const cpx<T> *__restrict__ src, int n) {
  const int lb = fftc::twiddle_split(n), nhi = n >> lb, cnt = nhi + (1 << lb);
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
  return fftc::TwoLevel<T, const cpx<T> *>{dst, dst + nhi, lb};
}

template <typename T, int E, int MAXR = 16, typename Tw>
__device__ __forceinline__ void lds_fft(cpx<T> *buf, const fftc::Layout &L, int sign, const Tw tw, int n_tw) {
  cpx<T> v[E];
  const int nt = blockDim.x;
  for (int Ns = 1; Ns < L.n;) {
    const int R = fftc::stage_radix(L.n, Ns, MAXR);
    // opaque copy of the thread id: keeps each stage's address arithmetic inside
    // the stage (hoisted out of the loop it is held for all stages and spills)
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    if (R == 16)
      fftc::stage_load<T, E, 16>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else if (R == 8)
      fftc::stage_load<T, E, 8>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else if (R == 4)
      fftc::stage_load<T, E, 4>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    else
      fftc::stage_load<T, E, 2>(buf, v, L, Ns, sign, tw, n_tw, tid, nt);
    __syncthreads();
    if (R == 16)
      fftc::stage_store<T, E, 16>(buf, v, L, Ns, tid, nt);
    else if (R == 8)
      fftc::stage_store<T, E, 8>(buf, v, L, Ns, tid, nt);
    else if (R == 4)
      fftc::stage_store<T, E, 4>(buf, v, L, Ns, tid, nt);
    else
      fftc::stage_store<T, E, 2>(buf, v, L, Ns, tid, nt);
    __syncthreads();
    Ns *= R;
  }
}

// Stockham stages of `count` contiguous length-N sequences (buffer index
// pad(s * N + k)) with N a compile-time constant: every stage's radix, stride
// and twiddle step fold to constants, and the element indexing to shifts.
// (the twiddle table has TWM * N entries)
template <typename T, int E, int MAXR, int N, int NS, int TWM = 1, typename Tw>
__device__ __forceinline__ void fixed_stages(cpx<T> *buf, int count, int sign, const Tw &tw) {
  if constexpr (NS < N) {
    constexpr int R = (NS * MAXR <= N) ? MAXR : N / NS;
    const fftc::Layout L{N, count, N, 1, false};
    const int nt = blockDim.x;
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    cpx<T> v[E];
    fftc::stage_load<T, E, R>(buf, v, L, NS, sign, tw, N * TWM, tid, nt);
    __syncthreads();
    fftc::stage_store<T, E, R>(buf, v, L, NS, tid, nt);
    __syncthreads();
    fixed_stages<T, E, MAXR, N, NS * R, TWM>(buf, count, sign, tw);
  }
}

// FFT of one sequence of length m = 2^lgm (m = nf / S of a row split S ways)
// with the twiddle table of the full length nf = S m.
template <typename T, int E, int MAXR, int S, typename Tw>
__device__ __forceinline__ void lds_fft_part(cpx<T> *buf, int lgm, int sign, const Tw &tw) {
  switch (lgm) {
    case 10: fixed_stages<T, E, MAXR, 1024, 1, S>(buf, 1, sign, tw); break;
    case 11: fixed_stages<T, E, MAXR, 2048, 1, S>(buf, 1, sign, tw); break;
    case 12: fixed_stages<T, E, MAXR, 4096, 1, S>(buf, 1, sign, tw); break;
    default: fixed_stages<T, E, MAXR, 8192, 1, S>(buf, 1, sign, tw); break;
  }
}

// FFT of `count` contiguous sequences of length n = 2^lg (6 <= lg <= 14) in LDS.
template <typename T, int E, int MAXR, typename Tw>
__device__ __forceinline__ void lds_fft_rows(cpx<T> *buf, int lg, int count, int sign, const Tw &tw) {
  switch (lg) {
    case 6: fixed_stages<T, E, MAXR, 64, 1>(buf, count, sign, tw); break;
    case 7: fixed_stages<T, E, MAXR, 128, 1>(buf, count, sign, tw); break;
    case 8: fixed_stages<T, E, MAXR, 256, 1>(buf, count, sign, tw); break;
    case 9: fixed_stages<T, E, MAXR, 512, 1>(buf, count, sign, tw); break;
    case 10: fixed_stages<T, E, MAXR, 1024, 1>(buf, count, sign, tw); break;
    case 11: fixed_stages<T, E, MAXR, 2048, 1>(buf, count, sign, tw); break;
    case 12: fixed_stages<T, E, MAXR, 4096, 1>(buf, count, sign, tw); break;
    case 13: fixed_stages<T, E, MAXR, 8192, 1>(buf, count, sign, tw); break;
    default: fixed_stages<T, E, MAXR, 16384, 1>(buf, count, sign, tw); break;
  }
}

__device__ __forceinline__ int image_index(int k, int N, int nf) {
  // grid index (FFT order) -> image index (centred), or -1 outside the image
  const int h = N / 2;
  return k < N - h ? k + h : (k >= nf - h ? k - nf + h : -1);
}
__device__ __forceinline__ int grid_index(int i, int N, int nf) {
  const int h = N / 2;
  return i >= h ? i - h : nf + i - h;
}

// Spread of one record into the workgroup's grid rows [row0, row0 + rpw) of plane p (LDS).
template <typename T, int KWM>
__device__ __forceinline__ void spread_rows(const Rec<T> &rc, T *ldsf, const Geom &g, int p, int row0,
                                            const Ker<T> &K) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2, rpw = g.rpw;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int ry = rc.iy - row0;
  if (ry < 0) ry += nf2;
  if (!(ry < rpw || ry + kw - 1 >= nf2)) return;
  const T vw = K.w(rc.dw, k);
  const T vre = rc.re * vw, vim = rc.im * vw;
  T kx[KWM];
  K.template taps<KWM>(rc.dx, kx);
  for (int jj = 0; jj < kw; ++jj) {
    int cy = ry + jj;
    if (cy >= nf2) cy -= nf2;
    if (cy >= rpw) continue;
    const T ky = K(rc.dy, jj);
    const T are = vre * ky, aim = vim * ky;
    const int base = cy * nf1;
#pragma unroll
    for (int i = 0; i < KWM; ++i) {
      if (i < kw) {
        int cx = rc.ix + i;
        if (cx >= nf1) cx -= nf1;
        const int a = 2 * fftc::pad(base + cx);
        atomicAdd(ldsf + a, are * kx[i]);
        atomicAdd(ldsf + a + 1, aim * kx[i]);
      }
    }
  }
}

// Interpolation of one record from the workgroup's rows of plane p (LDS): its
// partial sum over those rows, added to the record's FP64 accumulator.
template <typename T, int KWM>
__device__ __forceinline__ void interp_rows(const Rec<T> &rc, int idx, const cpx<T> *lds, const Geom &g, int p,
                                            int row0, double2 *__restrict__ vacc, const Ker<T> &K) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2, rpw = g.rpw;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int ry = rc.iy - row0;
  if (ry < 0) ry += nf2;
  if (!(ry < rpw || ry + kw - 1 >= nf2)) return;
  T kx[KWM];
  K.template taps<KWM>(rc.dx, kx);
  T sr = 0, si = 0;
  for (int jj = 0; jj < kw; ++jj) {
    int cy = ry + jj;
    if (cy >= nf2) cy -= nf2;
    if (cy >= rpw) continue;
    const T ky = K(rc.dy, jj);
    const int base = cy * nf1;
    T rr = 0, ri = 0;
#pragma unroll
    for (int i = 0; i < KWM; ++i) {
      if (i < kw) {
        int cx = rc.ix + i;
        if (cx >= nf1) cx -= nf1;
        const cpx<T> f = lds[fftc::pad(base + cx)];
        rr += f.x * kx[i];
        ri += f.y * kx[i];
      }
    }
    sr += rr * ky;
    si += ri * ky;
  }
  const T kz = K.w(rc.dw, k);
  atomicAdd(&vacc[idx].x, (double)(sr * kz));
  atomicAdd(&vacc[idx].y, (double)(si * kz));
}


// The intermediate planes live in a column-tiled layout: element (plane k,
// grid row k2, column slot s) at ((k * nct + s / kTile) * nf2 + k2) * kTile
// + s % kTile, so a row's kTile-column chunks are contiguous for the row
// kernels and a column's elements are kTile apart for the column kernels (the
// kTile columns of a tile run on one XCD, whose L2 serves the shared sectors).
constexpr int kTile = 1 << kTileLog2;

// Column slots of the tiled layout.  With whole rows the slot of image column
// col is col.  With rows split S ways, the columns of part h (grid index h mod
// S, col = S i + c_h with c_h = (h + N1/2) mod S) take slots h M + i, M = nf1 /
// (2 S) >= ceil(N1 / S), so a part writes whole kTile-column chunks; slots whose
// i is past the part's last column are unused (-1 here).
__device__ __forceinline__ int slot_column(int s, const Geom &g) {
  if (g.rsplit <= 1) return s < g.N1 ? s : -1;
  if (s >= g.nslot) return -1;
  const int S = g.rsplit, h = s >> g.slot_lg, i = s & ((1 << g.slot_lg) - 1);
  const int col = S * i + ((h + g.N1 / 2) & (S - 1));
  return col < g.N1 ? col : -1;
}

// whether a column workgroup's slots [s0, s0 + cpw) hold any image column
__device__ __forceinline__ bool slots_used(int s0, const Geom &g) {
  for (int c = 0; c < g.cpw; ++c)
    if (slot_column(s0 + c, g) >= 0) return true;
  return false;
}

// Workgroup id -> work item such that each XCD (workgroups are dispatched to
// the 8 XCDs round robin) takes one contiguous range of items.
__device__ __forceinline__ int xcd_group(int bid, int n) {
  const int q = n >> 3, r = n & 7, x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

template <typename T> struct MaxThreads { static constexpr int value = sizeof(T) == 8 ? 512 : 1024; };

// ms2dirty: spread into full grid rows held in LDS, then the row FFT (length
// nf1), keeping the N1 image columns, written to the column-tiled layout H.  One
// workgroup owns rpw rows of one w plane (rpw * nf1 = E * blockDim.x); rows are
// binned into full-width strips of T2 rows.
template <typename T, int E, int KWM>
__global__ void __launch_bounds__(MaxThreads<T>::value) spread_row_fft_kernel(const Rec<T> *__restrict__ recs,
                                                                             const int *__restrict__ off,
                                                                             cpx<T> *__restrict__ H,
                                                                             const cpx<T> *__restrict__ tw, Geom g,
                                                                             int p0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  T *ldsf = reinterpret_cast<T *>(smem_raw);
  const int rpw = g.rpw, nf1 = g.nf1, nt = blockDim.x;
  // rows in XCD-contiguous order: the rows sharing a tile's lines are written
  // through one L2, which merges their chunks into whole lines
  const int row0 = xcd_group(blockIdx.x, gridDim.x) * rpw, pl = blockIdx.y, p = p0 + pl;
  const int ntot = fftc::padded_size(rpw * nf1);
  const auto tv = lds_twiddles<T>(lds + ntot, tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, rpw);
  for (int i = threadIdx.x; i < ntot; i += nt) lds[i] = fftc::mk((T)0, (T)0);
  __syncthreads();
  const Ker<T> K = make_ker<T>(g);
  const int total = wt.total;
  const int N1 = g.N1, nout = rpw * N1, h1 = N1 / 2;
  const int lgt = 31 - __clz(g.nf2) + kTileLog2;  // log2(nf2 * kTile)
  cpx<T> *hb = H + ((size_t)pl * g.nct * g.nf2 + row0) * kTile;
  if (total == 0) {
    // no visibility reaches these rows of this plane: the transform is zero
    for (int i = threadIdx.x; i < nout; i += nt) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      hb[((col >> kTileLog2) << lgt) + (r << kTileLog2) + (col & (kTile - 1))] = fftc::mk((T)0, (T)0);
    }
    return;
  }
  for (int t0 = threadIdx.x; t0 < total; t0 += 2 * nt) {
    const int t1 = t0 + nt;
    const Rec<T> ra = recs[walk_record(wt, t0)];
    Rec<T> rb;
    if (t1 < total) rb = recs[walk_record(wt, t1)];
    spread_rows<T, KWM>(ra, ldsf, g, p, row0, K);
    if (t1 < total) spread_rows<T, KWM>(rb, ldsf, g, p, row0, K);
  }
  __syncthreads();
  lds_fft_rows<T, E, 16>(lds, 31 - __clz(nf1), rpw, +1, tv);
  // pruned write: all LDS reads first, then the stores (kTile-column chunks);
  // offsets from the row's base in 32 bits (a plane holds < 2^31 elements)
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  cpx<T> o[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      const int k1 = col >= h1 ? col - h1 : nf1 + col - h1;
      o[e] = lds[fftc::pad(r * nf1 + k1)];
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int r = rpw == 1 ? 0 : i / N1, col = i - r * N1;
      hb[((col >> kTileLog2) << lgt) + (r << kTileLog2) + (col & (kTile - 1))] = o[e];
    }
  }
}

// dirty2ms: row pass of the column-transformed planes (H as col_inv_kernel
// writes it; each row expanded to nf1 cells, zero outside the image columns),
// inverse row FFT in LDS, then interpolation of every record touching the workgroup's rows of
// plane p: per record a partial sum, added to its FP64 accumulator (atomics),
// so the full grid planes are never written.
template <typename T, int E, int KWM>
__global__ void __launch_bounds__(MaxThreads<T>::value) row_ifft_interp_kernel(const cpx<T> *__restrict__ H,
                                                                              const Rec<T> *__restrict__ recs,
                                                                              const int *__restrict__ off,
                                                                              double2 *__restrict__ vacc,
                                                                              const cpx<T> *__restrict__ tw,
                                                                              Geom g, int p0, int sign) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int rpw = g.rpw, nf1 = g.nf1, N1 = g.N1, nt = blockDim.x;
  // rows in XCD-contiguous order: the 16 rows sharing a column line run together
  const int row0 = xcd_group(blockIdx.x, gridDim.x) * rpw, pl = blockIdx.y, p = p0 + pl;
  const int lg = 31 - __clz(nf1), lg2 = 31 - __clz(g.nf2);
  cpx<T> v[E];
  const cpx<T> *hb = H + ((size_t)pl * g.nct * g.nf2 + row0) * kTile;
  const int lgt = lg2 + kTileLog2;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = threadIdx.x + e * nt;  // < rpw * nf1 = E * nt
    const int r = i >> lg, k1 = i & (nf1 - 1);
    const int col = image_index(k1, N1, nf1);
    v[e] = col >= 0 ? hb[((col >> kTileLog2) << lgt) + (r << kTileLog2) + (col & (kTile - 1))]
                    : fftc::mk((T)0, (T)0);
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(rpw * nf1), tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, rpw);
#pragma unroll
  for (int e = 0; e < E; ++e) lds[fftc::pad(threadIdx.x + e * nt)] = v[e];
  __syncthreads();
  const int total = wt.total;
  if (total == 0) return;  // no visibility reaches these rows of this plane
  // first record of this thread, loaded while the FFT runs
  int idx0 = 0;
  Rec<T> r0;
  if (threadIdx.x < total) {
    idx0 = walk_record(wt, threadIdx.x);
    r0 = recs[idx0];
  }
  lds_fft_rows<T, E, 16>(lds, lg, rpw, sign, tv);
  const Ker<T> K = make_ker<T>(g);
  if (threadIdx.x < total) interp_rows<T, KWM>(r0, idx0, lds, g, p, row0, vacc, K);
  for (int t = threadIdx.x + nt; t < total; t += nt) {
    const int idx = walk_record(wt, t);
    interp_rows<T, KWM>(recs[idx], idx, lds, g, p, row0, vacc, K);
  }
}

// Spread of one record into part h of a row split S ways (ms2dirty): part h
// holds z_h[n mod m] = sum over the row's cells n of x[n] w^(sign h n), whose
// m-point FFT gives the row transform at the indices S k + h (radix-S
// decimation in frequency; w = e^{2 pi i / nf1}, m = nf1 / S).
template <typename T, int KWM, typename Tw>
__device__ __forceinline__ void spread_row_part(const Rec<T> &rc, T *ldsf, const Tw &tv, const Geom &g, int p,
                                                int row0, int h, int mm, int sign, const Ker<T> &K) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int jj = row0 - rc.iy;
  if (jj < 0) jj += nf2;
  if (jj >= kw) return;
  const T a = K.w(rc.dw, k) * K(rc.dy, jj);
  const T vre = rc.re * a, vim = rc.im * a;
  T kxs[KWM];
  K.template taps<KWM>(rc.dx, kxs);
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    if (i < kw) {
      int cx = rc.ix + i;
      if (cx >= nf1) cx -= nf1;
      const T kx = kxs[i];
      T re = vre * kx, im = vim * kx;
      if (h) {
        const cpx<T> w = tv[(h * cx) & (nf1 - 1)];
        const T ws = (T)sign * w.y, t = re * w.x - im * ws;
        im = re * ws + im * w.x;
        re = t;
      }
      const int q = 2 * fftc::pad(cx & mm);
      atomicAdd(ldsf + q, re);
      atomicAdd(ldsf + q + 1, im);
    }
  }
}

// ms2dirty row pass for rows too long for two workgroups per CU: one workgroup
// per part h of a row split S ways (the row transform's outputs S k + h, see
// spread_row_part), each spreading the row's visibilities into an m-point LDS
// buffer, transforming it and writing its share of the N1 image columns (those
// whose grid index is h mod S).  Several such workgroups fit one CU, so one's
// memory phases overlap another's FFT.
template <typename T, int E, int KWM, int S>
__global__ void __launch_bounds__(MaxThreads<T>::value / S) spread_row_fft_part_kernel(
    const Rec<T> *__restrict__ recs, const int *__restrict__ off, cpx<T> *__restrict__ H,
    const cpx<T> *__restrict__ tw, Geom g, int p0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  T *ldsf = reinterpret_cast<T *>(smem_raw);
  const int nf1 = g.nf1, m = nf1 / S, nt = blockDim.x, sign = +1;
  // the parts of a row are consecutive items: same XCD, adjacent in time
  const int item = xcd_group(blockIdx.x, gridDim.x);
  const int row0 = item / S, h = item % S, pl = blockIdx.y, p = p0 + pl;
  const int ntot = fftc::padded_size(m);
  const auto tv = lds_twiddles<T>(lds + ntot, tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, 1);
  for (int i = threadIdx.x; i < ntot; i += nt) lds[i] = fftc::mk((T)0, (T)0);
  __syncthreads();
  const int total = wt.total;
  // this part's image columns: grid_index(col) = col - N1/2 (mod nf1) = h (mod S)
  const int N1 = g.N1, h1 = N1 / 2, c0 = (h + h1) % S, nout = (N1 - c0 + S - 1) / S;
  const int lgt = 31 - __clz(g.nf2) + kTileLog2;
  cpx<T> *hb = H + ((size_t)pl * g.nct * g.nf2 + row0) * kTile;
  const int hs = h << g.slot_lg;  // this part's first column slot
  if (total == 0) {
    for (int i = threadIdx.x; i < nout; i += nt) {
      const int sl = hs + i;
      hb[((sl >> kTileLog2) << lgt) + (sl & (kTile - 1))] = fftc::mk((T)0, (T)0);
    }
    return;
  }
  const Ker<T> K = make_ker<T>(g);
  for (int t0 = threadIdx.x; t0 < total; t0 += 2 * nt) {
    const int t1 = t0 + nt;
    const Rec<T> ra = recs[walk_record(wt, t0)];
    Rec<T> rb;
    if (t1 < total) rb = recs[walk_record(wt, t1)];
    spread_row_part<T, KWM>(ra, ldsf, tv, g, p, row0, h, m - 1, sign, K);
    if (t1 < total) spread_row_part<T, KWM>(rb, ldsf, tv, g, p, row0, h, m - 1, sign, K);
  }
  __syncthreads();
  lds_fft_part<T, E, 16, S>(lds, 31 - __clz(m), sign, tv);
  int tid = threadIdx.x;
  __asm__ volatile("" : "+v"(tid));
  constexpr int NO = E / 2;  // nout <= m / 2 = (E / 2) nt, since N1 <= nf1 / 2
  cpx<T> o[NO];
#pragma unroll
  for (int e = 0; e < NO; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int col = S * i + c0;
      const int k1 = col >= h1 ? col - h1 : nf1 + col - h1;
      o[e] = lds[fftc::pad(k1 / S)];
    }
  }
#pragma unroll
  for (int e = 0; e < NO; ++e) {
    const int i = tid + e * nt;
    if (i < nout) {
      const int sl = hs + i;
      hb[((sl >> kTileLog2) << lgt) + (sl & (kTile - 1))] = o[e];
    }
  }
}

// Interpolation of one record from part h of a row split S ways: the row's
// cells are x[n] = sum_h w^(sign h n) Y_h[n mod m] (w = e^{2 pi i / nf1}, m = nf1 / S),
// and part h holds Y_h; its share of the record's sum is added to the record's
// FP64 accumulator.
template <typename T, int KWM, typename Tw>
__device__ __forceinline__ void interp_row_part(const Rec<T> &rc, int idx, const cpx<T> *lds, const Tw &tv,
                                                const Geom &g, int p, int row0, int h, int mm, int sign,
                                                double2 *__restrict__ vacc, const Ker<T> &K) {
  const int kw = g.kw, nf1 = g.nf1, nf2 = g.nf2;
  int k = p - rc.iw;
  if (k < 0) k += g.P;
  if (k >= g.kww) return;
  int jj = row0 - rc.iy;
  if (jj < 0) jj += nf2;
  if (jj >= kw) return;
  T rr = 0, ri = 0;
  T kxs[KWM];
  K.template taps<KWM>(rc.dx, kxs);
#pragma unroll
  for (int i = 0; i < KWM; ++i) {
    if (i < kw) {
      int cx = rc.ix + i;
      if (cx >= nf1) cx -= nf1;
      const T kx = kxs[i];
      cpx<T> f = lds[fftc::pad(cx & mm)];
      if (h) {
        const cpx<T> w = tv[(h * cx) & (nf1 - 1)];
        f = fftc::mulw(f, w.x, (T)sign * w.y);
      }
      rr += f.x * kx;
      ri += f.y * kx;
    }
  }
  const T kyz = K(rc.dy, jj) * K.w(rc.dw, k);
  atomicAdd(&vacc[idx].x, (double)(rr * kyz));
  atomicAdd(&vacc[idx].y, (double)(ri * kyz));
}

// dirty2ms row pass for rows too long for two workgroups per CU: each row is
// split S ways by the index of its inverse-transform input, k = S k' + h
// (radix-S decimation in frequency), one workgroup per part: Y_h = FFT_m over
// X[S k' + h], so the parts of a row load disjoint columns of H, run m-point
// FFTs and add their shares to each record's accumulator.
template <typename T, int E, int KWM, int S>
__global__ void __launch_bounds__(MaxThreads<T>::value / S) row_ifft_interp_part_kernel(
    const cpx<T> *__restrict__ H, const Rec<T> *__restrict__ recs, const int *__restrict__ off,
    double2 *__restrict__ vacc, const cpx<T> *__restrict__ tw, Geom g, int p0, int sign) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ WalkTable wt;
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int nf1 = g.nf1, m = nf1 / S, N1 = g.N1, nt = blockDim.x;
  const int item = xcd_group(blockIdx.x, gridDim.x);
  const int row0 = item / S, h = item % S, pl = blockIdx.y, p = p0 + pl;
  const int lgm = 31 - __clz(m), lgt = 31 - __clz(g.nf2) + kTileLog2;
  const cpx<T> *hb = H + ((size_t)pl * g.nct * g.nf2 + row0) * kTile;
  cpx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = threadIdx.x + e * nt;  // < m = E * nt
    const int col = image_index(S * i + h, N1, nf1), sl = (h << g.slot_lg) + col / S;
    v[e] = col >= 0 ? hb[((sl >> kTileLog2) << lgt) + (sl & (kTile - 1))] : fftc::mk((T)0, (T)0);
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(m), tw, nf1);
  if (threadIdx.x < 64) walk_table_build(wt, off, g, p, row0, 1);
#pragma unroll
  for (int e = 0; e < E; ++e) lds[fftc::pad(threadIdx.x + e * nt)] = v[e];
  __syncthreads();
  const int total = wt.total;
  if (total == 0) return;
  int idx0 = 0;
  Rec<T> r0;
  if (threadIdx.x < total) {
    idx0 = walk_record(wt, threadIdx.x);
    r0 = recs[idx0];
  }
  lds_fft_part<T, E, 16, S>(lds, lgm, sign, tv);
  const Ker<T> K = make_ker<T>(g);
  if (threadIdx.x < total)
    interp_row_part<T, KWM>(r0, idx0, lds, tv, g, p, row0, h, m - 1, sign, vacc, K);
  for (int t = threadIdx.x + nt; t < total; t += nt) {
    const int idx = walk_record(wt, t);
    interp_row_part<T, KWM>(recs[idx], idx, lds, tv, g, p, row0, h, m - 1, sign, vacc, K);
  }
}

// z' (n - 1 - D) gamma h of image pixel (row, col), cuft.cu:53-92 pre-stage
__device__ __forceinline__ double pixel_zp(int row, int col, const Geom &g) {
  double nm1, n;
  pixel_n(row, col, g, nm1, n);
  return (nm1 - g.D) * g.gamma * g.h;
}

// ms2dirty column pass fused with the w-accumulation (w_term_dft, cuft.cu:520-572):
// a workgroup owns cpw grid columns (cpw * nf2 = E * blockDim.x) and walks the
// batch's nb planes: load the row-transformed columns, FFT over nf2, then for
// the N2 image rows accT[col][row] += F_p e^{+i z' (p - P/2)}: the phase
// argument and its reduction in FP64 (sin / cos of the reduced argument at the
// planes' precision) and the sum over planes in FP64 registers.  Image rows per thread:
// cpw * N2 <= 8 * blockDim.x (the plan requires 2 N2 <= nf2).
template <typename T, int E>
__global__ void __launch_bounds__(MaxThreads<T>::value) col_fwd_kernel(const cpx<T> *__restrict__ H,
                                                                      double2 *__restrict__ accT,
                                                                      const cpx<T> *__restrict__ tw, Geom g,
                                                                      int p0, int nb, int first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int cpw = g.cpw, nf2 = g.nf2, N2 = g.N2, nt = blockDim.x;
  const int s0 = xcd_group(blockIdx.x, gridDim.x) * cpw;  // first column slot
  if (!slots_used(s0, g)) return;
  const int lg = 31 - __clz(nf2), lgt = lg + kTileLog2;
  const int nout = cpw * N2;
  double ar[8], ai[8], zp[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    ar[e] = ai[e] = zp[e] = 0.0;
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) zp[e] = pixel_zp(row, col, g);
    }
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(cpw * nf2), tw, nf2);
  for (int k = 0; k < nb; ++k) {
    // opaque copy of the thread id: the per-plane addresses are formed inside
    // the loop rather than held in registers across it
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    const cpx<T> *hk = H + (size_t)k * g.nct * nf2 * kTile;
    cpx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + e * nt;
      const int ci = i >> lg, k2 = i & (nf2 - 1), sl = s0 + ci;
      v[e] = slot_column(sl, g) >= 0 ? hk[((sl >> kTileLog2) << lgt) + (k2 << kTileLog2) + (sl & (kTile - 1))]
                                     : fftc::mk((T)0, (T)0);
    }
    __syncthreads();  // the previous plane's LDS reads are done
#pragma unroll
    for (int e = 0; e < E; ++e) lds[fftc::pad(tid + e * nt)] = v[e];
    __syncthreads();
    lds_fft_rows<T, E, 16>(lds, lg, cpw, +1, tv);
    const double m = (double)(p0 + k - g.P / 2);
    int tid2 = threadIdx.x;
    __asm__ volatile("" : "+v"(tid2));
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int o = tid2 + e * nt;
      if (o < nout) {
        const int ci = o / N2, row = o - ci * N2;
        const cpx<T> f = lds[fftc::pad(ci * nf2 + grid_index(row, N2, nf2))];
        double sn, cs;
        plane_phase<T>(zp[e] * m, &sn, &cs);
        const double fr = (double)f.x, fi = (double)f.y;
        ar[e] += fr * cs - fi * sn;
        ai[e] += fr * sn + fi * cs;
      }
      // one element's phase at a time (bounds the live FP64 temporaries)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) {
        const size_t a = (size_t)col * N2 + row;
        double2 r = make_double2(ar[e], ai[e]);
        if (!first) {
          const double2 q = accT[a];
          r.x += q.x;
          r.y += q.y;
        }
        accT[a] = r;
      }
    }
  }
}

// dirty2ms column pass fused with the w-fill (w_term_idft, cuft.cu:575-612): a
// workgroup owns cpw grid columns and, for each of the batch's nb planes, places
// gT[col][row] e^{-i z' (p - P/2)} (FP64 phase argument) at the image rows of its columns
// (zeros elsewhere), runs the inverse FFT over nf2 and writes the full columns
// to the column-tiled H.
template <typename T, int E>
__global__ void __launch_bounds__(MaxThreads<T>::value) col_inv_kernel(const cpx<T> *__restrict__ gT,
                                                                      cpx<T> *__restrict__ H,
                                                                      const cpx<T> *__restrict__ tw, Geom g,
                                                                      int p0, int nb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
  const int cpw = g.cpw, nf2 = g.nf2, N2 = g.N2, nt = blockDim.x;
  const int s0 = xcd_group(blockIdx.x, gridDim.x) * cpw;  // first column slot
  if (!slots_used(s0, g)) return;
  const int lg = 31 - __clz(nf2);
  const int nout = cpw * N2, nzero = cpw * (nf2 - N2), h2 = N2 / 2;
  cpx<T> gv[8];
  double zp[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    gv[e] = fftc::mk((T)0, (T)0);
    zp[e] = 0.0;
    const int o = threadIdx.x + e * nt;
    if (o < nout) {
      const int ci = o / N2, row = o - ci * N2, col = slot_column(s0 + ci, g);
      if (col >= 0) {
        gv[e] = gT[(size_t)col * N2 + row];
        zp[e] = pixel_zp(row, col, g);
      }
    }
  }
  const auto tv = lds_twiddles<T>(lds + fftc::padded_size(cpw * nf2), tw, nf2);
  for (int k = 0; k < nb; ++k) {
    const double m = (double)(p0 + k - g.P / 2);
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    __syncthreads();  // the previous plane's LDS reads are done
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int o = tid + e * nt;
      if (o < nout) {
        const int ci = o / N2, row = o - ci * N2;
        double sn, cs;
        plane_phase<T>(-zp[e] * m, &sn, &cs);
        const double gr = (double)gv[e].x, gi = (double)gv[e].y;
        lds[fftc::pad(ci * nf2 + grid_index(row, N2, nf2))] =
            fftc::mk((T)(gr * cs - gi * sn), (T)(gr * sn + gi * cs));
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int z = tid; z < nzero; z += nt) {
      const int ci = z / (nf2 - N2), j = z - ci * (nf2 - N2);
      lds[fftc::pad(ci * nf2 + N2 - h2 + j)] = fftc::mk((T)0, (T)0);
    }
    __syncthreads();
    lds_fft_rows<T, E, 16>(lds, lg, cpw, -1, tv);
    int tid2 = threadIdx.x;
    __asm__ volatile("" : "+v"(tid2));
    cpx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[fftc::pad(tid2 + e * nt)];
    cpx<T> *hk = H + (size_t)k * g.nct * nf2 * kTile;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid2 + e * nt;
      const int ci = i >> lg, k2 = i & (nf2 - 1), sl = s0 + ci;
      if (slot_column(sl, g) >= 0) hk[((sl >> kTileLog2) << (lg + kTileLog2)) + (k2 << kTileLog2) + (sl & (kTile - 1))] = v[e];
    }
  }
}

// dirty2ms prologue for the LDS path: gT[col][row] (precision T, column-major) =
// dirty / n e^{-i C (z - D)} / (phihat1 phihat2 phihat3)
template <typename TI, typename T>
__global__ void __launch_bounds__(256) prologue_cols_kernel(const void *__restrict__ dirty, cpx<T> *__restrict__ gT,
                                                             Geom g, const double *__restrict__ fk1,
                                                             const double *__restrict__ fk2,
                                                             const double *__restrict__ q) {
  const long npix = (long)g.N1 * g.N2;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < npix; idx += (long)gridDim.x * blockDim.x) {
    const int col = (int)(idx / g.N2), row = (int)(idx - (long)col * g.N2);
    double scale, phase, s, c;
    pixel_correction(row, col, g, fk1, fk2, q, scale, phase);
    sincos(-phase, &s, &c);
    const double2 a = load_c<TI>(dirty, (long)row * g.N1 + col);
    gT[idx] = fftc::mk((T)((a.x * c - a.y * s) * scale), (T)((a.x * s + a.y * c) * scale));
  }
}

// ==========================================================================
// host side
// ==========================================================================
namespace {

// Grow-only device buffers, one set per device (calls on one device are serialised).
enum Slot {
  S_PART, S_CNT, S_HIST, S_OFF, S_BSUM, S_BINRANK, S_RECS, S_PERM, S_GRID, S_ACC, S_VACC, S_TAB, S_FFTW,
  S_H, S_TW, S_HC,
  S_NSLOT
};
struct Workspace {
  void *p[S_NSLOT] = {};
  size_t n[S_NSLOT] = {};
  int tw_key[3] = {-1, -1, -1};  // (nf1, nf2, sizeof(T)) of the twiddles held in S_TW
};
std::mutex g_ws_mu;
std::vector<Workspace> g_ws;

int ws_get(int dev, Slot s, size_t bytes, void **out) {
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

inline int imod_host(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

inline int nblocks(long n, int bs = 256, long cap = 1 << 20) {
  long b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
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
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
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
  g.rpw = g.cpw = g.nct = g.nslot = g.slot_lg = g.lds_row_bytes = g.lds_col_bytes = g.rsplit = g.lds_rowp_bytes = 0;
  auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
  const size_t cs = dbl ? 16 : 8;
  const size_t lds_max = tn.lds_bytes > 0 ? (size_t)tn.lds_bytes : size_t(160) << 10;
  const int max_threads = dbl ? 512 : 1024;
  if (tn.fft != 1 && pow2(g.nf1) && pow2(g.nf2) && g.nf1 >= 64 && g.nf2 >= 64 && g.nf1 <= 16384 &&
      g.nf2 <= 16384 && 2 * g.N1 <= g.nf1 && 2 * g.N2 <= g.nf2) {
    const int rpw = std::max(1, std::min(4096 / g.nf1, g.nf2));
    const int cpw = std::max(1, std::min(4096 / g.nf2, 16));
    // the row kernels' static LDS (bin walk table) comes on top of the row
    const size_t rb = (size_t)(fftc::padded_size(rpw * g.nf1) + fftc::twiddle_entries(g.nf1)) * cs;
    const size_t cb = (size_t)(fftc::padded_size(cpw * g.nf2) + fftc::twiddle_entries(g.nf2)) * cs;
    if (rb + 2048 <= lds_max && cb <= lds_max && rpw * g.nf1 / 16 <= max_threads &&
        cpw * g.nf2 / 16 <= max_threads && rpw * g.nf1 >= 16 * 64 && cpw * g.nf2 >= 16) {
      g.fft_lds = 1;
      g.rpw = rpw;
      g.cpw = cpw;
      g.nslot = g.N1;
      g.lds_row_bytes = (int)rb;
      g.lds_col_bytes = (int)cb;
      // rows too long for two workgroups per CU run split kRowSplit ways
      if (rpw == 1 && rb + 2048 > lds_max / 2 && g.nf1 >= 4096) {
        g.rsplit = kRowSplit;
        g.nslot = g.nf1 / 2;
        g.slot_lg = 31 - __builtin_clz(g.nf1 / (2 * kRowSplit));
        g.lds_rowp_bytes =
            (int)((size_t)(fftc::padded_size(g.nf1 / kRowSplit) + fftc::twiddle_entries(g.nf1)) * cs);
      }
      g.nct = (g.nslot + kTile - 1) / kTile;
    }
  }
  // planes per batch from the memory budget: full planes (rocFFT path) or, on
  // the LDS path, the compact plane plus the one or two column-pass buffers
  size_t per_plane = (size_t)g.nf1 * g.nf2 * cs;
  if (g.fft_lds) per_plane = (size_t)g.nf2 * g.nct * kTile * cs;
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
  double2 *acc = nullptr, *vacc = nullptr;
  const void *tw1 = nullptr, *tw2 = nullptr;
  const double *hc_d = nullptr;  // piecewise-polynomial kernel coefficients, FP64 | FP32
  const float *hc_f = nullptr;
};

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
      w_minmax_kernel<<<nbm[i], 256, 0, st>>>(pbs[i].uvw, pbs[i].nrow, pbs[i].flag, pbs[i].flag_stride,
                                               part + 2 * base[i], cnt + base[i]);
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
  if (key[0] != g.nf1 || key[1] != g.nf2 || key[2] != (int)sizeof(T)) {
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
    key[0] = g.nf1;
    key[1] = g.nf2;
    key[2] = (int)sizeof(T);
  }
  tw1 = t;
  tw2 = t + n1;
  return kOk;
}

template <typename K>
int allow_lds(K kernel, int bytes) {
  if (bytes > 65536) CURIG_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  return kOk;
}

template <typename T, int S>
int allow_parts(const Geom &g) {
  CURIG_TRY(allow_lds(spread_row_fft_part_kernel<T, 16, 8, S>, g.lds_rowp_bytes));
  CURIG_TRY(allow_lds(spread_row_fft_part_kernel<T, 16, 16, S>, g.lds_rowp_bytes));
  CURIG_TRY(allow_lds(spread_row_fft_part_kernel<T, 16, kMaxKernelWidth, S>, g.lds_rowp_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_part_kernel<T, 16, 8, S>, g.lds_rowp_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_part_kernel<T, 16, 16, S>, g.lds_rowp_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_part_kernel<T, 16, kMaxKernelWidth, S>, g.lds_rowp_bytes));
  return kOk;
}

template <typename T>
int allow_kernels(const Geom &g) {
  if (!g.fft_lds) return kOk;
  CURIG_TRY(allow_lds(spread_row_fft_kernel<T, 16, 8>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(spread_row_fft_kernel<T, 16, 16>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(spread_row_fft_kernel<T, 16, kMaxKernelWidth>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_kernel<T, 16, 8>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_kernel<T, 16, 16>, g.lds_row_bytes));
  CURIG_TRY(allow_lds(row_ifft_interp_kernel<T, 16, kMaxKernelWidth>, g.lds_row_bytes));
  if (g.rsplit == kRowSplit) CURIG_TRY((allow_parts<T, kRowSplit>(g)));
  CURIG_TRY(allow_lds(col_fwd_kernel<T, 16>, g.lds_col_bytes));
  CURIG_TRY(allow_lds(col_inv_kernel<T, 16>, g.lds_col_bytes));
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
    if (g.fft_lds) h = std::max(h, (size_t)g.nf2 * g.nct * kTile * g.Pb);
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
    const cpx<T> *t1, *t2;
    CURIG_TRY(twiddles<T>(dev, g0, st, t1, t2));
    b.tw1 = t1;
    b.tw2 = t2;
    CURIG_TRY(allow_kernels<T>(g0));
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

// Phase 3: binning of one problem (kernels only).
template <typename T>
void enqueue_binning(const Problem &pb, const ChanPlan &cp, const Bufs &b, hipStream_t st, bool with_value) {
  const Geom &g = cp.g;
  const long nbins = (long)g.nlayer * g.ntiles;
  const long nscan = (nbins + kScanTile - 1) / kScanTile;
  fill_int_kernel<<<nblocks(nbins), 256, 0, st>>>(b.hist, nbins, 0);
  if (pb.nrow > 0)
    bin_count_kernel<<<nblocks(pb.nrow), 256, 0, st>>>(pb.uvw, pb.nrow, pb.flag, pb.flag_stride, g, b.hist,
                                                       b.binrank);
  scan_reduce_kernel<<<(int)nscan, kScanBlock, 0, st>>>(b.hist, nbins, b.bsum);
  scan_sums_kernel<<<1, 1024, 0, st>>>(b.bsum, (int)nscan);
  scan_apply_kernel<<<(int)nscan, kScanBlock, 0, st>>>(b.hist, nbins, b.bsum, b.off, (int)cp.nvalid);
  if (pb.nrow > 0)
    bin_scatter_kernel<T><<<nblocks(pb.nrow), 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt,
                                                             pb.wgt_stride, g, b.binrank, b.off, (Rec<T> *)b.recs,
                                                             b.perm, with_value ? 1 : 0);
}

template <typename T, int KWM>
void launch_spread(const Rec<T> *recs, const int *off, v2_t<T> *grid, const Geom &g, int p0, int nb,
                   hipStream_t st) {
  dim3 grd(g.ntiles, nb);
  size_t lds = sizeof(T) * 2 * g.T1 * g.T2;
  spread_tile_kernel<T, KWM><<<grd, 256, lds, st>>>(recs, off, grid, g, p0);
}

template <typename T>
void spread_dispatch(const Rec<T> *recs, const int *off, v2_t<T> *grid, const Geom &g, int p0, int nb,
                     hipStream_t st) {
  if (g.kw <= 8)
    launch_spread<T, 8>(recs, off, grid, g, p0, nb, st);
  else if (g.kw <= 16)
    launch_spread<T, 16>(recs, off, grid, g, p0, nb, st);
  else
    launch_spread<T, kMaxKernelWidth>(recs, off, grid, g, p0, nb, st);
}

template <typename T>
void interp_dispatch(const Rec<T> *recs, const int *off, const v2_t<T> *grid, double2 *vacc, const Geom &g,
                     long nvalid, int p0, int nb, hipStream_t st) {
  if (nvalid <= 0) return;
  // the kernel reads its record ranges from the bin offsets; a batch touches at
  // most the records of (nb + kw - 1) planes' layers, rounded up to whole layers
  const double frac = std::min(1.0, (double)(nb + g.kww - 1 + 2 * g.Pbin) / g.P);
  const int nbk = std::max(1, std::min(nblocks((long)(nvalid * frac) + 1), 4096));
  if (g.kw <= 8)
    interp_kernel<T, 8><<<nbk, 256, 0, st>>>(recs, off, grid, vacc, g, p0, nb);
  else if (g.kw <= 16)
    interp_kernel<T, 16><<<nbk, 256, 0, st>>>(recs, off, grid, vacc, g, p0, nb);
  else
    interp_kernel<T, kMaxKernelWidth><<<nbk, 256, 0, st>>>(recs, off, grid, vacc, g, p0, nb);
}

template <typename T, int KWM>
void launch_spread_row(const Rec<T> *recs, const int *off, cpx<T> *H, const cpx<T> *tw1, const Geom &g, int p0,
                       int nb, hipStream_t st) {
  spread_row_fft_kernel<T, 16, KWM><<<dim3(g.nf2 / g.rpw, nb), g.rpw * g.nf1 / 16, g.lds_row_bytes, st>>>(
      recs, off, H, tw1, g, p0);
}

template <typename T, int S>
void launch_spread_parts(const Rec<T> *recs, const int *off, cpx<T> *H, const cpx<T> *tw1, const Geom &g, int p0,
                         int nb, hipStream_t st) {
  const dim3 grd(S * g.nf2, nb);
  const int nt = g.nf1 / (16 * S), lb = g.lds_rowp_bytes;
  if (g.kw <= 8)
    spread_row_fft_part_kernel<T, 16, 8, S><<<grd, nt, lb, st>>>(recs, off, H, tw1, g, p0);
  else if (g.kw <= 16)
    spread_row_fft_part_kernel<T, 16, 16, S><<<grd, nt, lb, st>>>(recs, off, H, tw1, g, p0);
  else
    spread_row_fft_part_kernel<T, 16, kMaxKernelWidth, S><<<grd, nt, lb, st>>>(recs, off, H, tw1, g, p0);
}

template <typename T, int S>
void launch_interp_parts(const Rec<T> *recs, const int *off, const cpx<T> *H, double2 *vacc, const cpx<T> *tw1,
                         const Geom &g, int p0, int nb, hipStream_t st) {
  const dim3 grd(S * g.nf2, nb);
  const int nt = g.nf1 / (16 * S), lb = g.lds_rowp_bytes;
  if (g.kw <= 8)
    row_ifft_interp_part_kernel<T, 16, 8, S><<<grd, nt, lb, st>>>(H, recs, off, vacc, tw1, g, p0, -1);
  else if (g.kw <= 16)
    row_ifft_interp_part_kernel<T, 16, 16, S><<<grd, nt, lb, st>>>(H, recs, off, vacc, tw1, g, p0, -1);
  else
    row_ifft_interp_part_kernel<T, 16, kMaxKernelWidth, S><<<grd, nt, lb, st>>>(H, recs, off, vacc, tw1, g, p0, -1);
}

// ms2dirty, one batch on the LDS path: spread + row FFT -> H (column-tiled)
template <typename T>
void lds_spread_rows(const Rec<T> *recs, const int *off, const Geom &g, int p0, int nb, cpx<T> *H,
                     const cpx<T> *tw1, hipStream_t st) {
  if (g.rsplit == kRowSplit) {
    launch_spread_parts<T, kRowSplit>(recs, off, H, tw1, g, p0, nb, st);
    return;
  }
  if (g.kw <= 8)
    launch_spread_row<T, 8>(recs, off, H, tw1, g, p0, nb, st);
  else if (g.kw <= 16)
    launch_spread_row<T, 16>(recs, off, H, tw1, g, p0, nb, st);
  else
    launch_spread_row<T, kMaxKernelWidth>(recs, off, H, tw1, g, p0, nb, st);
}

inline int col_groups(const Geom &g) { return (g.nslot + g.cpw - 1) / g.cpw; }

// ms2dirty, one batch on the LDS path: column FFT + w-accumulation into accT
template <typename T>
void lds_cols_forward(const Geom &g, int p0, int nb, const cpx<T> *H, double2 *accT, const cpx<T> *tw2, int first,
                      hipStream_t st) {
  col_fwd_kernel<T, 16><<<col_groups(g), g.cpw * g.nf2 / 16, g.lds_col_bytes, st>>>(H, accT, tw2, g, p0, nb, first);
}

// dirty2ms, one batch on the LDS path: w-fill + inverse column FFT -> H (column-tiled)
template <typename T>
void lds_cols_backward(const Geom &g, int p0, int nb, const cpx<T> *gT, cpx<T> *H, const cpx<T> *tw2,
                       hipStream_t st) {
  col_inv_kernel<T, 16><<<col_groups(g), g.cpw * g.nf2 / 16, g.lds_col_bytes, st>>>(gT, H, tw2, g, p0, nb);
}

// dirty2ms, one batch on the LDS path: inverse row FFT fused with the interpolation
template <typename T>
void lds_row_interp(const Rec<T> *recs, const int *off, const cpx<T> *H, double2 *vacc, const cpx<T> *tw1,
                    const Geom &g, int p0, int nb, hipStream_t st) {
  if (g.rsplit == kRowSplit) {
    launch_interp_parts<T, kRowSplit>(recs, off, H, vacc, tw1, g, p0, nb, st);
    return;
  }
  const dim3 grd(g.nf2 / g.rpw, nb);
  const int nt = g.rpw * g.nf1 / 16;
  if (g.kw <= 8)
    row_ifft_interp_kernel<T, 16, 8><<<grd, nt, g.lds_row_bytes, st>>>(H, recs, off, vacc, tw1, g, p0, -1);
  else if (g.kw <= 16)
    row_ifft_interp_kernel<T, 16, 16><<<grd, nt, g.lds_row_bytes, st>>>(H, recs, off, vacc, tw1, g, p0, -1);
  else
    row_ifft_interp_kernel<T, 16, kMaxKernelWidth><<<grd, nt, g.lds_row_bytes, st>>>(H, recs, off, vacc, tw1, g,
                                                                                      p0, -1);
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

// Phase 3: queue one ms2dirty problem (no waits and no allocation unless timing).
template <typename T>
int enqueue_ms2dirty(const Problem &pb, const ChanPlan &cp, const Bufs &b, hipStream_t st, Stats *sts) {
  const Geom &g = cp.g;
  Timer tm;
  tm.on = sts != nullptr;
  tm.s = st;
  int e0 = tm.mark();
  enqueue_binning<T>(pb, cp, b, st, true);
  CURIG_HIP(hipGetLastError());
  int e1 = tm.mark();
  const long npix = (long)g.N1 * g.N2;
  const Rec<T> *recs = (const Rec<T> *)b.recs;
  v2_t<T> *grid = (v2_t<T> *)b.grid;
  const double *fk1 = b.tabs, *fk2 = b.tabs + (g.nf1 / 2 + 1), *q = fk2 + (g.nf2 / 2 + 1);
  double t_spread = 0, t_fft = 0, t_wacc = 0;
  for (int bi = 0; bi < g.nbatch; ++bi) {
    const int p0 = bi * g.Pb, nb = std::min(g.Pb, g.P - p0);
    const int a = tm.mark();
    if (g.fft_lds) {
      // spread + row FFT counts as spread (t_grid); the column FFT fused with
      // the w-accumulation as FFT (t_fft)
      lds_spread_rows<T>(recs, b.off, g, p0, nb, (cpx<T> *)b.H, (const cpx<T> *)b.tw1, st);
      CURIG_HIP(hipGetLastError());
      const int bm = tm.mark();
      lds_cols_forward<T>(g, p0, nb, (const cpx<T> *)b.H, b.acc, (const cpx<T> *)b.tw2, bi == 0 ? 1 : 0, st);
      CURIG_HIP(hipGetLastError());
      const int c = tm.mark();
      if (tm.on) {
        CURIG_HIP(hipEventSynchronize(tm.ev[c]));
        t_spread += tm.ms(a, bm);
        t_fft += tm.ms(bm, c);
      }
      continue;
    }
    spread_dispatch<T>(recs, b.off, grid, g, p0, nb, st);
    const int bm = tm.mark();
    CURIG_TRY(batch_fft<T>(b.dev, g, nb, +1, grid, st));
    CURIG_HIP(hipGetLastError());
    const int c = tm.mark();
    wacc_kernel<T><<<nblocks(npix), 256, 0, st>>>(grid, b.acc, g, p0, nb, bi == 0 ? 1 : 0);
    CURIG_HIP(hipGetLastError());
    const int d = tm.mark();
    if (tm.on) {
      CURIG_HIP(hipEventSynchronize(tm.ev[d]));
      t_spread += tm.ms(a, bm);
      t_fft += tm.ms(bm, c);
      t_wacc += tm.ms(c, d);
    }
  }
  int e2 = tm.mark();
  if (pb.dbl)
    epilogue_kernel<double><<<nblocks(npix), 256, 0, st>>>(b.acc, pb.dirty, g, fk1, fk2, q, pb.accumulate ? 1 : 0,
                                                           g.fft_lds);
  else
    epilogue_kernel<float><<<nblocks(npix), 256, 0, st>>>(b.acc, pb.dirty, g, fk1, fk2, q, pb.accumulate ? 1 : 0,
                                                          g.fft_lds);
  CURIG_HIP(hipGetLastError());
  int e3 = tm.mark();
  if (sts) {
    CURIG_HIP(hipStreamSynchronize(st));
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
  enqueue_binning<T>(pb, cp, b, st, false);
  // every output row is written: flagged rows get 0
  if (pb.nrow > 0) {
    if (pb.dbl)
      zero_rows_kernel<double><<<nblocks(pb.nrow), 256, 0, st>>>(pb.vis, pb.vis_stride, pb.nrow);
    else
      zero_rows_kernel<float><<<nblocks(pb.nrow), 256, 0, st>>>(pb.vis, pb.vis_stride, pb.nrow);
  }
  const long npix = (long)g.N1 * g.N2;
  const size_t plane = (size_t)g.nf1 * g.nf2;
  const Rec<T> *recs = (const Rec<T> *)b.recs;
  v2_t<T> *grid = (v2_t<T> *)b.grid;
  double2 *gimg = b.acc;
  const double *fk1 = b.tabs, *fk2 = b.tabs + (g.nf1 / 2 + 1), *q = fk2 + (g.nf2 / 2 + 1);
  if (g.fft_lds) {
    if (pb.dbl)
      prologue_cols_kernel<double, T><<<nblocks(npix), 256, 0, st>>>(pb.dirty, (cpx<T> *)gimg, g, fk1, fk2, q);
    else
      prologue_cols_kernel<float, T><<<nblocks(npix), 256, 0, st>>>(pb.dirty, (cpx<T> *)gimg, g, fk1, fk2, q);
  } else if (pb.dbl) {
    prologue_kernel<double><<<nblocks(npix), 256, 0, st>>>(pb.dirty, gimg, g, fk1, fk2, q);
  } else {
    prologue_kernel<float><<<nblocks(npix), 256, 0, st>>>(pb.dirty, gimg, g, fk1, fk2, q);
  }
  fill_zero_c128_kernel<<<nblocks(cp.nvalid), 256, 0, st>>>(b.vacc, cp.nvalid);
  CURIG_HIP(hipGetLastError());
  int e1 = tm.mark();
  double t_interp = 0, t_fft = 0, t_fill = 0;
  for (int bi = 0; bi < g.nbatch; ++bi) {
    const int p0 = bi * g.Pb, nb = std::min(g.Pb, g.P - p0);
    const int a = tm.mark();
    if (g.fft_lds) {
      // w-fill + column iFFT counts as FFT (t_fft); the row iFFT fused with
      // the interpolation as interpolation (t_grid)
      lds_cols_backward<T>(g, p0, nb, (const cpx<T> *)gimg, (cpx<T> *)b.H, (const cpx<T> *)b.tw2, st);
      CURIG_HIP(hipGetLastError());
      const int c = tm.mark();
      if (cp.nvalid > 0)
        lds_row_interp<T>(recs, b.off, (const cpx<T> *)b.H, b.vacc, (const cpx<T> *)b.tw1, g, p0, nb, st);
      CURIG_HIP(hipGetLastError());
      const int d = tm.mark();
      if (tm.on) {
        CURIG_HIP(hipEventSynchronize(tm.ev[d]));
        t_fft += tm.ms(a, c);
        t_interp += tm.ms(c, d);
      }
      continue;
    }
    wfill_kernel<T><<<nblocks((long)plane), 256, 0, st>>>(gimg, grid, g, p0, nb);
    CURIG_HIP(hipGetLastError());
    const int bm = tm.mark();
    CURIG_TRY(batch_fft<T>(b.dev, g, nb, -1, grid, st));
    CURIG_HIP(hipGetLastError());
    const int c = tm.mark();
    interp_dispatch<T>(recs, b.off, grid, b.vacc, g, cp.nvalid, p0, nb, st);
    CURIG_HIP(hipGetLastError());
    const int d = tm.mark();
    if (tm.on) {
      CURIG_HIP(hipEventSynchronize(tm.ev[d]));
      t_fill += tm.ms(a, bm);
      t_fft += tm.ms(bm, c);
      t_interp += tm.ms(c, d);
    }
  }
  int e2 = tm.mark();
  if (cp.nvalid > 0) {
    if (pb.dbl)
      finalize_kernel<double><<<nblocks(cp.nvalid), 256, 0, st>>>(b.perm, b.vacc, (int)cp.nvalid, pb.uvw, pb.wgt,
                                                                   pb.wgt_stride, pb.vis, pb.vis_stride, g);
    else
      finalize_kernel<float><<<nblocks(cp.nvalid), 256, 0, st>>>(b.perm, b.vacc, (int)cp.nvalid, pb.uvw, pb.wgt,
                                                                  pb.wgt_stride, pb.vis, pb.vis_stride, g);
    CURIG_HIP(hipGetLastError());
  }
  int e3 = tm.mark();
  if (sts) {
    CURIG_HIP(hipStreamSynchronize(st));
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

void drop_graph(GraphCache &gc) {
  if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
  if (gc.graph) (void)hipGraphDestroy(gc.graph);
  gc.exec = nullptr;
  gc.graph = nullptr;
  gc.key = 0;
}

template <typename T>
int run_problems(bool to_image, const Problem *pbs, int n, hipStream_t st, const Tuning *tune, int graph_mode,
                 Stats *sts) {
  for (int i = 0; i < n; ++i) CURIG_TRY(check_args(pbs[i]));
  int dev = 0;
  CURIG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  std::vector<ChanPlan> cps;
  CURIG_TRY(plan_all(pbs, n, st, tune, dev, cps));
  Bufs b;
  CURIG_TRY((reserve<T>(cps, pbs, to_image, dev, st, b)));
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
  uint64_t key = 1469598103934665603ull;
  key = fnv(key, &to_image, sizeof(to_image));
  key = fnv(key, pbs, sizeof(Problem) * n);
  for (const ChanPlan &cp : cps) key = fnv(key, &cp.g, sizeof(Geom));
  key = fnv(key, &b, sizeof(Bufs));
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

}  // namespace

int ms2dirty_device(const Problem &pb, hipStream_t st, const Tuning *tune, Stats *sts) {
  return pb.dbl ? run_problems<double>(true, &pb, 1, st, tune, 0, sts)
                : run_problems<float>(true, &pb, 1, st, tune, 0, sts);
}

int dirty2ms_device(const Problem &pb, hipStream_t st, const Tuning *tune, Stats *sts) {
  return pb.dbl ? run_problems<double>(false, &pb, 1, st, tune, 0, sts)
                : run_problems<float>(false, &pb, 1, st, tune, 0, sts);
}

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
  double fc2pi = 2 * M_PI * pb.freq / kSpeedOfLight;
  int nbk = (int)std::min<long>(npix, 1 << 16);
  if (pb.dbl)
    explicit_kernel<double><<<nbk, 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt, pb.wgt_stride,
                                                 pb.flag, pb.flag_stride, g, fc2pi, pb.sign, pix, npix,
                                                 (double2 *)out);
  else
    explicit_kernel<float><<<nbk, 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt, pb.wgt_stride,
                                                pb.flag, pb.flag_stride, g, fc2pi, pb.sign, pix, npix,
                                                (double2 *)out);
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
    for (int s = 0; s < S_NSLOT; ++s)
      if (g_ws[d].p[s]) (void)hipFree(g_ws[d].p[s]);
  }
  g_ws.clear();
  (void)hipSetDevice(cur);
  fft_release_all();
}

}  // namespace curig
