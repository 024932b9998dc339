// Preparation kernels: w range of the rows, binning of rows into sorted
// records (count, scan, scatter), fills.
#include "dev_common.h"

namespace curig {

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
// taps (LDS path, else null): each record's kernel values at its KP cells per
// axis, [nrec][3][KP] (x, y, w; zero beyond the kernel width; the w taps are
// 1, 0, ... in the 2-D form), evaluated from the rounded offsets stored in the
// record, so the row kernels read what they would compute.
template <typename T, int KP>
__global__ void __launch_bounds__(256) bin_scatter_kernel(const double *__restrict__ uvw, int nrow,
                                                           const void *__restrict__ vis, long vs,
                                                           const void *__restrict__ wgt, long ws, Geom g,
                                                           const int2 *__restrict__ binrank,
                                                           const int *__restrict__ off, Rec<T> *__restrict__ recs,
                                                           int *__restrict__ perm, int with_value,
                                                           T *__restrict__ taps) {
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
    if (taps) {
      const Ker<T> K = make_ker<T>(g);
      T kx[KP], ky[KP], kz[KP];
      K.template taps<KP>(r.dx, kx);
      K.template taps<KP>(r.dy, ky);
      if (g.kww == 1) {
#pragma unroll
        for (int i = 0; i < KP; ++i) kz[i] = i == 0 ? (T)1 : (T)0;
      } else {
        K.template taps<KP>(r.dw, kz);
      }
      T *o = taps + (size_t)dst * (3 * KP);
#pragma unroll
      for (int i = 0; i < KP; ++i) {
        o[i] = kx[i];
        o[KP + i] = ky[i];
        o[2 * KP + i] = kz[i];
      }
    }
  }
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
void launch_w_minmax(int nblk, const double *uvw, int nrow, const uint8_t *flag, long fs, double *part, int *cnt,
                     hipStream_t st) {
  w_minmax_kernel<<<nblk, 256, 0, st>>>(uvw, nrow, flag, fs, part, cnt);
}

void launch_fill_zero_c128(double2 *p, long n, hipStream_t st) {
  fill_zero_c128_kernel<<<nblocks(n), 256, 0, st>>>(p, n);
}

void launch_zero_rows(bool dbl, void *vis, long stride, int nrow, hipStream_t st) {
  if (nrow <= 0) return;
  if (dbl)
    zero_rows_kernel<double><<<nblocks(nrow), 256, 0, st>>>(vis, stride, nrow);
  else
    zero_rows_kernel<float><<<nblocks(nrow), 256, 0, st>>>(vis, stride, nrow);
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
  if (pb.nrow > 0) {
    T *taps = g.fft_lds ? (T *)b.taps : nullptr;
    if (tap_width(g.kw) == 8)
      bin_scatter_kernel<T, 8><<<nblocks(pb.nrow), 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt,
                                                                  pb.wgt_stride, g, b.binrank, b.off,
                                                                  (Rec<T> *)b.recs, b.perm, with_value ? 1 : 0, taps);
    else
      bin_scatter_kernel<T, 16><<<nblocks(pb.nrow), 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt,
                                                                   pb.wgt_stride, g, b.binrank, b.off,
                                                                   (Rec<T> *)b.recs, b.perm, with_value ? 1 : 0,
                                                                   taps);
  }
}

template void enqueue_binning<float>(const Problem &, const ChanPlan &, const Bufs &, hipStream_t, bool);
template void enqueue_binning<double>(const Problem &, const ChanPlan &, const Bufs &, hipStream_t, bool);

}  // namespace curig
