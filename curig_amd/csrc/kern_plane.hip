// Full-plane path (grids that are not powers of two: tile-owner spread into
// whole planes, rocFFT, w-accumulation / w-fill, interpolation), the
// correction epilogue / prologues shared with the LDS path, the finalize
// scatter and the FP64 direct sums.
#include "dev_common.h"

namespace curig {

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

// Unroll factor of the kernel-rows loop: full for widths up to 16; wider
// kernels keep it rolled (fully unrolled, their kw x kw tap loops take minutes
// to compile and gain nothing measurable).
template <int KWM> struct RowUnroll { static constexpr int value = KWM <= 16 ? KWM : 1; };

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
#pragma unroll RowUnroll<KWM>::value  // wide kernels: the rows loop stays rolled
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
#pragma unroll RowUnroll<KWM>::value  // wide kernels: the rows loop stays rolled
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


// FP64 direct sum of the adjoint at selected rows (the transpose of
// explicit_kernel): out[q] = wgt_j sum_pix dirty[pix] / n e^{-i ph(j, pix)},
// j = rows[q].  Grid (rows, pixel chunks): each block adds its chunk's partial
// sum to out[q] (cleared by the launcher).
template <typename TI>
__global__ void __launch_bounds__(256) explicit_adj_kernel(const double *__restrict__ uvw,
                                                            const void *__restrict__ dirty,
                                                            const void *__restrict__ wgt, long ws, Geom g,
                                                            double fc2pi, int sign, const int64_t *__restrict__ rows,
                                                            long chunk, double2 *__restrict__ out) {
  __shared__ double sre[256], sim[256];
  const long q = blockIdx.x;
  const long j = rows[q];
  const double u = uvw[3 * j], v = uvw[3 * j + 1], w = uvw[3 * j + 2];
  const long npix = (long)g.N1 * g.N2;
  const long p0 = (long)blockIdx.y * chunk, p1 = p0 + chunk < npix ? p0 + chunk : npix;
  double re = 0.0, im = 0.0;
  for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int i = (int)(p / g.N1), k = (int)(p - (long)i * g.N1);
    double nm1, n;
    pixel_n(i, k, g, nm1, n);
    if (!g.divn) nm1 = 0.0, n = 1.0;
    const double l = (double)(i - g.N2 / 2) * g.dx, m = (double)(k - g.N1 / 2) * g.dy;
    double s, c;
    sincos(-fc2pi * (l * u + m * v + sign * w * nm1), &s, &c);
    const double2 d = load_c<TI>(dirty, p);
    re += (d.x * c - d.y * s) / n;
    im += (d.x * s + d.y * c) / n;
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
  if (threadIdx.x == 0) {
    const double wt = wgt ? load_r<TI>(wgt, j * ws) : 1.0;
    atomicAdd(&out[q].x, wt * sre[0]);
    atomicAdd(&out[q].y, wt * sim[0]);
  }
}

__global__ void zero_c128_kernel(double2 *__restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
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

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
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

template <typename T>
void launch_wacc(const v2_t<T> *grid, double2 *acc, const Geom &g, int p0, int nb, int first, hipStream_t st) {
  wacc_kernel<T><<<nblocks((long)g.N1 * g.N2), 256, 0, st>>>(grid, acc, g, p0, nb, first);
}

template <typename T>
void launch_wfill(const double2 *gimg, v2_t<T> *grid, const Geom &g, int p0, int nb, hipStream_t st) {
  wfill_kernel<T><<<nblocks((long)g.nf1 * g.nf2), 256, 0, st>>>(gimg, grid, g, p0, nb);
}

void launch_epilogue(bool dbl_out, const double2 *acc, void *out, const Geom &g, const double *fk1, const double *fk2,
                     const double *q, int accumulate, int colmajor, hipStream_t st) {
  const long npix = (long)g.N1 * g.N2;
  if (dbl_out)
    epilogue_kernel<double><<<nblocks(npix), 256, 0, st>>>(acc, out, g, fk1, fk2, q, accumulate, colmajor);
  else
    epilogue_kernel<float><<<nblocks(npix), 256, 0, st>>>(acc, out, g, fk1, fk2, q, accumulate, colmajor);
}

void launch_prologue(bool dbl_in, const void *dirty, double2 *gimg, const Geom &g, const double *fk1,
                     const double *fk2, const double *q, hipStream_t st) {
  const long npix = (long)g.N1 * g.N2;
  if (dbl_in)
    prologue_kernel<double><<<nblocks(npix), 256, 0, st>>>(dirty, gimg, g, fk1, fk2, q);
  else
    prologue_kernel<float><<<nblocks(npix), 256, 0, st>>>(dirty, gimg, g, fk1, fk2, q);
}

template <typename T>
void launch_prologue_cols(bool dbl_in, const void *dirty, cpx<T> *gT, const Geom &g, const double *fk1,
                          const double *fk2, const double *q, hipStream_t st) {
  const long npix = (long)g.N1 * g.N2;
  if (dbl_in)
    prologue_cols_kernel<double, T><<<nblocks(npix), 256, 0, st>>>(dirty, gT, g, fk1, fk2, q);
  else
    prologue_cols_kernel<float, T><<<nblocks(npix), 256, 0, st>>>(dirty, gT, g, fk1, fk2, q);
}

void launch_finalize(bool dbl, const int *perm, const double2 *vacc, int nrec, const double *uvw, const void *wgt,
                     long ws, void *vis, long vs, const Geom &g, hipStream_t st) {
  if (nrec <= 0) return;
  if (dbl)
    finalize_kernel<double><<<nblocks(nrec), 256, 0, st>>>(perm, vacc, nrec, uvw, wgt, ws, vis, vs, g);
  else
    finalize_kernel<float><<<nblocks(nrec), 256, 0, st>>>(perm, vacc, nrec, uvw, wgt, ws, vis, vs, g);
}

void launch_explicit_ms2dirty(bool dbl, const Problem &pb, const Geom &g, double fc2pi, const int64_t *pix, long npix,
                              double2 *out, hipStream_t st) {
  const int nbk = (int)std::min<long>(npix, 1 << 16);
  if (dbl)
    explicit_kernel<double><<<nbk, 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt, pb.wgt_stride,
                                                 pb.flag, pb.flag_stride, g, fc2pi, pb.sign, pix, npix, out);
  else
    explicit_kernel<float><<<nbk, 256, 0, st>>>(pb.uvw, pb.nrow, pb.vis, pb.vis_stride, pb.wgt, pb.wgt_stride,
                                                pb.flag, pb.flag_stride, g, fc2pi, pb.sign, pix, npix, out);
}

void launch_explicit_dirty2ms(bool dbl, const Problem &pb, const Geom &g, double fc2pi, const int64_t *rows,
                              long nsel, double2 *out, hipStream_t st) {
  if (nsel <= 0) return;
  zero_c128_kernel<<<nblocks(nsel), 256, 0, st>>>(out, nsel);
  const long npix = (long)g.N1 * g.N2;
  const long chunk = 1 << 16;
  const dim3 grd((unsigned)nsel, (unsigned)((npix + chunk - 1) / chunk));
  if (dbl)
    explicit_adj_kernel<double><<<grd, 256, 0, st>>>(pb.uvw, pb.dirty, pb.wgt, pb.wgt_stride, g, fc2pi, pb.sign, rows,
                                                     chunk, out);
  else
    explicit_adj_kernel<float><<<grd, 256, 0, st>>>(pb.uvw, pb.dirty, pb.wgt, pb.wgt_stride, g, fc2pi, pb.sign, rows,
                                                    chunk, out);
}

#define CURIG_PLANE_INST(T)                                                                                        \
  template void spread_dispatch<T>(const Rec<T> *, const int *, v2_t<T> *, const Geom &, int, int, hipStream_t);  \
  template void interp_dispatch<T>(const Rec<T> *, const int *, const v2_t<T> *, double2 *, const Geom &, long,    \
                                   int, int, hipStream_t);                                                         \
  template void launch_wacc<T>(const v2_t<T> *, double2 *, const Geom &, int, int, int, hipStream_t);             \
  template void launch_wfill<T>(const double2 *, v2_t<T> *, const Geom &, int, int, hipStream_t);                 \
  template void launch_prologue_cols<T>(bool, const void *, cpx<T> *, const Geom &, const double *,               \
                                        const double *, const double *, hipStream_t);
CURIG_PLANE_INST(float)
CURIG_PLANE_INST(double)

}  // namespace curig
