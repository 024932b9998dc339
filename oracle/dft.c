/*
 * oracle/dft.c -- FP64 explicit direct-sum DFT for ms2dirty / dirty2ms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP gridder:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it, and never as the thing measured or shipped.
 *
 * What it restates (reference citations are relative to the CuRIG tree):
 *   - src/astro/precomp.cu:242-275  explicit_gridder (GPU FP64 direct sum):
 *       dirty[pix] = sum_k vis_k exp(i 2pi f/c (l u + m v + (n-1) w)) / n
 *   - python/curagridder/test_wst.py:15-32 and explicit_test.py:13-54
 *       (explicit_gridder / explicit_degridder): pixel grid -ss/2 + arange(ss),
 *       nm1 = -(l^2+m^2)/(sqrt(1-l^2-m^2)+1), sign multiplies w.
 *   - src/astro/cugridder.cu:48-49 pixel sizes: dx = fov*pi/180/nx (pairs with u),
 *       dy = fov*pi/180/ny (pairs with v); image is [nx][ny], x-major.
 *   - src/astro/ra_exec.cu:98-105 and src/utils.cu:185-192: the weight multiplies
 *       the visibilities in BOTH directions ("divide" launches the multiply).
 *   Pixel centre uses integer halving (C's nx/2), as the reference's C code does.
 *
 * Conventions (sign = +1 cuRIG API, -1 NIFTY API):
 *   ms2dirty: dirty[i][k] = sum_j wgt_j vis_j e^{+i phi_jik} / n_ik
 *   dirty2ms: vis_j      = wgt_j sum_ik dirty[i][k] / n_ik e^{-i phi_jik}
 *   phi_jik = 2 pi f/c (l_i u_j + m_k v_j + sign w_j (n_ik - 1))
 *
 * Pinned by: tests/test_oracle.py (nrow=1 plane wave in closed form, uvw=0
 * constant image, exact adjointness, numpy brute force on tiny cases, the
 * committed golden fixtures in tests/golden/).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_C 299792458.0

static void pixel_lmn(int i, int k, int nx, int ny, double dx, double dy,
                      double *l, double *m, double *nm1, double *n) {
  double ll = (double)(i - nx / 2) * dx;
  double mm = (double)(k - ny / 2) * dy;
  double r2 = ll * ll + mm * mm;
  double s = sqrt(1.0 - r2);
  *l = ll;
  *m = mm;
  *nm1 = -r2 / (s + 1.0);
  *n = s;
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* Number of threads the oracle will use (after nthreads is applied). */
int oracle_threads(int nthreads) { return set_threads(nthreads); }

/* dirty: complex128 [nx][ny] interleaved (re, im).  vis: complex128 [nrow].
 * wgt may be NULL.  Overwrites dirty. */
void oracle_dft_ms2dirty(int nrow, int nx, int ny, double fov, double freq,
                         const double *uvw, const double *vis,
                         const double *wgt, double *dirty, int sign,
                         int nthreads) {
  set_threads(nthreads);
  const double dx = fov * M_PI / 180.0 / nx, dy = fov * M_PI / 180.0 / ny;
  const double s = 2.0 * M_PI * freq / ORACLE_C;
  const int64_t npix = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < npix; ++p) {
    int i = (int)(p / ny), k = (int)(p % ny);
    double l, m, nm1, n;
    pixel_lmn(i, k, nx, ny, dx, dy, &l, &m, &nm1, &n);
    double re = 0.0, im = 0.0;
    for (int j = 0; j < nrow; ++j) {
      double ph = s * (l * uvw[3 * j] + m * uvw[3 * j + 1] +
                       sign * uvw[3 * j + 2] * nm1);
      double sn, cs;
      sincos(ph, &sn, &cs);
      double wt = wgt ? wgt[j] : 1.0;
      double vr = vis[2 * j] * wt, vi = vis[2 * j + 1] * wt;
      re += vr * cs - vi * sn;
      im += vr * sn + vi * cs;
    }
    dirty[2 * p] = re / n;
    dirty[2 * p + 1] = im / n;
  }
}

/* Same sum at a subset of pixels: pix[q] = i*ny + k.  out: complex128 [npix]. */
void oracle_dft_ms2dirty_pixels(int nrow, int nx, int ny, double fov,
                                double freq, const double *uvw,
                                const double *vis, const double *wgt,
                                int npix, const int64_t *pix, double *out,
                                int sign, int nthreads) {
  set_threads(nthreads);
  const double dx = fov * M_PI / 180.0 / nx, dy = fov * M_PI / 180.0 / ny;
  const double s = 2.0 * M_PI * freq / ORACLE_C;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < npix; ++q) {
    int i = (int)(pix[q] / ny), k = (int)(pix[q] % ny);
    double l, m, nm1, n;
    pixel_lmn(i, k, nx, ny, dx, dy, &l, &m, &nm1, &n);
    double re = 0.0, im = 0.0;
    for (int j = 0; j < nrow; ++j) {
      double ph = s * (l * uvw[3 * j] + m * uvw[3 * j + 1] +
                       sign * uvw[3 * j + 2] * nm1);
      double sn, cs;
      sincos(ph, &sn, &cs);
      double wt = wgt ? wgt[j] : 1.0;
      double vr = vis[2 * j] * wt, vi = vis[2 * j + 1] * wt;
      re += vr * cs - vi * sn;
      im += vr * sn + vi * cs;
    }
    out[2 * q] = re / n;
    out[2 * q + 1] = im / n;
  }
}

/* vis: complex128 [nrow] output (overwritten).  dirty: complex128 [nx][ny]. */
void oracle_dft_dirty2ms(int nrow, int nx, int ny, double fov, double freq,
                         const double *uvw, double *vis, const double *wgt,
                         const double *dirty, int sign, int nthreads) {
  set_threads(nthreads);
  const double dx = fov * M_PI / 180.0 / nx, dy = fov * M_PI / 180.0 / ny;
  const double s = 2.0 * M_PI * freq / ORACLE_C;
  const int64_t npix = (int64_t)nx * ny;
  /* per-pixel l, m, n-1 and dirty/n, computed once */
  double *pl = (double *)malloc(sizeof(double) * 5 * (size_t)npix);
  for (int64_t p = 0; p < npix; ++p) {
    int i = (int)(p / ny), k = (int)(p % ny);
    double l, m, nm1, n;
    pixel_lmn(i, k, nx, ny, dx, dy, &l, &m, &nm1, &n);
    pl[5 * p] = l;
    pl[5 * p + 1] = m;
    pl[5 * p + 2] = nm1;
    pl[5 * p + 3] = dirty[2 * p] / n;
    pl[5 * p + 4] = dirty[2 * p + 1] / n;
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nrow; ++j) {
    double u = uvw[3 * j], v = uvw[3 * j + 1], w = uvw[3 * j + 2];
    double re = 0.0, im = 0.0;
    for (int64_t p = 0; p < npix; ++p) {
      const double *q = pl + 5 * p;
      double ph = -s * (q[0] * u + q[1] * v + sign * w * q[2]);
      double sn, cs;
      sincos(ph, &sn, &cs);
      re += q[3] * cs - q[4] * sn;
      im += q[3] * sn + q[4] * cs;
    }
    double wt = wgt ? wgt[j] : 1.0;
    vis[2 * j] = re * wt;
    vis[2 * j + 1] = im * wt;
  }
  free(pl);
}

/* NIFTY form (the caller interface the reference's python/curagridder/cursl.py
 * mirrors, with the arguments that file drops honoured): several channels,
 * separate pixel sizes, a mask, and the 2-D form.
 *   ms2dirty: dirty[i][k] = sum_{j,c} mask wgt ms e^{+i phi} / n
 *   dirty2ms: ms[j][c]    = mask wgt sum_{i,k} dirty[i][k] / n e^{-i phi}
 *   phi = 2 pi f_c/c (l_i u_j + m_k v_j - w_j (n_ik - 1))   (sign -1, cursl.py:75,108)
 * l_i = (i - nx/2) dx, m_k = (k - ny/2) dy.  apply_w == 0 is NIFTY's
 * do_wstacking=False: n - 1 taken as 0 and n as 1 (w ignored, no 1/n).
 * ms, wgt, mask: [nrow][nchan]; wgt and mask may be NULL (mask != 0 keeps a
 * sample).  to_image selects the direction; the output is overwritten. */
void oracle_dft_nifty(int to_image, int nrow, int nchan, int nx, int ny, double dx, double dy, const double *freq,
                      const double *uvw, double *ms, const double *wgt, const unsigned char *mask, int apply_w,
                      double *dirty, int nthreads) {
  set_threads(nthreads);
  const int64_t npix = (int64_t)nx * ny;
  double *pl = (double *)malloc(sizeof(double) * 4 * (size_t)npix);
  for (int64_t p = 0; p < npix; ++p) {
    int i = (int)(p / ny), k = (int)(p % ny);
    double l, m, nm1, n;
    pixel_lmn(i, k, nx, ny, dx, dy, &l, &m, &nm1, &n);
    if (!apply_w) nm1 = 0.0, n = 1.0;
    pl[4 * p] = l;
    pl[4 * p + 1] = m;
    pl[4 * p + 2] = nm1;
    pl[4 * p + 3] = n;
  }
  if (to_image) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < npix; ++p) {
      const double *q = pl + 4 * p;
      double re = 0.0, im = 0.0;
      for (int j = 0; j < nrow; ++j)
        for (int c = 0; c < nchan; ++c) {
          const int64_t s = (int64_t)j * nchan + c;
          if (mask && !mask[s]) continue;
          const double f = 2.0 * M_PI * freq[c] / ORACLE_C;
          double ph = f * (q[0] * uvw[3 * j] + q[1] * uvw[3 * j + 1] - uvw[3 * j + 2] * q[2]);
          double sn, cs;
          sincos(ph, &sn, &cs);
          double wt = wgt ? wgt[s] : 1.0;
          double vr = ms[2 * s] * wt, vi = ms[2 * s + 1] * wt;
          re += vr * cs - vi * sn;
          im += vr * sn + vi * cs;
        }
      dirty[2 * p] = re / q[3];
      dirty[2 * p + 1] = im / q[3];
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nrow; ++j)
      for (int c = 0; c < nchan; ++c) {
        const int64_t s = (int64_t)j * nchan + c;
        double re = 0.0, im = 0.0;
        if (!mask || mask[s]) {
          const double f = 2.0 * M_PI * freq[c] / ORACLE_C;
          for (int64_t p = 0; p < npix; ++p) {
            const double *q = pl + 4 * p;
            double ph = -f * (q[0] * uvw[3 * j] + q[1] * uvw[3 * j + 1] - uvw[3 * j + 2] * q[2]);
            double sn, cs;
            sincos(ph, &sn, &cs);
            double dr = dirty[2 * p] / q[3], di = dirty[2 * p + 1] / q[3];
            re += dr * cs - di * sn;
            im += dr * sn + di * cs;
          }
          double wt = wgt ? wgt[s] : 1.0;
          re *= wt;
          im *= wt;
        }
        ms[2 * s] = re;
        ms[2 * s + 1] = im;
      }
  }
  free(pl);
}
