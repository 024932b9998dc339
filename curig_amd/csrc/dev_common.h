// Device-side helpers shared by the gridder's kernel sources: the ES kernel,
// coordinate scaling, pixel geometry and typed loads / stores.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.h"

namespace curig {

#define M_1_2PI_D 0.159154943091895336  // include/utils.h:61
#define M_2PI_D 6.28318530717958648     // include/utils.h:62

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

}  // namespace curig
