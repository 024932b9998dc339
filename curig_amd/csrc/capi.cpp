// C ABI of libcurafft.so: the reference's four entry points over host buffers,
// its C++ plan API, and the additive device-resident / FP32 / multi-channel exports.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cufft_nufft.h"
#include "../../include/cugridder.h"
#include "../../include/curig_ext.h"
#include "curig_internal.h"
#include "gridder.h"

using namespace curig;

namespace {

std::mutex g_host_mu;
Tuning g_tuning;
bool g_stats_on = false;
Stats g_last_stats;

struct HostPath {
  hipStream_t stream = nullptr;
  void *buf[5] = {};
  size_t cap[5] = {};
};
std::vector<HostPath> g_host;

int host_buf(HostPath &h, int i, size_t bytes, void **out) {
  if (bytes == 0) bytes = 16;
  if (h.cap[i] < bytes) {
    if (h.buf[i]) CURIG_HIP(hipFree(h.buf[i]));
    h.buf[i] = nullptr;
    h.cap[i] = 0;
    hipError_t e = hipMalloc(&h.buf[i], bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed", bytes);
      return kErrAlloc;
    }
    h.cap[i] = bytes;
  }
  *out = h.buf[i];
  return kOk;
}

int host_path(HostPath **out) {
  int dev = 0;
  CURIG_HIP(hipGetDevice(&dev));
  if ((int)g_host.size() <= dev) g_host.resize(dev + 1);
  HostPath &h = g_host[dev];
  if (!h.stream) CURIG_HIP(hipStreamCreate(&h.stream));
  *out = &h;
  return kOk;
}

// Argument checks that need no device (same order of precedence as the reference:
// an epsilon below 1.1e-16 returns 1 before anything is computed).
int precheck(int nrow, int nx, int ny, double eps, double sigma) {
  KernelOpts ko;
  int ier = setup_conv_opts(ko, eps, sigma);
  if (ier != kOk) return ier;
  if (nrow < 0) {
    set_error("nrow %d < 0", nrow);
    return kErrNrow;
  }
  if (nx <= 0 || ny <= 0) {
    set_error("image size %d x %d invalid", nx, ny);
    return kErrSize;
  }
  return kOk;
}

Stats *stats_ptr() { return g_stats_on ? &g_last_stats : nullptr; }

// Host buffers in -> device -> gridder -> host buffers out.
int host_call(bool to_image, int nrow, int nx, int ny, double fov, double freq, const double *uvw, void *vis,
              const void *wgt, void *dirty, double eps, double sigma, int sign, bool dbl) {
  CURIG_TRY(precheck(nrow, nx, ny, eps, sigma));
  if (!dirty || (nrow > 0 && (!uvw || !vis))) {
    set_error("NULL data pointer");
    return kErrArg;
  }
  std::lock_guard<std::mutex> lk(g_host_mu);
  HostPath *h;
  CURIG_TRY(host_path(&h));
  const size_t cs = dbl ? 16 : 8, rs = dbl ? 8 : 4;
  const size_t nr = (size_t)nrow, npix = (size_t)nx * ny;
  void *d_uvw, *d_vis, *d_wgt = nullptr, *d_dirty;
  CURIG_TRY(host_buf(*h, 0, 24 * nr, &d_uvw));
  CURIG_TRY(host_buf(*h, 1, cs * nr, &d_vis));
  CURIG_TRY(host_buf(*h, 3, cs * npix, &d_dirty));
  hipStream_t st = h->stream;
  if (nrow > 0) CURIG_HIP(hipMemcpyAsync(d_uvw, uvw, 24 * nr, hipMemcpyHostToDevice, st));
  if (wgt && nrow > 0) {
    CURIG_TRY(host_buf(*h, 2, rs * nr, &d_wgt));
    CURIG_HIP(hipMemcpyAsync(d_wgt, wgt, rs * nr, hipMemcpyHostToDevice, st));
  }
  Problem pb;
  pb.nrow = nrow;
  pb.nx = nx;
  pb.ny = ny;
  pb.fov = fov;
  pb.freq = freq;
  pb.eps = eps;
  pb.sigma = sigma;
  pb.sign = sign;
  pb.dbl = dbl;
  pb.uvw = (const double *)d_uvw;
  pb.vis = d_vis;
  pb.wgt = d_wgt;
  pb.dirty = d_dirty;
  int ier;
  if (to_image) {
    if (nrow > 0) CURIG_HIP(hipMemcpyAsync(d_vis, vis, cs * nr, hipMemcpyHostToDevice, st));
    ier = ms2dirty_device(pb, st, &g_tuning, stats_ptr());
    if (ier != kOk) return ier;
    CURIG_HIP(hipMemcpyAsync(dirty, d_dirty, cs * npix, hipMemcpyDeviceToHost, st));
  } else {
    CURIG_HIP(hipMemcpyAsync(d_dirty, dirty, cs * npix, hipMemcpyHostToDevice, st));
    ier = dirty2ms_device(pb, st, &g_tuning, stats_ptr());
    if (ier != kOk) return ier;
    if (nrow > 0) CURIG_HIP(hipMemcpyAsync(vis, d_vis, cs * nr, hipMemcpyDeviceToHost, st));
  }
  CURIG_HIP(hipStreamSynchronize(st));
  return kOk;
}

Problem dev_problem(int nrow, int nx, int ny, double fov, double freq, const double *uvw, const void *vis,
                    const void *wgt, const void *dirty, double eps, double sigma, int sign, int precision) {
  Problem pb;
  pb.nrow = nrow;
  pb.nx = nx;
  pb.ny = ny;
  pb.fov = fov;
  pb.freq = freq;
  pb.eps = eps;
  pb.sigma = sigma;
  pb.sign = sign;
  pb.dbl = precision == 0;
  pb.uvw = uvw;
  pb.vis = const_cast<void *>(vis);
  pb.wgt = wgt;
  pb.dirty = const_cast<void *>(dirty);
  return pb;
}

int check_precision(int precision) {
  if (precision != 0 && precision != 1) {
    set_error("precision must be 0 (FP64) or 1 (FP32-mixed), got %d", precision);
    return kErrArg;
  }
  return kOk;
}

}  // namespace

// ---------------------------------------------------------------------------
// reference entry points (src/astro/cugridder.cu:587-613)
// ---------------------------------------------------------------------------
extern "C" int ms2dirty_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
                          curig_c128 *dirty, double epsilon, double sigma, int sign) {
  return host_call(true, nrow, nxdirty, nydirty, fov, freq, uvw, vis, nullptr, dirty, epsilon, sigma, sign, true);
}

extern "C" int ms2dirty_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
                          double *wgt, curig_c128 *dirty, double epsilon, double sigma, int sign) {
  return host_call(true, nrow, nxdirty, nydirty, fov, freq, uvw, vis, wgt, dirty, epsilon, sigma, sign, true);
}

extern "C" int dirty2ms_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
                          curig_c128 *dirty, double epsilon, double sigma, int sign) {
  return host_call(false, nrow, nxdirty, nydirty, fov, freq, uvw, vis, nullptr, dirty, epsilon, sigma, sign, true);
}

extern "C" int dirty2ms_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
                          double *wgt, curig_c128 *dirty, double epsilon, double sigma, int sign) {
  return host_call(false, nrow, nxdirty, nydirty, fov, freq, uvw, vis, wgt, dirty, epsilon, sigma, sign, true);
}

// ---------------------------------------------------------------------------
// additive exports
// ---------------------------------------------------------------------------
extern "C" int ms2dirty_f32(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw,
                            const void *vis, const float *wgt, void *dirty, double epsilon, double sigma, int sign) {
  return host_call(true, nrow, nxdirty, nydirty, fov, freq, uvw, const_cast<void *>(vis), wgt, dirty, epsilon, sigma,
                   sign, false);
}

extern "C" int dirty2ms_f32(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw, void *vis,
                            const float *wgt, const void *dirty, double epsilon, double sigma, int sign) {
  return host_call(false, nrow, nxdirty, nydirty, fov, freq, uvw, vis, wgt, const_cast<void *>(dirty), epsilon, sigma,
                   sign, false);
}

extern "C" int curig_ms2dirty_dev(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *d_uvw,
                                  const void *d_vis, const void *d_wgt, void *d_dirty, double epsilon, double sigma,
                                  int sign, int precision, void *stream) {
  CURIG_TRY(check_precision(precision));
  CURIG_TRY(precheck(nrow, nxdirty, nydirty, epsilon, sigma));
  Problem pb = dev_problem(nrow, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_dirty, epsilon, sigma, sign,
                           precision);
  return ms2dirty_device(pb, (hipStream_t)stream, &g_tuning, stats_ptr());
}

extern "C" int curig_ms2dirty_dev_noalloc(int nrow, int nxdirty, int nydirty, double fov, double freq,
                                          const double *d_uvw, const void *d_vis, const void *d_wgt, void *d_dirty,
                                          double epsilon, double sigma, int sign, int precision, void *stream) {
  CURIG_TRY(check_precision(precision));
  CURIG_TRY(precheck(nrow, nxdirty, nydirty, epsilon, sigma));
  Problem pb = dev_problem(nrow, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_dirty, epsilon, sigma, sign,
                           precision);
  return ms2dirty_device_noalloc(pb, (hipStream_t)stream, &g_tuning);
}

extern "C" int curig_dirty2ms_dev(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *d_uvw,
                                  void *d_vis, const void *d_wgt, const void *d_dirty, double epsilon, double sigma,
                                  int sign, int precision, void *stream) {
  CURIG_TRY(check_precision(precision));
  CURIG_TRY(precheck(nrow, nxdirty, nydirty, epsilon, sigma));
  Problem pb = dev_problem(nrow, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_dirty, epsilon, sigma, sign,
                           precision);
  return dirty2ms_device(pb, (hipStream_t)stream, &g_tuning, stats_ptr());
}

static int g_graph_mode = 1;

static int multi(bool to_image, int nrow, int nchan, int nx, int ny, double fov, const double *freq,
                 const double *d_uvw, const void *d_vis, const void *d_wgt, const uint8_t *d_flag, const void *d_dirty,
                 int per_channel, double eps, double sigma, int sign, int precision, void *stream,
                 const GridSpec *spec = nullptr) {
  CURIG_TRY(check_precision(precision));
  CURIG_TRY(precheck(nrow, nx, ny, eps, sigma));
  if (nchan <= 0 || !freq) {
    set_error("nchan must be > 0 with a frequency array");
    return kErrArg;
  }
  const size_t cs = precision == 0 ? 16 : 8, rs = precision == 0 ? 8 : 4;
  std::vector<Problem> pbs(nchan);
  for (int c = 0; c < nchan; ++c) {
    Problem pb = dev_problem(nrow, nx, ny, fov, freq[c], d_uvw, (const char *)d_vis + cs * c,
                             d_wgt ? (const char *)d_wgt + rs * c : nullptr, d_dirty, eps, sigma, sign, precision);
    pb.vis_stride = nchan;
    pb.wgt_stride = nchan;
    pb.flag = d_flag ? d_flag + c : nullptr;
    pb.flag_stride = nchan;
    if (per_channel) pb.dirty = (char *)const_cast<void *>(d_dirty) + cs * (size_t)nx * ny * c;
    pb.accumulate = to_image && !per_channel && c > 0;
    if (spec) pb.spec = *spec;
    pbs[c] = pb;
  }
  return multi_device(to_image, pbs.data(), nchan, (hipStream_t)stream, &g_tuning, g_graph_mode, stats_ptr());
}

extern "C" int curig_ms2dirty_multi_dev(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                                        const double *d_uvw, const void *d_vis, const void *d_wgt,
                                        const uint8_t *d_flag, void *d_dirty, int per_channel, double epsilon,
                                        double sigma, int sign, int precision, void *stream) {
  return multi(true, nrow, nchan, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_flag, d_dirty, per_channel,
               epsilon, sigma, sign, precision, stream);
}

extern "C" int curig_dirty2ms_multi_dev(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                                        const double *d_uvw, void *d_vis, const void *d_wgt, const uint8_t *d_flag,
                                        const void *d_dirty, int per_channel, double epsilon, double sigma, int sign,
                                        int precision, void *stream) {
  return multi(false, nrow, nchan, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_flag, d_dirty, per_channel,
               epsilon, sigma, sign, precision, stream);
}

// NIFTY-compatible host form: every channel of ms [nrow][nchan] gridded into
// one image (or the image degridded into every channel), pixel sizes per axis,
// optional grid size and mask, and the 2-D form without w-stacking.
static int nifty_call(bool to_image, int nrow, int nchan, const double *uvw, const double *freq, void *ms,
                      const void *wgt, const uint8_t *mask, int nx, int ny, double pixsize_x, double pixsize_y,
                      int nu, int nv, double eps, int do_wstacking, void *dirty, int precision) {
  CURIG_TRY(check_precision(precision));
  CURIG_TRY(precheck(nrow, nx, ny, eps, 2.0));
  if (nchan <= 0 || !freq || !dirty || (nrow > 0 && (!uvw || !ms))) {
    set_error("nchan must be > 0, with frequency, image and (for nrow > 0) uvw and ms arrays");
    return kErrArg;
  }
  if (!(pixsize_x > 0 && pixsize_y > 0 && std::isfinite(pixsize_x) && std::isfinite(pixsize_y))) {
    set_error("pixel sizes (%g, %g) must be positive and finite", pixsize_x, pixsize_y);
    return kErrNonFinite;
  }
  if ((nu == 0) != (nv == 0) || (nu != 0 && (nu <= nx || nv <= ny || (nu & 1) || (nv & 1)))) {
    set_error("grid %d x %d must be even and larger than the image %d x %d (or 0 x 0)", nu, nv, nx, ny);
    return kErrSize;
  }
  const double lmax = 0.5 * pixsize_x * nx, mmax = 0.5 * pixsize_y * ny;
  if (!(lmax * lmax + mmax * mmax < 1.0)) {
    set_error("field of view %g x %g rad reaches beyond the horizon (l^2 + m^2 >= 1)", 2 * lmax, 2 * mmax);
    return kErrArg;
  }
  GridSpec spec;
  spec.dx = pixsize_x;
  spec.dy = pixsize_y;
  spec.nu = nu;
  spec.nv = nv;
  spec.wstack = do_wstacking != 0;
  std::lock_guard<std::mutex> lk(g_host_mu);
  HostPath *h;
  CURIG_TRY(host_path(&h));
  const size_t cs = precision == 0 ? 16 : 8, rs = precision == 0 ? 8 : 4;
  const size_t nr = (size_t)nrow, ns = nr * nchan, npix = (size_t)nx * ny;
  void *d_uvw, *d_vis, *d_wgt = nullptr, *d_dirty, *d_flag = nullptr;
  CURIG_TRY(host_buf(*h, 0, 24 * nr, &d_uvw));
  CURIG_TRY(host_buf(*h, 1, cs * ns, &d_vis));
  CURIG_TRY(host_buf(*h, 3, cs * npix, &d_dirty));
  hipStream_t st = h->stream;
  if (nrow > 0) CURIG_HIP(hipMemcpyAsync(d_uvw, uvw, 24 * nr, hipMemcpyHostToDevice, st));
  if (wgt && nrow > 0) {
    CURIG_TRY(host_buf(*h, 2, rs * ns, &d_wgt));
    CURIG_HIP(hipMemcpyAsync(d_wgt, wgt, rs * ns, hipMemcpyHostToDevice, st));
  }
  std::vector<uint8_t> flag;
  if (mask && nrow > 0) {
    // mask != 0 keeps a sample; the gridder's flag != 0 drops it
    flag.resize(ns);
    for (size_t i = 0; i < ns; ++i) flag[i] = mask[i] == 0;
    CURIG_TRY(host_buf(*h, 4, ns, &d_flag));
    CURIG_HIP(hipMemcpyAsync(d_flag, flag.data(), ns, hipMemcpyHostToDevice, st));
  }
  if (to_image) {
    if (nrow > 0) CURIG_HIP(hipMemcpyAsync(d_vis, ms, cs * ns, hipMemcpyHostToDevice, st));
  } else {
    CURIG_HIP(hipMemcpyAsync(d_dirty, dirty, cs * npix, hipMemcpyHostToDevice, st));
  }
  CURIG_TRY(multi(to_image, nrow, nchan, nx, ny, 0.0, freq, (const double *)d_uvw, d_vis, d_wgt,
                  (const uint8_t *)d_flag, d_dirty, 0, eps, 2.0, -1, precision, st, &spec));
  if (to_image)
    CURIG_HIP(hipMemcpyAsync(dirty, d_dirty, cs * npix, hipMemcpyDeviceToHost, st));
  else if (nrow > 0)
    CURIG_HIP(hipMemcpyAsync(ms, d_vis, cs * ns, hipMemcpyDeviceToHost, st));
  CURIG_HIP(hipStreamSynchronize(st));
  return kOk;
}

extern "C" int curig_nifty_ms2dirty(int nrow, int nchan, const double *uvw, const double *freq, const void *ms,
                                    const void *wgt, const uint8_t *mask, int nxdirty, int nydirty,
                                    double pixsize_x, double pixsize_y, int nu, int nv, double epsilon,
                                    int do_wstacking, void *dirty, int precision) {
  return nifty_call(true, nrow, nchan, uvw, freq, const_cast<void *>(ms), wgt, mask, nxdirty, nydirty, pixsize_x,
                    pixsize_y, nu, nv, epsilon, do_wstacking, dirty, precision);
}

extern "C" int curig_nifty_dirty2ms(int nrow, int nchan, const double *uvw, const double *freq, void *ms,
                                    const void *wgt, const uint8_t *mask, int nxdirty, int nydirty,
                                    double pixsize_x, double pixsize_y, int nu, int nv, double epsilon,
                                    int do_wstacking, const void *dirty, int precision) {
  return nifty_call(false, nrow, nchan, uvw, freq, ms, wgt, mask, nxdirty, nydirty, pixsize_x, pixsize_y, nu, nv,
                    epsilon, do_wstacking, const_cast<void *>(dirty), precision);
}

// ---------------------------------------------------------------------------
// Generic NUFFT, types 1 and 2 in 1-3 dimensions, over the same kernels: the
// one-plane 2-D form with coordinates in radians (the grid phases themselves).
//   type 1: fk[k3][k2][k1] = sum_j c_j e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}
//   type 2: c_j = sum_k fk[k3][k2][k1] e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}
// with k' = k - N/2 (centred modes, N1 the fast index), as the reference's
// unit drivers define them (test/nufft_2d_1_test.cu, nufft_2d_2_test.cu,
// nufft_1d_test.cu ground truths).  Dimension 1 runs as N2 = 1; dimension 3
// as N3 passes of the 2-D transform over c_j e^{i iflag k3' w_j}.
// ---------------------------------------------------------------------------
__global__ void pack_nufft_kernel(const double *u, const double *v, double *uvw, int n, double s) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    uvw[3 * j] = v ? s * v[j] : 0.0;  // slow image axis (N2) pairs with v
    uvw[3 * j + 1] = s * u[j];       // fast image axis (N1) pairs with u
    uvw[3 * j + 2] = 0.0;
  }
}

template <typename T> struct Cx;
template <> struct Cx<double> { using type = double2; };
template <> struct Cx<float> { using type = float2; };

// out_j = c_j e^{i t w_j}
template <typename T>
__global__ void phase_mul_kernel(const typename Cx<T>::type *c, const double *w, typename Cx<T>::type *out, int n,
                                 double t) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    double sn, cs;
    sincos(t * w[j], &sn, &cs);
    const double a = c[j].x, b = c[j].y;
    out[j].x = (T)(a * cs - b * sn);
    out[j].y = (T)(a * sn + b * cs);
  }
}

// acc_j (+)= x_j e^{i t w_j}
template <typename T>
__global__ void phase_acc_kernel(typename Cx<T>::type *acc, const typename Cx<T>::type *x, const double *w, int n,
                                 double t, int first) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    double sn, cs;
    sincos(t * w[j], &sn, &cs);
    const double a = x[j].x, b = x[j].y;
    double re = a * cs - b * sn, im = a * sn + b * cs;
    if (!first) {
      re += acc[j].x;
      im += acc[j].y;
    }
    acc[j].x = (T)re;
    acc[j].y = (T)im;
  }
}

template <typename T>
static int nufft_run(int type, int dim, int N1, int N2, int N3, int M, const double *d_u, const double *d_v,
                     const double *d_w, void *d_c, void *d_fk, int iflag, double eps, double sigma, hipStream_t st,
                     const Tuning &tn) {
  using C2 = typename Cx<T>::type;
  const bool dbl = sizeof(T) == 8;
  const int n2 = dim >= 2 ? N2 : 1, n3 = dim == 3 ? N3 : 1;
  double *d_uvw = nullptr;
  C2 *tmp = nullptr;
  int ier = kOk;
  const int nb = std::max(1, std::min((M + 255) / 256, 65536));
  if (M > 0) {
    CURIG_HIP(hipMalloc(&d_uvw, sizeof(double) * 3 * (size_t)M));
    // type 1 runs ms2dirty (e^{+i x.k}); type 2 runs dirty2ms (e^{-i x.k})
    pack_nufft_kernel<<<nb, 256, 0, st>>>(d_u, dim >= 2 ? d_v : nullptr, d_uvw, M, type == 1 ? iflag : -iflag);
    ier = hipGetLastError() == hipSuccess ? kOk : kErrHip;
    if (ier == kOk && dim == 3 && hipMalloc(&tmp, sizeof(C2) * (size_t)M) != hipSuccess) ier = kErrAlloc;
  }
  Problem pb;
  pb.nrow = M;
  pb.nx = n2;
  pb.ny = N1;
  pb.freq = 1.0;
  pb.eps = eps;
  pb.sigma = sigma;
  pb.sign = -1;
  pb.dbl = dbl;
  pb.uvw = d_uvw;
  pb.spec.radians = true;
  pb.spec.wstack = false;
  const size_t slab = (size_t)N1 * n2;
  for (int k3 = 0; k3 < n3 && ier == kOk; ++k3) {
    const double t = (double)iflag * (double)(k3 - n3 / 2);
    pb.dirty = (C2 *)d_fk + slab * k3;
    if (type == 1) {
      pb.vis = d_c;
      if (dim == 3 && M > 0) {
        phase_mul_kernel<T><<<nb, 256, 0, st>>>((const C2 *)d_c, d_w, tmp, M, t);
        pb.vis = tmp;
      }
      ier = ms2dirty_device(pb, st, &tn, stats_ptr());
    } else {
      pb.vis = dim == 3 ? (void *)tmp : d_c;
      ier = dirty2ms_device(pb, st, &tn, stats_ptr());
      if (ier == kOk && dim == 3 && M > 0) phase_acc_kernel<T><<<nb, 256, 0, st>>>((C2 *)d_c, tmp, d_w, M, t, k3 == 0);
    }
  }
  if (ier == kOk && hipStreamSynchronize(st) != hipSuccess) ier = kErrHip;
  if (d_uvw) (void)hipFree(d_uvw);
  if (tmp) (void)hipFree(tmp);
  return ier;
}

static int nufft_checked(int type, int dim, int N1, int N2, int N3, int M, const double *d_u, const double *d_v,
                         const double *d_w, void *d_c, void *d_fk, int iflag, double epsilon, double sigma,
                         int precision, void *stream, const Tuning &tn) {
  CURIG_TRY(check_precision(precision));
  KernelOpts ko;
  CURIG_TRY(setup_conv_opts(ko, epsilon, sigma));
  if ((type != 1 && type != 2) || dim < 1 || dim > 3 || (iflag != 1 && iflag != -1)) {
    set_error("NUFFT type %d, dimension %d, iflag %d: types 1 and 2, dimensions 1-3, iflag +-1", type, dim, iflag);
    return kErrArg;
  }
  if (M < 0) {
    set_error("M %d < 0", M);
    return kErrNrow;
  }
  if (N1 <= 0 || (dim >= 2 && N2 <= 0) || (dim == 3 && N3 <= 0)) {
    set_error("mode counts %d x %d x %d invalid", N1, N2, N3);
    return kErrSize;
  }
  if (!d_fk || (M > 0 && (!d_c || !d_u || (dim >= 2 && !d_v) || (dim == 3 && !d_w)))) {
    set_error("NULL data pointer");
    return kErrArg;
  }
  return precision == 0 ? nufft_run<double>(type, dim, N1, N2, N3, M, d_u, d_v, d_w, d_c, d_fk, iflag, epsilon,
                                            sigma, (hipStream_t)stream, tn)
                        : nufft_run<float>(type, dim, N1, N2, N3, M, d_u, d_v, d_w, d_c, d_fk, iflag, epsilon,
                                           sigma, (hipStream_t)stream, tn);
}

extern "C" int curig_nufft_dev(int type, int dim, int N1, int N2, int N3, int M, const double *d_u,
                               const double *d_v, const double *d_w, void *d_c, void *d_fk, int iflag,
                               double epsilon, double sigma, int precision, void *stream) {
  return nufft_checked(type, dim, N1, N2, N3, M, d_u, d_v, d_w, d_c, d_fk, iflag, epsilon, sigma, precision, stream,
                       g_tuning);
}

// Type 3 in one dimension, inputs x_j (M) and outputs s_i (N), through the w
// axis of the same kernels: rows (0, 0, +-x_j) and a 1 x N image whose pixels
// carry the outputs as their n - 1 (cuft.cu:313-350 type-3 setting; the
// reference's drivers test/nufft_1d_3_1_test.cu, nufft_1d_3_2_test.cu):
//   direction 1: f_i = sum_j c_j e^{i iflag s_i x_j}
//   direction 0: c_j = sum_i f_i e^{i iflag s_i x_j}
__global__ void pack_type3_kernel(const double *x, double *uvw, int n, double s) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    uvw[3 * j] = 0.0;
    uvw[3 * j + 1] = 0.0;
    uvw[3 * j + 2] = s * x[j];
  }
}

static int nufft3_checked(int M, int N, const double *d_x, const double *d_s, void *d_c, void *d_f, int iflag,
                          int direction, double epsilon, double sigma, int precision, void *stream, const Tuning &tn) {
  CURIG_TRY(check_precision(precision));
  KernelOpts ko;
  CURIG_TRY(setup_conv_opts(ko, epsilon, sigma));
  if (iflag != 1 && iflag != -1) {
    set_error("iflag %d: +-1", iflag);
    return kErrArg;
  }
  if (M < 0) {
    set_error("M %d < 0", M);
    return kErrNrow;
  }
  if (N <= 0) {
    set_error("N %d: at least one output", N);
    return kErrSize;
  }
  if (!d_s || !d_f || (M > 0 && (!d_x || !d_c))) {
    set_error("NULL data pointer");
    return kErrArg;
  }
  hipStream_t st = (hipStream_t)stream;
  // the outputs' range sets the plan (one wait)
  std::vector<double> hs(N);
  CURIG_HIP(hipMemcpyAsync(hs.data(), d_s, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  CURIG_HIP(hipStreamSynchronize(st));
  double smin = hs[0], smax = hs[0];
  for (double v : hs) {
    if (!std::isfinite(v)) {
      set_error("non-finite output coordinate");
      return kErrNonFinite;
    }
    smin = std::min(smin, v);
    smax = std::max(smax, v);
  }
  double *d_uvw = nullptr;
  if (M > 0) {
    CURIG_HIP(hipMalloc(&d_uvw, sizeof(double) * 3 * (size_t)M));
    // direction 1 runs ms2dirty (e^{+i w z}), direction 0 dirty2ms (e^{-i w z})
    const int nb = std::max(1, std::min((M + 255) / 256, 65536));
    pack_type3_kernel<<<nb, 256, 0, st>>>(d_x, d_uvw, M, direction == 1 ? iflag : -iflag);
  }
  Problem pb;
  pb.nrow = M;
  pb.nx = 1;
  pb.ny = N;
  pb.freq = 1.0;
  pb.eps = epsilon;
  pb.sigma = sigma;
  pb.sign = 1;
  pb.dbl = precision == 0;
  pb.uvw = d_uvw;
  pb.vis = d_c;
  pb.dirty = d_f;
  pb.spec.radians = true;
  pb.spec.zout = d_s;
  pb.spec.zmin = smin;
  pb.spec.zmax = smax;
  int ier = direction == 1 ? ms2dirty_device(pb, st, &tn, stats_ptr()) : dirty2ms_device(pb, st, &tn, stats_ptr());
  if (ier == kOk && hipStreamSynchronize(st) != hipSuccess) ier = kErrHip;
  if (d_uvw) (void)hipFree(d_uvw);
  return ier;
}

extern "C" int curig_nufft3_dev(int M, int N, const double *d_x, const double *d_s, void *d_c, void *d_f, int iflag,
                                int direction, double epsilon, double sigma, int precision, void *stream) {
  return nufft3_checked(M, N, d_x, d_s, d_c, d_f, iflag, direction, epsilon, sigma, precision, stream, g_tuning);
}

extern "C" int curig_explicit_ms2dirty_dev(int nrow, int nxdirty, int nydirty, double fov, double freq,
                                           const double *d_uvw, const void *d_vis, const void *d_wgt, int precision,
                                           long npix, const int64_t *d_pix, void *d_out, int sign, void *stream) {
  CURIG_TRY(check_precision(precision));
  if (!d_out) {
    set_error("NULL output");
    return kErrArg;
  }
  Problem pb = dev_problem(nrow, nxdirty, nydirty, fov, freq, d_uvw, d_vis, d_wgt, d_out, 1e-6, 2.0, sign, precision);
  return explicit_ms2dirty_device(pb, npix, d_pix, d_out, (hipStream_t)stream);
}

extern "C" int curig_explicit_dirty2ms_dev(int nrow, int nxdirty, int nydirty, double fov, double freq,
                                           const double *d_uvw, const void *d_dirty, const void *d_wgt, int precision,
                                           long nsel, const int64_t *d_rows, void *d_out, int sign, void *stream) {
  CURIG_TRY(check_precision(precision));
  if (!d_out || !d_dirty) {
    set_error("NULL image or output");
    return kErrArg;
  }
  Problem pb = dev_problem(nrow, nxdirty, nydirty, fov, freq, d_uvw, nullptr, d_wgt, d_dirty, 1e-6, 2.0, sign,
                           precision);
  pb.vis = d_out;  // unused by the direct sum; satisfies the argument check
  return explicit_dirty2ms_device(pb, nsel, d_rows, d_out, (hipStream_t)stream);
}

extern "C" const char *curig_last_error(void) { return last_error(); }

extern "C" int curig_plan_info(int nrow_valid, int nxdirty, int nydirty, double fov, double freq, double epsilon,
                               double sigma, int sign, double wmin, double wmax, int precision, int *ints,
                               double *reals) {
  CURIG_TRY(check_precision(precision));
  if (nxdirty <= 0 || nydirty <= 0) {
    set_error("image size %d x %d invalid", nxdirty, nydirty);
    return kErrSize;
  }
  Geom g;
  KernelOpts ko;
  CURIG_TRY(plan_geometry(nrow_valid, nxdirty, nydirty, fov, freq, epsilon, sigma, sign, wmin, wmax, precision == 0,
                          &g_tuning, &g, &ko));
  if (ints) {
    int v[10] = {g.kw, g.nf1, g.nf2, g.P, g.Pb, g.nbatch, g.T1, g.T2, g.fft_lds, g.rsplit};
    memcpy(ints, v, sizeof(v));
  }
  if (reals) {
    double r[8] = {ko.beta, ko.c, g.gamma, g.h, g.C, g.D, g.dx, g.dy};
    memcpy(reals, r, sizeof(r));
  }
  return kOk;
}

extern "C" int curig_nifty_plan_info(int nrow_valid, int nxdirty, int nydirty, double pixsize_x, double pixsize_y,
                                     double freq, double epsilon, int nu, int nv, int do_wstacking, double wmin,
                                     double wmax, int precision, int *ints, double *reals) {
  CURIG_TRY(check_precision(precision));
  if (nxdirty <= 0 || nydirty <= 0) {
    set_error("image size %d x %d invalid", nxdirty, nydirty);
    return kErrSize;
  }
  if (!(pixsize_x > 0 && pixsize_y > 0) || (nu == 0) != (nv == 0)) {
    set_error("pixel sizes must be positive and the grid size given for both axes or neither");
    return kErrArg;
  }
  GridSpec spec;
  spec.dx = pixsize_x;
  spec.dy = pixsize_y;
  spec.nu = nu;
  spec.nv = nv;
  spec.wstack = do_wstacking != 0;
  Geom g;
  KernelOpts ko;
  CURIG_TRY(plan_geometry(nrow_valid, nxdirty, nydirty, 0.0, freq, epsilon, 2.0, -1, wmin, wmax, precision == 0,
                          &g_tuning, &g, &ko, &spec));
  if (ints) {
    int v[10] = {g.kw, g.nf1, g.nf2, g.P, g.Pb, g.nbatch, g.T1, g.T2, g.fft_lds, g.rsplit};
    memcpy(ints, v, sizeof(v));
  }
  if (reals) {
    double r[8] = {ko.beta, ko.c, g.gamma, g.h, g.C, g.D, g.dx, g.dy};
    memcpy(reals, r, sizeof(r));
  }
  return kOk;
}

extern "C" int curig_set_kernel_eval(int method) {
  if (method != 0 && method != 1) {
    set_error("kernel evaluation method %d: 0 (exp of sqrt) or 1 (piecewise polynomial)", method);
    return kErrArg;
  }
  g_tuning.kereval = method;
  return kOk;
}

extern "C" int curig_kernel_horner(double epsilon, double sigma, int precision, int *deg, double *coef,
                                   int max_coef) {
  CURIG_TRY(check_precision(precision));
  KernelOpts ko;
  CURIG_TRY(setup_conv_opts(ko, epsilon, sigma));
  std::vector<double> c;
  const int d = horner_fit(ko, horner_tol(epsilon, precision == 0), c);
  if (d < 0) {
    set_error("no polynomial of degree <= %d fits the kernel", kMaxHornerDeg);
    return kErrArg;
  }
  if (deg) *deg = d;
  if (coef) {
    if ((int)c.size() > max_coef) {
      set_error("%zu coefficients, room for %d", c.size(), max_coef);
      return kErrArg;
    }
    memcpy(coef, c.data(), sizeof(double) * c.size());
  }
  return kOk;
}

extern "C" int curig_kernel_fseries(double epsilon, double sigma, int nf, double *out) {
  KernelOpts ko;
  int ier = setup_conv_opts(ko, epsilon, sigma);
  if (ier == kErrSigma) return ier;
  if (nf < 2 || !out) {
    set_error("nf must be >= 2");
    return kErrArg;
  }
  fseries_grid(ko, nf, out);
  return ier;
}

extern "C" int curig_gauss_legendre(int n, double *x, double *w) {
  if (n < 1 || !x || !w) {
    set_error("n must be >= 1");
    return kErrArg;
  }
  gauss_legendre(n, x, w);
  return kOk;
}

extern "C" int curig_set_tuning(int tile_x, int tile_y, int planes_per_batch, int planes_per_bin) {
  if (tile_x < 0 || tile_y < 0 || planes_per_batch < 0 || planes_per_bin < 0 || tile_x > 256 || tile_y > 256 ||
      (size_t)tile_x * tile_y > 4096 || planes_per_batch > 4096 || planes_per_bin > 4096) {
    set_error("invalid tuning %d x %d, %d planes per batch, %d per bin", tile_x, tile_y, planes_per_batch,
              planes_per_bin);
    return kErrArg;
  }
  g_tuning.T1 = tile_x;
  g_tuning.T2 = tile_y;
  g_tuning.Pb = planes_per_batch;
  g_tuning.Pbin = planes_per_bin;
  return kOk;
}

extern "C" int curig_set_fft_path(int path) {
  if (path < 0 || path > 1) {
    set_error("fft path must be 0 (auto) or 1 (rocFFT)");
    return kErrArg;
  }
  g_tuning.fft = path;
  return kOk;
}

extern "C" int curig_set_graph_mode(int on) {
  g_graph_mode = on != 0;
  return kOk;
}

extern "C" int curig_enable_stats(int on) {
  g_stats_on = on != 0;
  return kOk;
}

extern "C" int curig_get_stats(int *ints, double *reals) {
  const Stats &s = g_last_stats;
  if (ints) {
    int v[10] = {s.kw, s.nf1, s.nf2, s.P, s.Pb, s.nbatch, s.T1, s.T2, s.fft_lds, s.rsplit};
    memcpy(ints, v, sizeof(v));
  }
  if (reals) {
    double r[8] = {(double)s.nvis, s.t_prep, s.t_grid, s.t_fft, s.t_plane, s.t_final, s.t_total, (double)s.timed};
    memcpy(reals, r, sizeof(r));
  }
  return kOk;
}

extern "C" void curig_release(void) {
  std::lock_guard<std::mutex> lk(g_host_mu);
  for (auto &h : g_host) {
    for (void *b : h.buf)
      if (b) (void)hipFree(b);
    if (h.stream) (void)hipStreamDestroy(h.stream);
  }
  g_host.clear();
  release_workspace();
}

// ---------------------------------------------------------------------------
// C++ plan API (src/astro/cugridder.cu:105-307)
// ---------------------------------------------------------------------------
__global__ void pack_uvw_kernel(const double *u, const double *v, const double *w, double *uvw, int n, double s) {
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    uvw[3 * j] = u[j] * s;
    uvw[3 * j + 1] = v[j] * s;
    uvw[3 * j + 2] = w[j] * s;
  }
}

int gridder_setting(int N1, int N2, int method, int kerevalmeth, int w_term_method, PCS tol, int direction,
                    double sigma, int iflag, int batchsize, int M, int channel, PCS fov, visibility *pointer_v,
                    PCS *d_u, PCS *d_v, PCS *d_w, CUCPX *d_c, curafft_plan *plan, ragridder_plan *gridder_plan) {
  (void)iflag;
  (void)batchsize;
  if (!plan || !gridder_plan || !pointer_v || !pointer_v->frequency) {
    set_error("NULL plan or visibility struct");
    return kErrArg;
  }
  if (w_term_method != 1 || method != 0 || (kerevalmeth != 0 && kerevalmeth != 1)) {
    set_error("only improved w-stacking (w_term_method 1), method 0 and kerevalmeth 0 (exp) or 1 (Horner) "
              "are provided");
    return kErrArg;
  }
  if (plan->opts.precision != 0 && plan->opts.precision != 1) {
    set_error("opts.precision must be 0 (FP64) or 1 (FP32-mixed)");
    return kErrArg;
  }
  const bool dbl = plan->opts.precision == 0;
  KernelOpts ko;
  int ier = setup_conv_opts(ko, tol, sigma);
  if (ier != kOk) return ier;
  if (M < 0 || N1 <= 0 || N2 <= 0) {
    set_error("invalid sizes");
    return M < 0 ? kErrNrow : kErrSize;
  }
  plan->tol = tol;
  plan->opts.upsampfac = sigma;
  plan->opts.gpu_kerevalmeth = kerevalmeth;  // cugridder.cu:146
  plan->opts.gpu_gridder_method = method;
  plan->copts.kw = ko.kw;
  plan->copts.direction = direction;
  plan->copts.pirange = 1;
  plan->copts.upsampfac = sigma;
  plan->copts.ES_beta = ko.beta;
  plan->copts.ES_halfwidth = ko.halfwidth;
  plan->copts.ES_c = ko.c;
  plan->iflag = direction > 0 ? 1 : -1;
  plan->type = direction > 0 ? 1 : 2;
  plan->dim = 3;
  plan->mode_flag = 1;
  plan->M = M;
  plan->ms = N1;
  plan->mt = N2;
  plan->mu = 1;
  plan->execute_flow = 1;
  plan->d_u = d_u;
  plan->d_v = d_v;
  plan->d_w = d_w;
  plan->d_c = d_c;
  gridder_plan->kv = *pointer_v;
  gridder_plan->fov = fov;
  gridder_plan->width = N1;
  gridder_plan->height = N2;
  gridder_plan->nrow = M;
  gridder_plan->channel = channel;
  gridder_plan->w_term_method = w_term_method;
  gridder_plan->speedoflight = (int)kSpeedOfLight;
  gridder_plan->pixelsize_x = fov / 180.0 * M_PI / (double)N2;
  gridder_plan->pixelsize_y = fov / 180.0 * M_PI / (double)N1;
  // pack u, v, w (metres; pirange = 1 means already multiplied by 2 pi)
  double s = pointer_v->pirange ? 1.0 / (2 * M_PI) : 1.0;
  CURIG_HIP(hipMalloc(&plan->d_uvw, sizeof(double) * 3 * (size_t)std::max(M, 1)));
  if (M > 0) {
    pack_uvw_kernel<<<(M + 255) / 256, 256>>>(d_u, d_v, d_w, plan->d_uvw, M, s);
    CURIG_HIP(hipGetLastError());
  }
  if (pointer_v->weight && M > 0) {
    // the weights in the precision of the plan's data
    CURIG_HIP(hipMalloc(&plan->d_wgt, sizeof(double) * M));
    if (dbl) {
      CURIG_HIP(hipMemcpy(plan->d_wgt, pointer_v->weight, sizeof(double) * M, hipMemcpyHostToDevice));
    } else {
      std::vector<float> wf(pointer_v->weight, pointer_v->weight + M);
      CURIG_HIP(hipMemcpy(plan->d_wgt, wf.data(), sizeof(float) * M, hipMemcpyHostToDevice));
    }
  }
  // plan sizes for the caller (w range from the host copy when present)
  double wmin = 0, wmax = 0;
  if (M > 0) {
    std::vector<double> hw(M);
    CURIG_HIP(hipMemcpy(hw.data(), d_w, sizeof(double) * M, hipMemcpyDeviceToHost));
    wmin = wmax = hw[0] * s;
    for (double x : hw) {
      wmin = std::min(wmin, x * s);
      wmax = std::max(wmax, x * s);
    }
  }
  Geom g;
  Tuning tn = g_tuning;
  tn.kereval = kerevalmeth;
  ier = plan_geometry(M, N2, N1, fov, pointer_v->frequency[0], tol, sigma, pointer_v->sign, wmin, wmax, dbl, &tn, &g,
                      nullptr);
  if (ier != kOk) return ier;
  plan->nf1 = g.nf1;
  plan->nf2 = g.nf2;
  plan->nf3 = g.P;
  plan->batchsize = g.Pb;
  plan->ta.i_center[0] = g.C;
  plan->ta.o_center[0] = g.D;
  plan->ta.gamma[0] = g.gamma;
  plan->ta.h[0] = g.h;
  gridder_plan->num_w = g.P;
  return kOk;
}

int gridder_execution(curafft_plan *plan, ragridder_plan *gridder_plan) {
  if (!plan || !gridder_plan || !plan->fk || (plan->M > 0 && !plan->d_c)) {
    set_error("plan->fk (image) and plan->d_c (visibilities) must be set");
    return kErrArg;
  }
  Problem pb;
  pb.nrow = plan->M;
  pb.nx = plan->mt;
  pb.ny = plan->ms;
  pb.fov = gridder_plan->fov;
  pb.freq = gridder_plan->kv.frequency[0];
  pb.sigma = plan->copts.upsampfac;
  pb.sign = gridder_plan->kv.sign;
  pb.dbl = plan->opts.precision == 0;  // as gridder_setting planned
  pb.uvw = plan->d_uvw;
  pb.vis = plan->d_c;
  pb.wgt = plan->d_wgt;
  pb.dirty = plan->fk;
  pb.eps = plan->tol;
  hipStream_t st = plan->streams ? plan->streams[0] : nullptr;
  Tuning tn = g_tuning;
  tn.kereval = plan->opts.gpu_kerevalmeth == 1 ? 1 : 0;
  int ier = plan->copts.direction == 1 ? ms2dirty_device(pb, st, &tn, stats_ptr())
                                       : dirty2ms_device(pb, st, &tn, stats_ptr());
  if (ier != kOk) return ier;
  CURIG_HIP(hipStreamSynchronize(st));
  return kOk;
}

int gridder_destroy(curafft_plan *plan, ragridder_plan *gridder_plan) {
  if (plan) {
    if (plan->d_uvw) (void)hipFree(plan->d_uvw);
    if (plan->d_wgt) (void)hipFree(plan->d_wgt);
    if (plan->d_u) (void)hipFree(plan->d_u);
    if (plan->d_v) (void)hipFree(plan->d_v);
    if (plan->d_w) (void)hipFree(plan->d_w);
    if (plan->d_c) (void)hipFree(plan->d_c);
    if (plan->fk) (void)hipFree(plan->fk);
    free(plan);
  }
  if (gridder_plan) {
    free(gridder_plan->dirty_image);
    free(gridder_plan->kv.u);
    free(gridder_plan->kv.v);
    free(gridder_plan->kv.w);
    free(gridder_plan->kv.vis);
    free(gridder_plan->kv.frequency);
    free(gridder_plan->kv.weight);
    free(gridder_plan);
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// C++ generic-NUFFT plan API (src/fourier/cuft.cu:249-444)
// ---------------------------------------------------------------------------
int cunufft_setting(int N1, int N2, int N3, int M, int kerevalmeth, int method, int direction, PCS tol, PCS sigma,
                    int type, int dim, PCS *d_u, PCS *d_v, PCS *d_w, CUCPX *d_c, curafft_plan *plan) {
  if (!plan) {
    set_error("NULL plan");
    return kErrArg;
  }
  KernelOpts ko;
  CURIG_TRY(setup_conv_opts(ko, tol, sigma));
  if (!((type == 1 || type == 2) && dim >= 1 && dim <= 3) && !(type == 3 && dim == 1)) {
    set_error("NUFFT type %d, dimension %d: types 1 and 2 in dimensions 1-3, type 3 in dimension 1", type, dim);
    return kErrArg;
  }
  if (M < 0 || N1 <= 0 || (dim >= 2 && N2 <= 0) || (dim == 3 && N3 <= 0)) {
    set_error("M %d, modes %d x %d x %d invalid", M, N1, N2, N3);
    return M < 0 ? kErrNrow : kErrSize;
  }
  plan->opts.upsampfac = sigma;
  plan->opts.gpu_sort = 1;
  plan->opts.gpu_kerevalmeth = kerevalmeth;
  plan->opts.gpu_gridder_method = method;
  plan->copts.kw = ko.kw;
  plan->copts.direction = direction;
  plan->copts.pirange = 1;
  plan->copts.upsampfac = sigma;
  plan->copts.ES_beta = ko.beta;
  plan->copts.ES_halfwidth = ko.halfwidth;
  plan->copts.ES_c = ko.c;
  plan->type = type;
  plan->dim = dim;
  plan->mode_flag = 1;
  plan->iflag = direction > 0 ? 1 : -1;  // cuft.cu:284-285
  plan->ms = N1;
  plan->mt = dim >= 2 ? N2 : 1;
  plan->mu = dim == 3 ? N3 : 1;
  plan->M = M;
  plan->d_u = d_u;
  plan->d_v = d_v;
  plan->d_w = d_w;
  plan->d_c = d_c;
  plan->tol = tol;
  plan->batchsize = 1;
  plan->execute_flow = 1;
  if (type == 3) {
    // type 3: inputs d_u (M), outputs plan->d_x (N1), set by the caller before
    // this call (cuft.cu:313-330); the grid follows from both ranges at exec
    if (!plan->d_x || (M > 0 && !d_u)) {
      set_error("type 3 needs plan->d_x (the N1 output coordinates) and d_u");
      return kErrArg;
    }
    plan->nf1 = plan->nf2 = plan->nf3 = 0;
    return kOk;
  }
  // grid of the 2-D pass (the third dimension is a sum over its N3 modes)
  GridSpec sp;
  sp.radians = true;
  sp.wstack = false;
  Geom g;
  CURIG_TRY(plan_geometry(M, plan->mt, N1, 0.0, 1.0, tol, sigma, -1, 0.0, 0.0, plan->opts.precision == 0, &g_tuning,
                          &g, nullptr, &sp));
  plan->nf1 = g.nf1;
  plan->nf2 = g.nf2;
  plan->nf3 = 1;
  return kOk;
}

int cunufft_exec(curafft_plan *plan) {
  if (!plan || !plan->fk) {
    set_error("plan->fk (the mode array) must be set");
    return kErrArg;
  }
  Tuning tn = g_tuning;
  tn.kereval = plan->opts.gpu_kerevalmeth == 1 ? 1 : 0;
  if (plan->type == 3)
    return nufft3_checked(plan->M, plan->ms, plan->d_u, plan->d_x, plan->d_c, plan->fk, plan->iflag,
                          plan->copts.direction > 0 ? 1 : 0, plan->tol, plan->copts.upsampfac, plan->opts.precision,
                          plan->streams ? plan->streams[0] : nullptr, tn);
  return nufft_checked(plan->type, plan->dim, plan->ms, plan->mt, plan->mu, plan->M, plan->d_u, plan->d_v,
                       plan->d_w, plan->d_c, plan->fk, plan->iflag, plan->tol, plan->copts.upsampfac,
                       plan->opts.precision, plan->streams ? plan->streams[0] : nullptr, tn);
}

int curafft_free(curafft_plan *plan) {
  if (!plan) return kOk;
  // the plan owns the arrays given to cunufft_setting and plan->fk (cuft.cu:446-478)
  PCS **own[] = {&plan->d_u, &plan->d_v, &plan->d_w};
  for (PCS **p : own)
    if (*p) (void)hipFree(*p), *p = nullptr;
  if (plan->d_c) (void)hipFree(plan->d_c), plan->d_c = nullptr;
  if (plan->fw) (void)hipFree(plan->fw), plan->fw = nullptr;
  if (!plan->opts.gpu_conv_only && plan->fk) (void)hipFree(plan->fk), plan->fk = nullptr;
  return kOk;
}
