// Test driver of the C++ plan APIs, written the way the reference's drivers
// use them (test/w_s_gridder_test.cu:166-271, test/w_s_degridder_test.cu:
// 170-276, test/nufft_2d_1_test.cu:180-327): host arrays copied to device
// buffers that the plan takes over, gridder_setting -> gridder_execution ->
// gridder_destroy, or cunufft_setting -> cunufft_exec -> curafft_free.  Built
// by the package build into a shared library the tests call through ctypes,
// so each run is checked against the oracle in Python.
#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cufft_nufft.h"
#include "cugridder.h"

namespace {

template <typename T> T *dev_copy(const T *h, size_t n) {
  T *d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
  if (n && hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return d;
}

template <typename T> T *host_dup(const T *h, size_t n) {
  if (!h) return nullptr;
  T *p = (T *)malloc(sizeof(T) * (n ? n : 1));
  memcpy(p, h, sizeof(T) * n);
  return p;
}

// complex128 host data as the plan's precision: complex128, or complex64 bytes
std::vector<char> as_precision(const std::complex<double> *h, size_t n, int precision) {
  std::vector<char> out(n * (precision ? 8 : 16));
  if (precision == 0) {
    memcpy(out.data(), h, n * 16);
  } else {
    std::complex<float> *f = (std::complex<float> *)out.data();
    for (size_t i = 0; i < n; ++i) f[i] = std::complex<float>((float)h[i].real(), (float)h[i].imag());
  }
  return out;
}

void from_precision(const std::vector<char> &in, std::complex<double> *h, size_t n, int precision) {
  if (precision == 0) {
    memcpy(h, in.data(), n * 16);
  } else {
    const std::complex<float> *f = (const std::complex<float> *)in.data();
    for (size_t i = 0; i < n; ++i) h[i] = std::complex<double>(f[i].real(), f[i].imag());
  }
}

}  // namespace

extern "C" {

// direction 1: vis -> dirty (ms2dirty); 0: dirty -> vis (dirty2ms).  u, v, w in
// metres (pirange 0) or times 2 pi (pirange 1); wgt may be NULL.  Returns the
// first nonzero code of the plan calls, or -1 for a driver failure.
int drv_gridder(int direction, int nrow, int nx, int ny, double fov, double freq, const double *u, const double *v,
                const double *w, int pirange, int sign, const double *wgt, std::complex<double> *vis,
                std::complex<double> *dirty, double eps, double sigma, int kerevalmeth, int precision) {
  const size_t cs = precision ? 8 : 16, npix = (size_t)nx * ny;
  double *d_u = dev_copy(u, nrow), *d_v = dev_copy(v, nrow), *d_w = dev_copy(w, nrow);
  std::vector<char> hv = as_precision(vis, direction == 1 ? nrow : 0, precision);
  hv.resize((size_t)nrow * cs);
  CUCPX *d_vis = nullptr;
  if (!d_u || !d_v || !d_w || hipMalloc(&d_vis, cs * (nrow ? nrow : 1)) != hipSuccess) return -1;
  if (direction == 1 && nrow && hipMemcpy(d_vis, hv.data(), cs * nrow, hipMemcpyHostToDevice) != hipSuccess) return -1;
  curafft_plan *plan = (curafft_plan *)calloc(1, sizeof(curafft_plan));
  ragridder_plan *gplan = (ragridder_plan *)calloc(1, sizeof(ragridder_plan));
  plan->opts.precision = precision;
  visibility *pv = (visibility *)malloc(sizeof(visibility));
  pv->u = host_dup(u, nrow);
  pv->v = host_dup(v, nrow);
  pv->w = host_dup(w, nrow);
  pv->vis = host_dup(vis, nrow);
  pv->weight = host_dup(wgt, nrow);
  pv->flag = nullptr;
  pv->frequency = host_dup(&freq, 1);
  pv->pirange = pirange;
  pv->sign = sign;
  int ier = gridder_setting(ny, nx, 0, kerevalmeth, 1, eps, direction, sigma, 0, 1, nrow, 1, fov, pv, d_u, d_v, d_w,
                            d_vis, plan, gplan);
  free(pv);
  if (ier == 0) {
    CUCPX *d_fk = nullptr;
    if (hipMalloc(&d_fk, cs * npix) != hipSuccess) return -1;
    plan->fk = d_fk;
    std::vector<char> hd = as_precision(dirty, direction == 0 ? npix : 0, precision);
    hd.resize(npix * cs);
    if (direction == 0 && hipMemcpy(d_fk, hd.data(), cs * npix, hipMemcpyHostToDevice) != hipSuccess) return -1;
    gplan->dirty_image = (std::complex<double> *)malloc(sizeof(std::complex<double>) * npix);
    ier = gridder_execution(plan, gplan);
    if (ier == 0) {
      if (direction == 1) {
        if (hipMemcpy(hd.data(), d_fk, cs * npix, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        from_precision(hd, dirty, npix, precision);
      } else if (nrow) {
        if (hipMemcpy(hv.data(), d_vis, cs * nrow, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        from_precision(hv, vis, nrow, precision);
      }
    }
  }
  const int ierd = gridder_destroy(plan, gplan);
  return ier ? ier : ierd;
}

// Generic NUFFT through cunufft_setting / cunufft_exec / curafft_free.  Types
// 1 / 2: coordinates u, v, w (radians, as many as dim needs), c [M], fk
// [N3][N2][N1]; type 1 runs c -> fk, type 2 fk -> c, with iflag +1 for
// direction 1 and -1 for direction 0 (cuft.cu:284-285).  Type 3 (dim 1):
// inputs u [M], outputs s [N1]; direction 1: fk = sum c e^{+i s u}, 0: c =
// sum fk e^{-i s u}.
int drv_nufft(int type, int dim, int N1, int N2, int N3, int M, const double *u, const double *v, const double *w,
              const double *s, int direction, double eps, double sigma, int kerevalmeth, int precision,
              std::complex<double> *c, std::complex<double> *fk) {
  const size_t cs = precision ? 8 : 16;
  const size_t nmodes = (size_t)N1 * (dim >= 2 ? N2 : 1) * (dim == 3 ? N3 : 1);
  double *d_u = dev_copy(u, M);
  double *d_v = dim >= 2 ? dev_copy(v, M) : nullptr;
  double *d_w = dim == 3 ? dev_copy(w, M) : nullptr;
  const bool c_in = type == 1 || (type == 3 && direction == 1);  // c is the input
  std::vector<char> hc = as_precision(c, c_in ? M : 0, precision);
  hc.resize((size_t)M * cs);
  CUCPX *d_c = nullptr;
  if (!d_u || hipMalloc(&d_c, cs * (M ? M : 1)) != hipSuccess) return -1;
  if (c_in && M && hipMemcpy(d_c, hc.data(), cs * M, hipMemcpyHostToDevice) != hipSuccess) return -1;
  curafft_plan *plan = (curafft_plan *)calloc(1, sizeof(curafft_plan));
  plan->opts.precision = precision;
  if (type == 3) plan->d_x = dev_copy(s, N1);
  int ier = cunufft_setting(N1, N2, N3, M, kerevalmeth, 0, direction, eps, sigma, type, dim, d_u, d_v, d_w, d_c,
                            plan);
  if (ier == 0) {
    CUCPX *d_fk = nullptr;
    if (hipMalloc(&d_fk, cs * nmodes) != hipSuccess) return -1;
    plan->fk = d_fk;
    std::vector<char> hf = as_precision(fk, c_in ? 0 : nmodes, precision);
    hf.resize(nmodes * cs);
    if (!c_in && hipMemcpy(d_fk, hf.data(), cs * nmodes, hipMemcpyHostToDevice) != hipSuccess) return -1;
    ier = cunufft_exec(plan);
    if (ier == 0) {
      if (c_in) {
        if (hipMemcpy(hf.data(), d_fk, cs * nmodes, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        from_precision(hf, fk, nmodes, precision);
      } else if (M) {
        if (hipMemcpy(hc.data(), d_c, cs * M, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        from_precision(hc, c, M, precision);
      }
    }
  }
  const int ierf = curafft_free(plan);
  if (plan->d_x) (void)hipFree(plan->d_x);
  free(plan);
  return ier ? ier : ierf;
}

}  // extern "C"
