/* Public C ABI of the gridder (libcurafft.so).
 *
 * Drop-in for include/cugridder.h:8-21 of the reference: the four extern "C"
 * entry points keep their names, argument order, meaning and return codes, so
 * the reference's ctypes binding (python/curagridder/curagridder.py:47-67,
 * cursl.py:29-49) loads this library unchanged.  Host pointers in, host
 * pointers out; each call is synchronous.
 *
 * Conventions (src/astro/cugridder.cu:330-613):
 *   uvw    [nrow][3] metres, C order
 *   vis    [nrow] complex128 (in for ms2dirty, out for dirty2ms)
 *   wgt    [nrow] float64 or absent (_1 variants); multiplies the visibilities in
 *          both directions
 *   dirty  [nxdirty][nydirty] complex128, x-major (out for ms2dirty, in for dirty2ms)
 *   fov    degrees (pixel size fov*pi/180/nxdirty along x, /nydirty along y)
 *   freq   Hz (one channel)
 *   sign   multiplies w: -1 NIFTY convention, +1 cuRIG convention
 *   ms2dirty: dirty[i][k] = sum_j wgt_j vis_j e^{+2 pi i f/c (l_i u_j + m_k v_j + sign w_j (n_ik - 1))} / n_ik
 *   dirty2ms: vis_j = wgt_j sum_ik dirty[i][k] / n_ik e^{-2 pi i f/c (...)}
 *   (both to the accuracy epsilon of the ES-kernel NUFFT)
 *
 * Return codes: 0 ok; 1 epsilon < 1.1e-16 (nothing computed, as the reference's
 * wrapper returns early, cugridder.cu:403-406); 2 sigma <= 1; 3 bad image size;
 * 4 nrow < 0; 5 bad argument (NULL pointer, sign); 6 HIP error; 7 device
 * allocation failed; 8 rocFFT error; 9 non-finite uvw/fov/freq.  The reference
 * exits the process on device errors (helper_cuda.h:583-595); this library
 * returns the code instead and curig_last_error() (curig_ext.h) gives the text.
 */
#ifndef CURIG_CUGRIDDER_H
#define CURIG_CUGRIDDER_H

#ifdef __cplusplus
#include <complex>
#include "curafft_plan.h"
#include "ragridder_plan.h"
typedef std::complex<double> curig_c128;

/* C++ plan API (src/astro/cugridder.cu:105-307).  Device buffers d_u, d_v, d_w
 * (metres) and d_c are owned by the plan after gridder_setting; set plan->fk to
 * the device image before gridder_execution.  gridder_destroy frees the plan's
 * device buffers, the host arrays in gridder_plan->kv and both plan structs
 * (allocate them with malloc / new of the POD struct), as the reference does.
 * kerevalmeth: 0 exp(sqrt()) kernel values, 1 Horner's rule on a piecewise
 * polynomial.  plan->opts.precision (set before gridder_setting; 0 after a
 * memset) selects the data type of d_c and plan->fk: 0 complex128 (the
 * reference's CUCPX), 1 complex64 (FP32-mixed); kv.weight stays float64 on the
 * host either way. */
int gridder_setting(int N1, int N2, int method, int kerevalmeth, int w_term_method, PCS tol, int direction,
                    double sigma, int iflag, int batchsize, int M, int channel, PCS fov, visibility *pointer_v,
                    PCS *d_u, PCS *d_v, PCS *d_w, CUCPX *d_c, curafft_plan *plan, ragridder_plan *gridder_plan);
int gridder_execution(curafft_plan *plan, ragridder_plan *gridder_plan);
int gridder_destroy(curafft_plan *plan, ragridder_plan *gridder_plan);

extern "C" {
#else
typedef double _Complex curig_c128;
#endif

/* cugridder.cu:594-599 */
int ms2dirty_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
               curig_c128 *dirty, double epsilon, double sigma, int sign);
/* cugridder.cu:587-592 */
int ms2dirty_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
               double *wgt, curig_c128 *dirty, double epsilon, double sigma, int sign);
/* cugridder.cu:601-606 */
int dirty2ms_1(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
               curig_c128 *dirty, double epsilon, double sigma, int sign);
/* cugridder.cu:608-613 */
int dirty2ms_2(int nrow, int nxdirty, int nydirty, double fov, double freq, double *uvw, curig_c128 *vis,
               double *wgt, curig_c128 *dirty, double epsilon, double sigma, int sign);

#ifdef __cplusplus
}
#endif

#endif
