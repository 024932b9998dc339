/* Additive exports of libcurafft.so (no counterpart among the reference's four
 * C entry points; they do not change those).  Conventions and return codes as
 * in cugridder.h.
 *
 *  - FP32-mixed form: complex64 visibilities / image and float32 weights; the ES
 *    kernel and the gridding / interpolation multiply-accumulate run in FP32,
 *    while uvw scaling, the w phase and the sum over w planes stay FP64.
 *  - Device-resident forms: all array pointers are device pointers on the current
 *    HIP device; work is queued on `stream` (a hipStream_t, NULL = default stream).
 *    The call returns after queuing, except for one internal wait for the w range.
 *  - Multi-channel forms (NIFTY layout): vis / wgt / flag are [nrow][nchan].
 *  - precision: 0 = FP64 (complex128 / float64), 1 = FP32-mixed (complex64 / float32).
 */
#ifndef CURIG_EXT_H
#define CURIG_EXT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FP32-mixed host entry points; same arguments as ms2dirty_2 / dirty2ms_2 with
 * vis, dirty complex64 and wgt float32 (wgt may be NULL). */
int ms2dirty_f32(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw, const void *vis,
                 const float *wgt, void *dirty, double epsilon, double sigma, int sign);
int dirty2ms_f32(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *uvw, void *vis,
                 const float *wgt, const void *dirty, double epsilon, double sigma, int sign);

/* Device-resident single channel. */
int curig_ms2dirty_dev(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *d_uvw,
                       const void *d_vis, const void *d_wgt, void *d_dirty, double epsilon, double sigma, int sign,
                       int precision, void *stream);
int curig_dirty2ms_dev(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *d_uvw,
                       void *d_vis, const void *d_wgt, const void *d_dirty, double epsilon, double sigma, int sign,
                       int precision, void *stream);

/* curig_ms2dirty_dev without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * working buffer comes from a fixed 128 MiB scratch array that each device
 * holds once the library has run there; the call takes as many w planes per
 * batch as fit.  Returns 10 (with curig_last_error) when the working set
 * exceeds the scratch or the plan is off the LDS FFT path (power-of-two grid of
 * at most 16384 per axis, image at most half the grid, kernel width <= 16).
 * Calls share the scratch: the library serialises them on the host and orders
 * calls on different streams after each other; the output is complete once
 * `stream` reaches the call's end. */
int curig_ms2dirty_dev_noalloc(int nrow, int nxdirty, int nydirty, double fov, double freq, const double *d_uvw,
                               const void *d_vis, const void *d_wgt, void *d_dirty, double epsilon, double sigma,
                               int sign, int precision, void *stream);

/* Device-resident multi-channel.  freq: host array [nchan].  d_flag: uint8
 * [nrow][nchan], nonzero excludes the sample (NULL: none).  ms2dirty writes
 * nchan images [nchan][nx][ny] when per_channel != 0, else one image summed over
 * channels.  dirty2ms reads the matching layout and writes every sample (flagged
 * samples get 0). */
int curig_ms2dirty_multi_dev(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                             const double *d_uvw, const void *d_vis, const void *d_wgt, const uint8_t *d_flag,
                             void *d_dirty, int per_channel, double epsilon, double sigma, int sign, int precision,
                             void *stream);
int curig_dirty2ms_multi_dev(int nrow, int nchan, int nxdirty, int nydirty, double fov, const double *freq,
                             const double *d_uvw, void *d_vis, const void *d_wgt, const uint8_t *d_flag,
                             const void *d_dirty, int per_channel, double epsilon, double sigma, int sign,
                             int precision, void *stream);

/* NIFTY-compatible host form (the caller interface of the reference's
 * python/curagridder/cursl.py:55-119, which passes only freq[0] and ignores
 * rad_pix_y, nx, ny, do_wstacking and flags; these entries honour all of them).
 * Host pointers; sign -1 and the NIFTY conventions:
 *   dirty[i][k] = sum_{row,ch} mask wgt ms e^{+i 2pi f_ch/c (u l_i + v m_k - w (n-1))} / n
 * with l_i = (i - nx/2) pixsize_x, m_k = (k - ny/2) pixsize_y.  dirty2ms is the
 * adjoint and writes ms [nrow][nchan] (masked samples get 0).
 *   ms, wgt, mask: [nrow][nchan] (wgt / mask may be NULL); mask != 0 keeps a sample.
 *   nu, nv: grid size along u and v (even, larger than nx, ny), fixing the
 *     upsampling factor to min(nu/nx, nv/ny); 0, 0: upsampling factor 2.
 *   do_wstacking == 0: the 2-D transform, w taken as 0 and no division by n.
 *   precision 0: ms / dirty complex128, wgt float64; 1: complex64 / float32.
 * ms2dirty returns the complex image; the real part is NIFTY's dirty image. */
int curig_nifty_ms2dirty(int nrow, int nchan, const double *uvw, const double *freq, const void *ms,
                         const void *wgt, const uint8_t *mask, int nxdirty, int nydirty, double pixsize_x,
                         double pixsize_y, int nu, int nv, double epsilon, int do_wstacking, void *dirty,
                         int precision);
int curig_nifty_dirty2ms(int nrow, int nchan, const double *uvw, const double *freq, void *ms, const void *wgt,
                         const uint8_t *mask, int nxdirty, int nydirty, double pixsize_x, double pixsize_y, int nu,
                         int nv, double epsilon, int do_wstacking, const void *dirty, int precision);

/* FP64 direct sum on the GPU (the reference's explicit_gridder,
 * src/astro/precomp.cu:242-315, with the weight applied): out[q] (complex128) =
 * dirty pixel d_pix[q] = i*ny + k; with d_pix == NULL, every pixel (npix ignored).
 * vis / wgt precision as `precision`. */
int curig_explicit_ms2dirty_dev(int nrow, int nxdirty, int nydirty, double fov, double freq,
                                const double *d_uvw, const void *d_vis, const void *d_wgt, int precision,
                                long npix, const int64_t *d_pix, void *d_out, int sign, void *stream);

/* FP64 direct sum of the adjoint (dirty2ms) on the GPU at nsel chosen rows:
 * out[q] (complex128) = wgt_j sum_{i,k} dirty[i][k] / n_ik
 * e^{-2 pi i f/c (u_j l_i + v_j m_k + sign w_j (n_ik - 1))}, j = d_rows[q].
 * dirty / wgt precision as `precision`. */
int curig_explicit_dirty2ms_dev(int nrow, int nxdirty, int nydirty, double fov, double freq,
                                const double *d_uvw, const void *d_dirty, const void *d_wgt, int precision,
                                long nsel, const int64_t *d_rows, void *d_out, int sign, void *stream);

/* Generic NUFFT on device pointers (the work behind cunufft_exec, declared in
 * cufft_nufft.h): type 1 (c -> fk) or 2 (fk -> c), dimension 1-3, modes
 * N1 (fast) x N2 x N3 (unused dimensions ignored), M points with coordinates
 * d_u, d_v, d_w (radians; d_v, d_w only as the dimension needs), iflag +-1,
 * fk [N3][N2][N1] and c [M] complex (precision 0: complex128, 1: complex64).
 * Synchronous on `stream`. */
int curig_nufft_dev(int type, int dim, int N1, int N2, int N3, int M, const double *d_u, const double *d_v,
                    const double *d_w, void *d_c, void *d_fk, int iflag, double epsilon, double sigma, int precision,
                    void *stream);

/* Type-3 NUFFT in one dimension on device pointers (the work behind
 * cunufft_exec for type 3): inputs x [M], outputs s [N] (radians, FP64),
 *   direction 1: f_i = sum_j c_j e^{i iflag s_i x_j}   (c -> f)
 *   direction 0: c_j = sum_i f_i e^{i iflag s_i x_j}   (f -> c)
 * c [M], f [N] complex (precision 0: complex128, 1: complex64).  Computed on
 * the w axis of the gridder (one wait for the range of s).  Synchronous. */
int curig_nufft3_dev(int M, int N, const double *d_x, const double *d_s, void *d_c, void *d_f, int iflag,
                     int direction, double epsilon, double sigma, int precision, void *stream);

/* Text of the last error on the calling thread. */
const char *curig_last_error(void);

/* Host-only plan computation (no device access): for nrow_valid unflagged rows
 * whose w (metres) spans [wmin, wmax].  ints[10] = kw, nf1, nf2, nf3 (w planes),
 * planes per batch, batches, tile width, tile height, FFT path (1: pruned LDS
 * FFT, 0: rocFFT), row split (parts per grid row, 0: whole rows); reals[8] = beta, c, gamma, h, i_center, o_center,
 * pixelsize_x, pixelsize_y. */
int curig_plan_info(int nrow_valid, int nxdirty, int nydirty, double fov, double freq, double epsilon,
                    double sigma, int sign, double wmin, double wmax, int precision, int *ints, double *reals);
/* Host-only: the same for the NIFTY form (sign -1; pixel sizes, grid size and
 * do_wstacking as in curig_nifty_ms2dirty). */
int curig_nifty_plan_info(int nrow_valid, int nxdirty, int nydirty, double pixsize_x, double pixsize_y,
                          double freq, double epsilon, int nu, int nv, int do_wstacking, double wmin, double wmax,
                          int precision, int *ints, double *reals);
/* Host-only: correction factors phihat(k) (-1)^k, k = 0..nf/2, of the kernel
 * normalised to peak 1 (multiply by e^beta for the reference's unnormalised kernel). */
int curig_kernel_fseries(double epsilon, double sigma, int nf, double *out);
/* Kernel evaluation of later calls: 0 = exp(beta (sqrt(1 - c x^2) - 1)) per tap
 * (default), 1 = Horner's rule on a piecewise polynomial fitted at plan time
 * (the reference's gpu_kerevalmeth = 1, include/curafft_opts.h:23, declared
 * there without an implementation).  The plan APIs take it from
 * opts.gpu_kerevalmeth instead. */
int curig_set_kernel_eval(int method);
/* Host-only: that polynomial for (epsilon, sigma) at a precision: *deg and
 * kw (deg + 1) coefficients, interval i = 0..kw-1 covering x in
 * [-kw/2 + i, -kw/2 + i + 1), powers of z = 2 (x + kw/2 - i) - 1 highest
 * first; max_coef is the room in coef. */
int curig_kernel_horner(double epsilon, double sigma, int precision, int *deg, double *coef, int max_coef);
/* Host-only: n-point Gauss-Legendre rule on [-1, 1]. */
int curig_gauss_legendre(int n, double *x, double *w);

/* Tuning of later calls (0 = default): tile width / height of the bins (for
 * power-of-two grids only the height is used: bins are full-width row strips),
 * w planes per batch (FFT batch) and w planes per bin layer.  The autotuner
 * (curig_amd.autotune) sets these. */
int curig_set_tuning(int tile_x, int tile_y, int planes_per_batch, int planes_per_bin);
/* FFT path of later calls: 0 = pruned LDS FFT where the grid is a power of two
 * and a row fits one workgroup's LDS (else rocFFT), 1 = always rocFFT. */
int curig_set_fft_path(int path);
/* Multi-channel calls: 1 (default) = capture the whole per-channel sequence into
 * one hipGraph and replay it while the same call repeats (power-of-two grids);
 * 0 = queue the per-channel kernels directly. */
int curig_set_graph_mode(int on);
/* 1: time the stages of later calls with hipEvents (adds synchronisation;
 * in graph mode only the total is timed). */
int curig_enable_stats(int on);
/* Stats of the last call on this process: ints[10] = kw, nf1, nf2, nf3, planes
 * per batch, batches, tile width, tile height, FFT path (1 = LDS), row split
 * (parts per grid row in the LDS row kernels, 0 = whole rows); reals[8] = rows gridded, t_prep,
 * t_grid (spread / interp), t_fft, t_plane (w accumulate / w fill), t_final,
 * t_total (ms), timed flag. */
int curig_get_stats(int *ints, double *reals);

/* Frees cached device buffers and rocFFT plans. */
void curig_release(void);

#ifdef __cplusplus
}
#endif

#endif
