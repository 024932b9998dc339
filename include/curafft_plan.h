/* Plan structures of the gridder.
 * Replaces include/curafft_plan.h:16-122 and include/curafft_opts.h:4-39 of the
 * reference: same struct and field names and meanings; the cuFFT handle and CUDA
 * stream become rocFFT / HIP types, and a few fields are added at the end of
 * curafft_opts / curafft_plan (marked "added"). */
#ifndef CURIG_CURAFFT_PLAN_H
#define CURIG_CURAFFT_PLAN_H

#include <hip/hip_runtime_api.h>
#include <rocfft/rocfft.h>

#include "datatype.h"

#define MAX_KERNEL_WIDTH 28

struct t3attrib {
  PCS i_half_width[3]; /* half-width of the effective w range (radians)       */
  PCS o_half_width[3]; /* half-width of n-1 over the image                    */
  PCS i_center[3];     /* centre of the effective w range                     */
  PCS o_center[3];     /* centre of n-1 over the image                        */
  PCS gamma[3];        /* w rescale factor, w' = (w - i_center) / gamma       */
  PCS h[3];            /* w-plane spacing 2 pi / nf3                          */
};

struct conv_opts {
  int kw;        /* kernel width (cells)                                     */
  int direction; /* 1: visibilities -> image, 0: image -> visibilities       */
  int pirange;   /* 1: coordinates in [-pi, pi)                              */
  PCS upsampfac; /* sigma                                                    */
  PCS ES_beta;
  PCS ES_halfwidth;
  PCS ES_c; /* 4 / kw^2 */
};

struct curafft_opts {
  double upsampfac;
  int gpu_device_id;
  int gpu_sort;        /* rows are always binned by (w-plane batch, tile)      */
  int gpu_binsizex;    /* tile width along the grid's fast axis (0: default)   */
  int gpu_binsizey;    /* tile height (0: default)                             */
  int gpu_binsizez;    /* w planes per batch (0: from the memory budget)       */
  int gpu_kerevalmeth; /* 0: exp(sqrt()) evaluation                            */
  int gpu_conv_only;
  int gpu_gridder_method;
  int precision; /* added: 0 FP64, 1 FP32-mixed (complex64 data, FP32 grid)   */
#ifdef __cplusplus
  curafft_opts()
      : upsampfac(2.0), gpu_device_id(0), gpu_sort(1), gpu_binsizex(0), gpu_binsizey(0),
        gpu_binsizez(0), gpu_kerevalmeth(0), gpu_conv_only(0), gpu_gridder_method(0), precision(0) {}
#endif
};

struct curafft_plan {
  struct curafft_opts opts;
  struct conv_opts copts;
  rocfft_plan fftplan;   /* cached by the library; not owned by the plan         */
  rocfft_plan fftplan_l;
  hipStream_t *streams;  /* streams[0] is used when non-NULL                     */

  struct t3attrib ta;

  PCS *d_u; /* device u, v, w (metres), as passed to gridder_setting            */
  PCS *d_v;
  PCS *d_w;
  CUCPX *d_c; /* device visibilities                                           */

  PCS *d_x;
  PCS *d_y;
  PCS *d_z;

  int dim;
  int type;      /* 1: ms2dirty (NU -> U), 2: dirty2ms                          */
  int mode_flag; /* 1: CMCL (centred) mode ordering                            */
  int M;
  int nf1; /* grid, fast axis (v)                                              */
  int nf2; /* grid, slow axis (u)                                              */
  int nf3; /* number of w planes                                               */
  int ms;  /* image size N1 (= ny)                                             */
  int mt;  /* image size N2 (= nx)                                             */
  int mu;
  int ntransf;
  int iflag; /* +1 ms2dirty, -1 dirty2ms                                       */
  int batchsize;
  int execute_flow;

  int byte_now;
  PCS *fwkerhalf1;
  PCS *fwkerhalf2;
  PCS *fwkerhalf3;

  CUCPX *fw;
  CUCPX *fk; /* device dirty image [nx][ny]                                    */

  int *idxnupts;
  int *sortidx;
  int3 *cell_loc;
  int *binsize;
  int *binstartpts;

  PCS *d_uvw; /* added: packed [M][3] copy of (d_u, d_v, d_w) in metres       */
  PCS *d_wgt; /* added: device copy of the weights (kv.weight)               */
  PCS tol;    /* added: requested accuracy (epsilon)                          */
};

#endif
