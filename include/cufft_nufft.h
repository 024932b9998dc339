/* Generic NUFFT plan API (C++ linkage), replacing the reference's
 * include/cuft.h:7-8 (cunufft_setting) and src/fourier/cuft.cu:426-478
 * (cunufft_exec, curafft_free) with the same names and arguments.
 *
 * Types 1 and 2 in 1-3 dimensions and type 3 in one dimension, as the reference:
 *   type 1: fk[k3][k2][k1] = sum_j c_j e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}
 *   type 2: c_j = sum_k fk[k3][k2][k1] e^{i iflag (k1' u_j + k2' v_j + k3' w_j)}
 *   type 3: outputs s_i = plan->d_x[i] (N1 of them, set before cunufft_setting):
 *           direction 1: fk_i = sum_j c_j e^{i s_i u_j}; direction 0: c_j = sum_i fk_i e^{-i s_i u_j}
 * k' = k - N/2 (centred modes), N1 the fast index, u, v, w in radians (any
 * real value: the transform is 2 pi periodic), iflag = +1 for direction > 0,
 * else -1 (src/fourier/cuft.cu:284-285).  plan->fk is set by the caller before
 * cunufft_exec (as the reference's drivers do); cunufft_exec runs the whole
 * transform (the reference's runs only the spreading step, cuft.cu:426-444)
 * and returns after it completes.  curafft_free releases d_u, d_v, d_w, d_c,
 * fw and fk (unless opts.gpu_conv_only), as the reference's does, but not the
 * plan struct.  Precision: opts.precision 0 (complex128) or 1 (FP32-mixed,
 * complex64 data; coordinates stay FP64).  Return codes as in cugridder.h. */
#ifndef CURIG_CUFFT_NUFFT_H
#define CURIG_CUFFT_NUFFT_H

#include "curafft_plan.h"
#include "datatype.h"

#ifdef __cplusplus
int cunufft_setting(int N1, int N2, int N3, int M, int kerevalmeth, int method, int direction, PCS tol, PCS sigma,
                    int type, int dim, PCS *d_u, PCS *d_v, PCS *d_w, CUCPX *d_c, curafft_plan *plan);
int cunufft_exec(curafft_plan *plan);
int curafft_free(curafft_plan *plan);
#endif

#endif
