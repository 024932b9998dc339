/* Radio-astronomy plan structures.
 * Replaces include/ragridder_plan.h:6-54 of the reference (same names and fields). */
#ifndef CURIG_RAGRIDDER_PLAN_H
#define CURIG_RAGRIDDER_PLAN_H

#include "datatype.h"

struct visibility {
  PCS *u; /* host copies of u, v, w (metres); freed by gridder_destroy          */
  PCS *v;
  PCS *w;
  CPX *vis;
  PCS *weight; /* host weights or NULL, multiplied in both directions           */
  bool *flag;
  PCS *frequency; /* frequency[0] (Hz) is the channel gridded                 */
  int pirange;    /* 0: u, v, w in metres; 1: already multiplied by 2 pi        */
  int sign;       /* multiplies w: -1 NIFTY convention, +1 cuRIG convention    */
};

struct ragridder_plan {
  struct visibility kv;
  PCS fov;
  PCS pixelsize_x; /* fov pi / 180 / nx, pairs with u                          */
  PCS pixelsize_y; /* fov pi / 180 / ny, pairs with v                          */
  PCS w_max;
  PCS w_min;
  PCS w_s_r;
  int num_w;
  PCS w_0;
  PCS delta_w;
  int speedoflight;
  CPX *dirty_image;
  int width;  /* N1 = ny */
  int height; /* N2 = nx */
  int channel;
  int nrow;
  int cur_channel;
  int w_term_method; /* 1: improved w-stacking */
};

#endif
