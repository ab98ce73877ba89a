/* Entry points of the solvers this project adds beside the reference's CG
 * (no counterpart in TeaLeaf/kernel_interface.h; the reference host never
 * calls them).  Same conventions as kernel_interface.h: scalar results are
 * accumulated (+=) into the driver's variables. */
#pragma once
#ifndef TEA_KERNEL_INTERFACE_EXT_H
#define TEA_KERNEL_INTERFACE_EXT_H

#include "chunk.h"
#include "settings.h"

/* Mixed-precision CG: FP32 p, r, w, kx, ky; FP64 u, dot products, alpha,
 * beta and a periodic FP64 true-residual refresh.  One chunk, all faces
 * external. */
void run_mpcg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                   double* rr, double* bb);
void run_mpcg_calc_w(Chunk* chunk, Settings* settings, double alpha,
                     int apply_u, double beta, int apply_p, double* pw);
void run_mpcg_calc_r(Chunk* chunk, Settings* settings, double alpha,
                     double* rr);
void run_mpcg_refresh(Chunk* chunk, Settings* settings, double alpha,
                      int apply_u, double* rr, double* rr_true);

#endif
