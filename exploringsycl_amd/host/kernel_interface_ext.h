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
 * external.  calc_w returns p.w and p.r, calc_r and refresh r.r and r.w,
 * for the flexible CG(1) recursion alpha = p.r / p.w, beta = -r.w / p.w. */
void run_mpcg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                   double* rr, double* bb);
void run_mpcg_calc_w(Chunk* chunk, Settings* settings, double alpha,
                     int apply_u, double beta, int apply_p, double* pw,
                     double* pr);
void run_mpcg_calc_r(Chunk* chunk, Settings* settings, double alpha,
                     double* rr, double* rw);
void run_mpcg_refresh(Chunk* chunk, Settings* settings, double alpha,
                      int apply_u, double* rr, double* rr_true, double* rw);

/* Flexible CG, FCG(1), with the FP32 vectors and FP64 scalars of the
 * mixed-precision CG.  z = M(r): Jacobi, or Jacobi plus a two-level
 * coarse-grid correction (the learned network, or Jacobi sweeps).
 *   run_fcg_attach: allocates the FCG state; params != NULL loads the
 *     network (count parameters, feature normalisation norm[4], coarsening
 *     factor coarsen) on up to `levels` coarse levels (coarsen * 4^l).
 *   run_fcg_init: as run_mpcg_init, plus the network's coefficient features.
 *   run_fcg_precondition: z = M(r) (with the coarse correction if use_net,
 *     else Jacobi); zw += z.w.
 *   run_fcg_calc_w: u += alpha p_old if apply_u; p = z + beta p_old; w = A p;
 *     pw += p.w, pr += p.r.
 *   run_fcg_calc_r_precondition: r -= alpha w, rr += r.r, then z = M(r),
 *     zw += z.w, with one wait for both results.
 * The residual update and the FP64 refresh are run_mpcg_calc_r and
 * run_mpcg_refresh. */
void run_fcg_attach(Chunk* chunk, Settings* settings, const float* params,
                    size_t count, const float* norm, int coarsen, int levels);
/*   run_fcg_attach_sweeps: as run_fcg_attach, with `sweeps` Jacobi sweeps
 *     on each level's Galerkin coarse operator (from zero) in place of the
 *     network: the non-learned two-level preconditioner (coarsen 8). */
void run_fcg_attach_sweeps(Chunk* chunk, Settings* settings, int sweeps,
                           int coarsen, int levels);
void run_fcg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                  double* rr, double* bb);
void run_fcg_precondition(Chunk* chunk, Settings* settings, int use_net,
                          double* zw);
void run_fcg_calc_w(Chunk* chunk, Settings* settings, double alpha,
                    int apply_u, double beta, double* pw, double* pr);
void run_fcg_calc_r_precondition(Chunk* chunk, Settings* settings,
                                 double alpha, int use_net, double* rr,
                                 double* zw);

/* Copies the chunk's u (x*y values, halo included) to host memory. */
void run_read_u(Chunk* chunk, Settings* settings, double* host);

#endif
