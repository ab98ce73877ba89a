/* Host-side solver drivers: each loops over the rank's chunks and calls the
 * kernel_interface entry points (TeaLeaf/drivers/drivers.h:5-72). */
#pragma once
#ifndef TEA_DRIVERS_H
#define TEA_DRIVERS_H

#include "chunk.h"

void set_chunk_data_driver(Chunk* chunk, Settings* settings);
void set_chunk_state_driver(Chunk* chunk, Settings* settings, State* states);
void kernel_initialise_driver(Chunk* chunks, Settings* settings);
void kernel_finalise_driver(Chunk* chunks, Settings* settings);

void halo_update_driver(Chunk* chunks, Settings* settings, int depth);
void remote_halo_driver(Chunk* chunks, Settings* settings, int depth);

void cg_driver(Chunk* chunks, Settings* settings, double rx, double ry,
               double* error);
void cg_init_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                    double* rro);
void cg_main_step_driver(Chunk* chunks, Settings* settings, int tt,
                         double* rro, double* error);

/* The reference's other solvers (drivers_jacobi_cheby_ppcg.cpp): Jacobi,
 * Chebyshev after CG steps that estimate the eigenvalue bounds, and CG
 * preconditioned by ppcg_inner_steps Chebyshev steps (PPCG). */
void jacobi_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                   double* error);
void cheby_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                  double* error);
void ppcg_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                 double* error);

/* Mixed-precision CG (this project's addition): relative-residual stop
 * (rel_eps, default 1e-10) confirmed by the FP64 true residual. */
void mpcg_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                 double* error);

/* Flexible CG with the learned preconditioner (FCG_SOLVER, network file
 * from fcg_weights=) or with Jacobi alone (JACOBI_PCG_SOLVER); the same
 * stop as mpcg_driver.  A learned solve whose |r| falls by less than
 * stall_factor over stall_window iterations restarts as Jacobi-PCG. */
void fcg_driver(Chunk* chunks, Settings* settings, double rx, double ry,
                double* error);

void field_summary_driver(Chunk* chunks, Settings* settings,
                          bool solve_finished);
void store_energy_driver(Chunk* chunk, Settings* settings);
void solve_finished_driver(Chunk* chunks, Settings* settings);

#endif
