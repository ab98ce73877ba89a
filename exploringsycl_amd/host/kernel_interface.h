/* The kernel-backend boundary: the 24 entry points every TeaLeaf backend
 * exports (TeaLeaf/kernel_interface.h:13-71).  Names, argument order and
 * meaning are the reference's, so a backend that implements them links
 * against either this host or the reference's own host code.
 *
 * Scalar out-parameters follow the reference SYCL backend: rro and pw are
 * added to the caller's variable (cg.cpp:133,194); rrn, the 2-norm, the
 * Jacobi error and the field summary (vol, mass, ie, temp) are assigned
 * (cg.cpp:253, solver_methods.cpp:116, jacobi.cpp:116,
 * field_summary.cpp:147-150).  The drivers zero them first, and each process
 * drives one chunk, so the two agree. */
#pragma once
#ifndef TEA_KERNEL_INTERFACE_H
#define TEA_KERNEL_INTERFACE_H

#include "chunk.h"
#include "settings.h"

void run_set_chunk_data(Chunk* chunk, Settings* settings);
void run_set_chunk_state(Chunk* chunk, Settings* settings, State* states);
void run_kernel_initialise(Chunk* chunk, Settings* settings);
void run_kernel_finalise(Chunk* chunk, Settings* settings);

void run_local_halos(Chunk* chunk, Settings* settings, int depth);
void run_pack_or_unpack(Chunk* chunk, Settings* settings, int depth, int face,
                        bool pack, FieldBufferType field, double* buffer);
void run_store_energy(Chunk* chunk, Settings* settings);
void run_field_summary(Chunk* chunk, Settings* settings, double* vol,
                       double* mass, double* ie, double* temp);

void run_cg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                 double* rro);
void run_cg_calc_w(Chunk* chunk, Settings* settings, double* pw);
void run_cg_calc_ur(Chunk* chunk, Settings* settings, double alpha,
                    double* rrn);
void run_cg_calc_p(Chunk* chunk, Settings* settings, double beta);

void run_cheby_init(Chunk* chunk, Settings* settings);
void run_cheby_iterate(Chunk* chunk, Settings* settings, double alpha,
                       double beta);

void run_jacobi_init(Chunk* chunk, Settings* settings, double rx, double ry);
void run_jacobi_iterate(Chunk* chunk, Settings* settings, double* error);

void run_ppcg_init(Chunk* chunk, Settings* settings);
void run_ppcg_inner_iteration(Chunk* chunk, Settings* settings, double alpha,
                              double beta);

void run_copy_u(Chunk* chunk, Settings* settings);
void run_calculate_residual(Chunk* chunk, Settings* settings);
void run_calculate_2norm(Chunk* chunk, Settings* settings,
                         FieldBufferType buffer, double* norm);
void run_finalise(Chunk* chunk, Settings* settings);

#endif
