/* Rank-level communication for the host driver.  One process drives one GPU
 * and one chunk here, so every reduction over ranks is the identity (the
 * reference's NO_MPI build, TeaLeaf/comms.c:84-101); the multi-GPU layout
 * is described in the design notes, not built. */
#pragma once
#ifndef TEA_COMMS_H
#define TEA_COMMS_H

#include "settings.h"

void barrier();
void abort_comms();
void finalise_comms();
void initialise_comms(int argc, char** argv);
void initialise_ranks(Settings* settings);
void sum_over_ranks(Settings* settings, double* a);
void min_over_ranks(Settings* settings, double* a);

#endif
