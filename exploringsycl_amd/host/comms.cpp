// Single-rank communication layer: one process, one GPU, one chunk.  All
// rank reductions are the identity, as in the reference's NO_MPI build
// (TeaLeaf/comms.c:84-101).
#include "comms.h"

#include <stdlib.h>

void initialise_comms(int argc, char** argv) { (void)argc; (void)argv; }
void initialise_ranks(Settings* settings) {
  settings->rank = MASTER;
  settings->num_ranks = 1;
}
void finalise_comms() {}
void sum_over_ranks(Settings* settings, double* a) { (void)settings; (void)a; }
void min_over_ranks(Settings* settings, double* a) { (void)settings; (void)a; }
void barrier() {}
void abort_comms() { exit(1); }
