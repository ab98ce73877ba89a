/* Oracle (CPU) backend half of the kernel_interface.h plugin contract:
 * fields are host arrays.  TEST INFRASTRUCTURE ONLY. */
#pragma once
#ifndef TEA_ORACLE_CHUNK_EXTENSION_H
#define TEA_ORACLE_CHUNK_EXTENSION_H

typedef double* FieldBufferType;

typedef struct ChunkExtension {
  FieldBufferType comms_buffer;
  /* Mixed-precision CG arrays (this project's addition): FP32 p (two
   * buffers, mp_cur is current), r, w, kx, ky. */
  float* mp_p[2];
  int mp_cur;
  float* mp_r;
  float* mp_w;
  float* mp_kx;
  float* mp_ky;
  /* Flexible CG: the preconditioned residual z, and the coarse levels of
   * the Jacobi-sweep coarse correction (NULL: Jacobi only). */
  float* fcg_z;
  struct OracleCoarse* coarse;
} ChunkExtension;

#endif
