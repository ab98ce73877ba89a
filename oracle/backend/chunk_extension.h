/* Oracle (CPU) backend half of the kernel_interface.h plugin contract:
 * fields are host arrays.  TEST INFRASTRUCTURE ONLY. */
#pragma once
#ifndef TEA_ORACLE_CHUNK_EXTENSION_H
#define TEA_ORACLE_CHUNK_EXTENSION_H

typedef double* FieldBufferType;

typedef struct ChunkExtension {
  FieldBufferType comms_buffer;
} ChunkExtension;

#endif
