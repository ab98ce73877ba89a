/* Backend half of the kernel_interface.h plugin contract for the HIP build:
 * the field handle type and the per-chunk extension
 * (TeaLeaf/c_kernels/sycl/chunk_extension.h:6-13 for the SYCL backend).
 *
 * FieldBufferType is a device address of FP64 cells.  The host never
 * dereferences it; it only hands it back across the interface (e.g.
 * run_calculate_2norm(chunk, settings, chunk->r, &norm)). */
#pragma once
#ifndef TEA_HIP_CHUNK_EXTENSION_H
#define TEA_HIP_CHUNK_EXTENSION_H

typedef double* FieldBufferType;

struct tea_hip_chunk;

typedef struct ChunkExtension {
  FieldBufferType comms_buffer;   /* face buffer, max(x, y) * halo_depth */
  struct tea_hip_chunk* hip;      /* device state of this chunk */
} ChunkExtension;

#endif
