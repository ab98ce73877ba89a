// The host side of a chunk.  This host runs one chunk per process, covering
// the whole grid, so a chunk carries only its padded extents, its first dt,
// its place in the grid and its four neighbours, all of which are the domain
// boundary.  The face buffers of a multi-rank halo exchange stay null; the
// backend fills in the extension when the kernels are initialised.
#include "chunk.h"

#include <algorithm>

void initialise_chunk(Chunk* chunk, Settings* settings, int x, int y) {
  *chunk = Chunk{};
  const int halo_cells = 2 * settings->halo_depth;
  chunk->x = x + halo_cells;
  chunk->y = y + halo_cells;
  chunk->dt_init = settings->dt_init;

  chunk->left = 0;
  chunk->right = x;
  chunk->bottom = 0;
  chunk->top = y;
  chunk->neighbours = new int[NUM_FACES];
  std::fill(chunk->neighbours, chunk->neighbours + NUM_FACES, EXTERNAL_FACE);

  chunk->ext = new ChunkExtension{};
}

void finalise_chunk(Chunk* chunk) {
  delete[] chunk->neighbours;
  delete chunk->ext;
  chunk->neighbours = nullptr;
  chunk->ext = nullptr;
}
