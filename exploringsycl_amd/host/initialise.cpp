// Application set-up (control flow of TeaLeaf/initialise.c:11-134): read
// tea.in, split the grid into chunks, then let the backend allocate, build
// the mesh, apply the states, fill the depth-2 halos and copy energy0 into
// energy.
#include <float.h>

#include "application.h"
#include "drivers.h"
#include "tea_host.h"

static void decompose_field(Settings* settings, Chunk* chunks);

State* g_states = NULL;

void initialise_application(Chunk** chunks, Settings* settings) {
  State* states = NULL;
  read_config(settings, &states);
  g_states = states;
  // One chunk per process (one GPU): halo exchange between chunks is not
  // part of this build, and the backend assigns the scalars it returns.
  if (settings->num_chunks_per_rank != 1)
    die(__LINE__, __FILE__, "num_chunks_per_rank must be 1 (one chunk per GPU)\n");

  *chunks = (Chunk*)calloc(settings->num_chunks_per_rank, sizeof(Chunk));
  decompose_field(settings, *chunks);
  kernel_initialise_driver(*chunks, settings);
  set_chunk_data_driver(*chunks, settings);
  set_chunk_state_driver(*chunks, settings, states);

  reset_fields_to_exchange(settings);
  settings->fields_to_exchange[FIELD_DENSITY] = true;
  settings->fields_to_exchange[FIELD_ENERGY0] = true;
  settings->fields_to_exchange[FIELD_ENERGY1] = true;
  halo_update_driver(*chunks, settings, 2);

  store_energy_driver(*chunks, settings);
}

// Chooses the x-by-y chunk grid whose chunks have the smallest
// perimeter-to-area ratio, gives the first (cells % chunks) columns/rows one
// extra cell, and sets each local chunk's extents and neighbours.
static void decompose_field(Settings* settings, Chunk* chunks) {
  settings->num_chunks = settings->num_ranks * settings->num_chunks_per_rank;
  const int num_chunks = settings->num_chunks;
  const double xc = (double)settings->grid_x_cells;
  const double yc = (double)settings->grid_y_cells;
  double best = DBL_MAX;
  int x_chunks = 0, y_chunks = 0;
  for (int xx = 1; xx <= num_chunks; ++xx) {
    if (num_chunks % xx) continue;
    const int yy = num_chunks / xx;
    const double perimeter = ((xc / xx) * (xc / xx) + (yc / yy) * (yc / yy)) * 2;
    const double area = (xc / xx) * (yc / yy);
    if (perimeter / area < best) {
      best = perimeter / area;
      x_chunks = xx;
      y_chunks = yy;
    }
  }
  if (!x_chunks || !y_chunks)
    die(__LINE__, __FILE__, "Failed to decompose the field with given parameters.\n");

  const int dx = settings->grid_x_cells / x_chunks;
  const int dy = settings->grid_y_cells / y_chunks;
  const int mod_x = settings->grid_x_cells % x_chunks;
  const int mod_y = settings->grid_y_cells % y_chunks;
  int add_y_prev = 0;
  for (int yy = 0; yy < y_chunks; ++yy) {
    const int add_y = yy < mod_y;
    int add_x_prev = 0;
    for (int xx = 0; xx < x_chunks; ++xx) {
      const int add_x = xx < mod_x;
      for (int cc = 0; cc < settings->num_chunks_per_rank; ++cc) {
        const int chunk = xx + yy * x_chunks;
        const int rank = cc + settings->rank * settings->num_chunks_per_rank;
        if (rank != chunk) continue;
        Chunk* c = &chunks[cc];
        initialise_chunk(c, settings, dx + add_x, dy + add_y);
        c->left = xx * dx + add_x_prev;
        c->right = c->left + dx + add_x;
        c->bottom = yy * dy + add_y_prev;
        c->top = c->bottom + dy + add_y;
        c->neighbours[CHUNK_LEFT] = xx == 0 ? EXTERNAL_FACE : chunk - 1;
        c->neighbours[CHUNK_RIGHT] = xx == x_chunks - 1 ? EXTERNAL_FACE : chunk + 1;
        c->neighbours[CHUNK_BOTTOM] = yy == 0 ? EXTERNAL_FACE : chunk - x_chunks;
        c->neighbours[CHUNK_TOP] = yy == y_chunks - 1 ? EXTERNAL_FACE : chunk + x_chunks;
      }
      add_x_prev += add_x;
    }
    add_y_prev += add_y;
  }
}
