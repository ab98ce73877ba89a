// Application set-up.  One process drives one GPU and one chunk that covers
// the whole grid: the deck is read, the chunk is sized, and the backend then
// allocates its fields, builds the mesh and applies the states.  The order of
// the backend calls is the one every backend expects from its host
// (TeaLeaf/initialise.c:11-31).
#include "application.h"
#include "drivers.h"
#include "tea_host.h"

State* g_states = nullptr;

void restore_initial_state(Chunk* chunks, Settings* settings) {
  set_chunk_state_driver(chunks, settings, g_states);

  // Reflect density and both energies into the depth-2 halos, then start
  // the run from energy = energy0.
  reset_fields_to_exchange(settings);
  for (int field : {FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1})
    settings->fields_to_exchange[field] = true;
  halo_update_driver(chunks, settings, 2);
  store_energy_driver(chunks, settings);
}

void initialise_application(Chunk** chunks, Settings* settings) {
  read_config(settings, &g_states);
  if (settings->num_ranks != 1 || settings->num_chunks_per_rank != 1)
    die(__LINE__, __FILE__,
        "this host runs one chunk on one rank (num_chunks_per_rank = %d, ranks = %d)\n",
        settings->num_chunks_per_rank, settings->num_ranks);
  settings->num_chunks = 1;

  Chunk* chunk = new Chunk{};
  initialise_chunk(chunk, settings, settings->grid_x_cells, settings->grid_y_cells);
  *chunks = chunk;

  kernel_initialise_driver(chunk, settings);
  set_chunk_data_driver(chunk, settings);
  restore_initial_state(chunk, settings);
}
