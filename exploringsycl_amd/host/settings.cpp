// The settings a run starts from before tea.in is read.  A deck that omits a
// key gets these values, so they are the reference's defaults
// (TeaLeaf/settings.h:11-45); everything not named below starts at zero,
// false or null.
#include "settings.h"

#include <stdint.h>

#include <algorithm>
#include <string>

namespace {

// The file-name and solver-name strings are 256-byte buffers that the
// command line and the deck overwrite in place.
char* name_buffer(const std::string& text) {
  char* buffer = static_cast<char*>(calloc(256, 1));
  text.copy(buffer, 255);
  return buffer;
}

}  // namespace

void set_default_settings(Settings* settings) {
  Settings s{};

  // Mesh: 10 x 10 cells on [0, 100] x [0, 100], halos two cells deep.
  s.grid_x_min = 0.0;
  s.grid_y_min = 0.0;
  s.grid_x_max = 100.0;
  s.grid_y_max = 100.0;
  s.grid_x_cells = 10;
  s.grid_y_cells = 10;
  s.halo_depth = 2;

  // Time: steps of 0.1 up to t = 10; the step count is unbounded unless the
  // deck sets end_step.
  s.dt_init = 0.1;
  s.end_time = 10.0;
  s.end_step = INT32_MAX;
  s.summary_frequency = 10;

  // Solver: CG on K = density, stopping at an absolute |r|^2 of 1e-15 or
  // after 10000 iterations.  presteps, eps_lim and ppcg_inner_steps steer
  // the switch from CG to Chebyshev or PPCG.
  s.solver = CG_SOLVER;
  s.solver_name = name_buffer("CG");
  s.coefficient = CONDUCTIVITY;
  s.max_iters = 10000;
  s.eps = 1.0e-15;
  s.presteps = 30;
  s.eps_lim = 1.0e-5;
  s.ppcg_inner_steps = 10;
  s.check_result = true;
  s.kernel_language = C;

  // One rank driving one chunk.
  s.rank = 0;
  s.num_ranks = 1;
  s.num_chunks = 1;
  s.num_chunks_per_rank = 1;

  s.tea_in_filename = name_buffer("tea.in");
  s.tea_out_filename = name_buffer("tea.out");
  s.test_problem_filename = name_buffer("tea.problems");
  s.fields_to_exchange = static_cast<bool*>(calloc(NUM_FIELDS, sizeof(bool)));

  *settings = s;
}

void reset_fields_to_exchange(Settings* settings) {
  std::fill(settings->fields_to_exchange, settings->fields_to_exchange + NUM_FIELDS, false);
}

bool is_fields_to_exchange(Settings* settings) {
  return std::any_of(settings->fields_to_exchange,
                     settings->fields_to_exchange + NUM_FIELDS,
                     [](bool flagged) { return flagged; });
}
