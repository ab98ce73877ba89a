// The timestep loop (control flow of TeaLeaf/diffuse.c:10-112): per step,
// refresh the depth-2 halos of energy and density, run the solver, finish
// the step, and take a field summary every summary_frequency steps; after
// the last step, a final summary checked against tea.problems.
#include <time.h>

#include "application.h"
#include "comms.h"
#include "drivers.h"
#include "tea_host.h"

static double now_seconds() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

// run_start: when the first timestep of this diffuse() call began.
static void solve(Chunk* chunks, Settings* settings, int tt, double run_start) {
  print_and_log(settings, "\nTimestep %d\n", tt + 1);
  const double t0 = now_seconds();
  g_run.steps.push_back({tt + 1, 0, 0, 0, -1, 0.0, 0.0, 0.0});

  double dt = settings->dt_init;
  for (int cc = 0; cc < settings->num_chunks_per_rank; ++cc)
    if (chunks[cc].dt_init < dt) dt = chunks[cc].dt_init;
  min_over_ranks(settings, &dt);

  const double rx = dt / (settings->dx * settings->dx);
  const double ry = dt / (settings->dy * settings->dy);

  reset_fields_to_exchange(settings);
  settings->fields_to_exchange[FIELD_ENERGY1] = true;
  settings->fields_to_exchange[FIELD_DENSITY] = true;
  halo_update_driver(chunks, settings, 2);

  double error = 1e+10;
  switch (settings->solver) {
    case JACOBI_SOLVER:
      jacobi_driver(chunks, settings, rx, ry, &error);
      break;
    case CHEBY_SOLVER:
      cheby_driver(chunks, settings, rx, ry, &error);
      break;
    case PPCG_SOLVER:
      ppcg_driver(chunks, settings, rx, ry, &error);
      break;
    case CG_SOLVER:
      cg_driver(chunks, settings, rx, ry, &error);
      break;
    case MPCG_SOLVER:
      mpcg_driver(chunks, settings, rx, ry, &error);
      break;
    case FCG_SOLVER:
    case JACOBI_PCG_SOLVER:
      fcg_driver(chunks, settings, rx, ry, &error);
      break;
    default:
      die(__LINE__, __FILE__,
          "solver '%s' is not built in this host\n",
          settings->solver_name);
  }
  g_run.steps.back().error = error;

  solve_finished_driver(chunks, settings);

  if (tt % settings->summary_frequency == 0)
    field_summary_driver(chunks, settings, false);

  const double t1 = now_seconds();
  const double step_seconds = t1 - t0;
  g_run.steps.back().seconds = step_seconds;
  print_and_log(settings, "Wallclock: \t\t%.3lfs\n", t1 - run_start);
  print_and_log(settings, "Avg. time per cell: \t%.6e\n",
                step_seconds / ((double)settings->grid_x_cells * settings->grid_y_cells));
  print_and_log(settings, "Error: \t\t\t%.6e\n", error);
}

void diffuse(Chunk* chunks, Settings* settings) {
  const double run_start = now_seconds();
  for (int tt = 0; tt < settings->end_step; ++tt)
    solve(chunks, settings, tt, run_start);
  field_summary_driver(chunks, settings, true);
}
