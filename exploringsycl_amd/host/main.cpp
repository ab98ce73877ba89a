// tealeaf: the host program.  Usage:
//   tealeaf [--in tea.in] [--out tea.out] [--problems tea.problems]
//           [-s|--solver cg] [-x CELLS] [-y CELLS] [--json FILE]
//           [--warmup-steps W] [--quiet] [--field-out FILE]
// Reads the deck, sets up the chunk through the kernel backend it is linked
// with, runs end_step timesteps and checks the final temperature against the
// tea.problems table (TeaLeaf/main.c:9-99 for the reference's flow).
// --warmup-steps W first runs W untimed steps, puts the initial state back
// and then times the end_step steps; --json writes the run record.
#include <string.h>
#include <time.h>

#include <vector>

#include "application.h"
#include "comms.h"
#include "drivers.h"
#include "kernel_interface.h"
#include "kernel_interface_ext.h"
#include "tea_host.h"

RunRecord g_run;
HostOptions g_opts;
SolverExtras g_extra;

static double now_seconds() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void usage(Settings* settings) {
  printf("tealeaf <options>\n");
  printf("options:\n");
  printf("\t-solver, --solver, -s: 'cg' (the other reference solvers are not built)\n");
  printf("\t-x CELLS, -y CELLS: grid size (override tea.in)\n");
  printf("\t--in FILE, --out FILE, --problems FILE: input deck, log, known answers\n");
  printf("\t--json FILE: write the run record (iterations, summaries, timing)\n");
  printf("\t--warmup-steps W: run W untimed steps first, then reset and time\n");
  printf("\t--quiet: write the log only\n");
  printf("\t--field-out FILE: write the final interior u as raw float64 (row-major)\n");
  (void)settings;
}

static void set_name(char* dst, const char* src) {
  strncpy(dst, src, 255);
  dst[255] = 0;
}

// Command-line settings that must be known before the deck is read.
static void parse_args(Settings* settings, int argc, char** argv) {
  for (int aa = 1; aa < argc; ++aa) {
    const char* a = argv[aa];
    const bool has_val = aa + 1 < argc;
    if ((!strcmp(a, "-solver") || !strcmp(a, "--solver") || !strcmp(a, "-s")) && has_val) {
      const char* v = argv[++aa];
      if (!strcmp(v, "cg")) settings->solver = CG_SOLVER;
      else if (!strcmp(v, "cheby")) settings->solver = CHEBY_SOLVER;
      else if (!strcmp(v, "ppcg")) settings->solver = PPCG_SOLVER;
      else if (!strcmp(v, "jacobi")) settings->solver = JACOBI_SOLVER;
    } else if (!strcmp(a, "-x") && has_val) {
      g_opts.x_cells = atoi(argv[++aa]);
    } else if (!strcmp(a, "-y") && has_val) {
      g_opts.y_cells = atoi(argv[++aa]);
    } else if (!strcmp(a, "--in") && has_val) {
      set_name(settings->tea_in_filename, argv[++aa]);
    } else if (!strcmp(a, "--out") && has_val) {
      set_name(settings->tea_out_filename, argv[++aa]);
    } else if (!strcmp(a, "--problems") && has_val) {
      set_name(settings->test_problem_filename, argv[++aa]);
    } else if (!strcmp(a, "--json") && has_val) {
      g_opts.json_path = argv[++aa];
    } else if (!strcmp(a, "--warmup-steps") && has_val) {
      g_opts.warmup_steps = atoi(argv[++aa]);
    } else if (!strcmp(a, "--quiet")) {
      g_opts.quiet = true;
    } else if (!strcmp(a, "--field-out") && has_val) {
      g_opts.field_out = argv[++aa];
    } else if (!strcmp(a, "-h") || !strcmp(a, "-help") || !strcmp(a, "--help")) {
      usage(settings);
      exit(0);
    } else {
      fprintf(stderr, "unknown argument '%s' (try --help)\n", a);
      exit(2);
    }
  }
}

static void write_field(const char* path, Chunk* chunks, Settings* settings);

int main(int argc, char** argv) {
  initialise_comms(argc, argv);
  Settings* settings = (Settings*)malloc(sizeof(Settings));
  set_default_settings(settings);
  initialise_ranks(settings);
  parse_args(settings, argc, argv);

  Chunk* chunks = NULL;
  initialise_application(&chunks, settings);

  if (g_opts.warmup_steps > 0) {
    const int end_step = settings->end_step;
    const bool check = settings->check_result;
    settings->end_step = g_opts.warmup_steps;
    settings->check_result = false;
    diffuse(chunks, settings);
    settings->end_step = end_step;
    settings->check_result = check;
    const std::string coarse_solve = g_run.coarse_solve;
    g_run = RunRecord();
    g_run.coarse_solve = coarse_solve;
    g_run.warmup_steps = g_opts.warmup_steps;
    restore_initial_state(chunks, settings);
    print_and_log(settings, "\n--- warm-up done, timed run follows ---\n");
  }

  if (tea_backend_mark) tea_backend_mark("timed_begin");
  const double t0 = now_seconds();
  diffuse(chunks, settings);
  if (tea_backend_mark) tea_backend_mark("timed_end");
  g_run.timed_seconds = now_seconds() - t0;

  if (!g_opts.json_path.empty()) write_run_json(g_opts.json_path.c_str(), settings);
  if (!g_opts.field_out.empty()) write_field(g_opts.field_out.c_str(), chunks, settings);

  kernel_finalise_driver(chunks, settings);
  finalise_chunk(chunks);
  delete chunks;
  free(g_states);
  if (settings->tea_out_fp) fclose(settings->tea_out_fp);
  finalise_comms();
  return g_run.check_found && !g_run.check_passed ? 3 : EXIT_SUCCESS;
}

// The interior of chunk 0's u, row by row, as raw float64.
static void write_field(const char* path, Chunk* chunks, Settings* settings) {
  Chunk* c = &chunks[0];
  std::vector<double> u((size_t)c->x * c->y);
  run_read_u(c, settings, u.data());
  FILE* fp = fopen(path, "wb");
  if (!fp) die(__LINE__, __FILE__, "Could not open %s\n", path);
  const int h = settings->halo_depth;
  for (int jj = h; jj < c->y - h; ++jj)
    fwrite(&u[(size_t)jj * c->x + h], sizeof(double), (size_t)(c->x - 2 * h), fp);
  fclose(fp);
}

void write_run_json(const char* path, Settings* settings) {
  FILE* fp = fopen(path, "w");
  if (!fp) die(__LINE__, __FILE__, "Could not open %s\n", path);
  fprintf(fp, "{\n  \"x_cells\": %d, \"y_cells\": %d, \"end_step\": %d,\n",
          settings->grid_x_cells, settings->grid_y_cells, settings->end_step);
  fprintf(fp, "  \"coefficient\": %d, \"eps\": %.17g, \"max_iters\": %d,\n",
          settings->coefficient, settings->eps, settings->max_iters);
  fprintf(fp, "  \"solver\": \"%s\", \"rel_eps\": %.17g,\n", settings->solver_name,
          g_extra.rel_eps);
  fprintf(fp, "  \"warmup_steps\": %d, \"timed_seconds\": %.9f,\n",
          g_run.warmup_steps, g_run.timed_seconds);
  fprintf(fp, "  \"coarse_solve\": \"%s\",\n", g_run.coarse_solve.c_str());
  fprintf(fp, "  \"steps\": [");
  for (size_t i = 0; i < g_run.steps.size(); ++i) {
    const StepRecord& s = g_run.steps[i];
    fprintf(fp, "%s\n    {\"step\": %d, \"iterations\": %d, \"executed\": %d, "
                "\"refreshes\": %d, \"fallback_at\": %d, \"fallback_beta\": %.17g, "
                "\"rel_residual\": %.17g, "
                "\"error\": %.17g, \"seconds\": %.9f, \"solver_iterations\": %d, "
                "\"estimated\": %d, \"eigmin\": %.17g, \"eigmax\": %.17g}",
            i ? "," : "", s.step, s.iterations, s.executed, s.refreshes,
            s.fallback_at, s.fallback_beta, s.rel_residual, s.error, s.seconds,
            s.solver_iterations, s.estimated, s.eigmin, s.eigmax);
  }
  fprintf(fp, "\n  ],\n  \"summaries\": [");
  for (size_t i = 0; i < g_run.summaries.size(); ++i) {
    const SummaryRecord& s = g_run.summaries[i];
    fprintf(fp, "%s\n    {\"step\": %d, \"final\": %s, \"vol\": %.17g, "
                "\"mass\": %.17g, \"ie\": %.17g, \"temp\": %.17g}",
            i ? "," : "", s.step, s.final_check ? "true" : "false", s.vol,
            s.mass, s.ie, s.temp);
  }
  fprintf(fp, "\n  ],\n  \"check_found\": %s, \"check_value\": %.17g, "
              "\"check_passed\": %s\n}\n",
          g_run.check_found ? "true" : "false", g_run.check_value,
          g_run.check_passed ? "true" : "false");
  fclose(fp);
}
