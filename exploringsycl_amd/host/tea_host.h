/* Host-only additions: the run record written as JSON (per-step iteration
 * counts and field summaries), the bench controls, and the optional backend
 * hook used to bracket the timed region.  None of this crosses
 * kernel_interface.h. */
#pragma once
#ifndef TEA_HOST_H
#define TEA_HOST_H

#include <string>
#include <vector>

#include "chunk.h"

struct StepRecord {
  int step;           /* 1-based timestep */
  int iterations;     /* as printed: "CG: N iterations" */
  int executed;       /* CG iterations actually run (printed + 1 on convergence) */
  int refreshes;      /* FP64 true-residual refreshes (mixed precision) */
  int fallback_at;    /* FCG: iteration at which Jacobi took over (-1: never) */
  double fallback_beta = 0.0; /* FCG: the beta of that iteration (the restart) */
  double rel_residual; /* final |b - A u| / |b| where the solver computes it */
  double error;       /* last rr of the solve */
  double seconds;     /* wall time of the whole timestep */
  /* Chebyshev / PPCG: iterations after the switch from CG as printed, the
   * printed Chebyshev estimate, and the eigenvalue bounds (-1 / 0: none). */
  int solver_iterations = -1;
  int estimated = -1;
  double eigmin = 0.0, eigmax = 0.0;
};

struct SummaryRecord {
  int step;           /* timestep after which it was taken (0 = initial) */
  bool final_check;   /* the end-of-run summary checked against tea.problems */
  double vol, mass, ie, temp;
};

struct RunRecord {
  std::vector<StepRecord> steps;
  std::vector<SummaryRecord> summaries;
  double check_value = 0.0;   /* expected temperature, if found */
  bool check_found = false;
  bool check_passed = false;
  double timed_seconds = 0.0; /* the steps after warm-up, device-synchronised */
  int warmup_steps = 0;
  std::string coarse_solve = "none"; /* FCG: "network", "jacobi sweeps N" or "none" */
};

extern RunRecord g_run;

/* Extra command-line controls (not in the reference). */
struct HostOptions {
  std::string json_path;   /* --json FILE: write g_run as JSON at the end */
  int warmup_steps = 0;    /* --warmup-steps W: run W steps, reset, then time */
  bool quiet = false;      /* --quiet: log to tea.out only */
  std::string field_out;   /* --field-out FILE: final interior u, raw float64 */
  int x_cells = 0;         /* -x CELLS: overrides the deck's x_cells (0: none) */
  int y_cells = 0;         /* -y CELLS: overrides the deck's y_cells (0: none) */
};

extern HostOptions g_opts;

/* Solver settings beyond the reference's tea.in keys (parsed from tea.in;
 * kept out of Settings so its layout stays the reference's). */
struct SolverExtras {
  double rel_eps = 0.0;      /* rel_eps=: stop when |r|/|b| < rel_eps (0: off) */
  int refresh_interval = 50; /* refresh_interval=: FP64 true residual period */
  std::string fcg_weights;   /* fcg_weights=: the preconditioner network file */
  int fcg_levels = 2;        /* fcg_levels=: coarse levels (1 or 2) */
  int fcg_sweeps = 0;        /* fcg_coarse_sweeps=: N > 0 replaces the network by N
                                Jacobi sweeps on each coarse operator */
  int stall_window = 400;    /* stall_window=: FCG stagnation window (iterations) */
  double stall_factor = 4.0; /* stall_factor=: least |r| reduction over the window */
};

/* The learned preconditioner's network file: one ASCII line
 *   "tea-nn-weights 1 <count> <coarsen> <m0> <s0> <m1> <s1>\n"
 * followed by <count> little-endian float32 parameters.  Returns false (and
 * sets *why) if the file cannot be read or is malformed. */
struct NetWeights {
  std::vector<float> params;
  float norm[4] = {0.f, 1.f, 0.f, 1.f};
  int coarsen = 0;
};
bool read_net_weights(const char* path, NetWeights* out, std::string* why);

extern SolverExtras g_extra;

/* The states read from tea.in (kept for --warmup-steps resets). */
extern State* g_states;

/* Optional backend hook (weak): called with "timed_begin" and "timed_end"
 * around the timed steps so a backend can reset and report its kernel
 * timers.  Backends that do not define it are unaffected. */
extern "C" void tea_backend_mark(const char* event) __attribute__((weak));

void write_run_json(const char* path, Settings* settings);

#endif
