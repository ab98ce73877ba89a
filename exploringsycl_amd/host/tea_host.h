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
  double rel_residual; /* final |b - A u| / |b| where the solver computes it */
  double error;       /* last rr of the solve */
  double seconds;     /* wall time of the whole timestep */
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
};

extern RunRecord g_run;

/* Extra command-line controls (not in the reference). */
struct HostOptions {
  std::string json_path;   /* --json FILE: write g_run as JSON at the end */
  int warmup_steps = 0;    /* --warmup-steps W: run W steps, reset, then time */
  bool quiet = false;      /* --quiet: log to tea.out only */
};

extern HostOptions g_opts;

/* Solver settings beyond the reference's tea.in keys (parsed from tea.in;
 * kept out of Settings so its layout stays the reference's). */
struct SolverExtras {
  double rel_eps = 0.0;      /* rel_eps=: stop when |r|/|b| < rel_eps (0: off) */
  int refresh_interval = 50; /* refresh_interval=: FP64 true residual period */
};

extern SolverExtras g_extra;

/* The states read from tea.in (kept for --warmup-steps resets). */
extern State* g_states;

/* Optional backend hook (weak): called with "timed_begin" and "timed_end"
 * around the timed steps so a backend can reset and report its kernel
 * timers.  Backends that do not define it are unaffected. */
extern "C" void tea_backend_mark(const char* event) __attribute__((weak));

void write_run_json(const char* path, Settings* settings);

#endif
