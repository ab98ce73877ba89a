/* Run settings, as parsed from tea.in.  The Settings and State layouts are
 * the reference's (TeaLeaf/settings.h:64-145) so that a kernel backend built
 * against these headers is link- and layout-compatible with the reference
 * host; the NEW solver switches live in the separate SolverExtras block that
 * only this host uses. */
#pragma once
#ifndef TEA_SETTINGS_H
#define TEA_SETTINGS_H

#include "shared.h"

#define NUM_FIELDS 6

/* The first four are the reference's; the rest are this project's
 * additions (appended, so the reference's values are unchanged). */
typedef enum {
  JACOBI_SOLVER,
  CG_SOLVER,
  CHEBY_SOLVER,
  PPCG_SOLVER,
  MPCG_SOLVER,       /* mixed-precision CG */
  FCG_SOLVER,        /* flexible CG with the learned preconditioner */
  JACOBI_PCG_SOLVER  /* diagonally preconditioned CG */
} Solver;

typedef enum { C, FORTRAN } Kernel_Language;

typedef struct Settings {
  struct Profile* kernel_profile;
  struct Profile* application_profile;
  struct Profile* wallclock_profile;

  FILE* tea_out_fp;

  int rank;
  int end_step;
  int presteps;
  int max_iters;
  int coefficient;
  int ppcg_inner_steps;
  int summary_frequency;
  int halo_depth;
  int num_states;
  int num_chunks;
  int num_chunks_per_rank;
  int num_ranks;
  bool* fields_to_exchange;

  bool is_offload;

  bool error_switch;
  bool check_result;
  bool preconditioner;

  double eps;
  double dt_init;
  double end_time;
  double eps_lim;

  char* tea_in_filename;
  char* tea_out_filename;
  char* test_problem_filename;

  Solver solver;
  char* solver_name;

  Kernel_Language kernel_language;

  int grid_x_cells;
  int grid_y_cells;

  double grid_x_min;
  double grid_y_min;
  double grid_x_max;
  double grid_y_max;

  double dx;
  double dy;
} Settings;

typedef enum { RECTANGULAR, CIRCULAR, POINT } Geometry;

typedef struct {
  bool defined;
  double density;
  double energy;
  double x_min;
  double y_min;
  double x_max;
  double y_max;
  double radius;
  Geometry geometry;
} State;

void set_default_settings(Settings* settings);
void reset_fields_to_exchange(Settings* settings);
bool is_fields_to_exchange(Settings* settings);

#endif
