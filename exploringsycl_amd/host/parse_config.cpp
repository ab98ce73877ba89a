// tea.in reader.  The grammar is TeaLeaf's (TeaLeaf/parse_config.c:14-312):
//  - a line whose first non-blank characters are a key ("x_cells", "eps",
//    ...) sets that key; its value is the first token after the key that
//    starts with a letter or digit (so "eps 1.0e-15" and "eps=1.0e-15" both
//    work, and a leading minus sign is not part of the value);
//  - a line starting with a switch name ("use_cg", "check_result", ...)
//    turns that switch on;
//  - "state N density=.. energy=.. geometry=rectangle xmin=.. ..." defines
//    state N; the bounds of states 2..n are moved inwards by 1/100 of a cell
//    (parse_config.c:222-260) so that a bound on a cell edge cannot fall on
//    either side by rounding.
// -x/-y on the command line override the deck's x_cells / y_cells.
#include <ctype.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "application.h"
#include "tea_host.h"

namespace {

// True if `line`, after leading blanks, starts with `key`.
bool starts_with(const char* key, const char* line) {
  while (*line && isspace((unsigned char)*line)) ++line;
  return strncmp(line, key, strlen(key)) == 0;
}

// The value following the first occurrence of `key` in `line`: the first
// whitespace-free token that begins with a letter or a digit.  A missing
// value is fatal, as in the reference.  The occurrence is found by the same
// forward scan the reference uses (a partial match that fails restarts at
// the failing character).
std::string value_after(const char* line, const char* key) {
  const size_t n = strlen(line), klen = strlen(key);
  size_t matched = 0;
  for (size_t i = 0; i < n; ++i) {
    if (matched == klen) {
      if (isalnum((unsigned char)line[i])) {
        size_t j = i;
        while (j < n && !isspace((unsigned char)line[j])) ++j;
        return std::string(line + i, j - i);
      }
    } else {
      matched = line[i] == key[matched] ? matched + 1 : 0;
    }
  }
  die(__LINE__, __FILE__, "Failed to find a value for key '%s'\n", key);
  return std::string();
}

bool get_int(const char* key, const char* line, int* out) {
  if (!starts_with(key, line)) return false;
  *out = atoi(value_after(line, key).c_str());
  return true;
}

bool get_double(const char* key, const char* line, double* out) {
  if (!starts_with(key, line)) return false;
  *out = atof(value_after(line, key).c_str());
  return true;
}

std::vector<std::string> read_lines(FILE* fp) {
  std::vector<std::string> lines;
  char* buf = NULL;
  size_t cap = 0;
  while (getline(&buf, &cap, fp) != -1) lines.emplace_back(buf);
  free(buf);
  return lines;
}

void read_settings(const std::vector<std::string>& lines, Settings* s) {
  for (const std::string& ln : lines) {
    const char* line = ln.c_str();
    // Keys, in the reference's precedence order ("epslim" before "eps").
    if (get_double("initial_timestep", line, &s->dt_init)) continue;
    if (get_double("end_time", line, &s->end_time)) continue;
    if (get_int("end_step", line, &s->end_step)) continue;
    if (get_double("xmin", line, &s->grid_x_min)) continue;
    if (get_double("ymin", line, &s->grid_y_min)) continue;
    if (get_double("xmax", line, &s->grid_x_max)) continue;
    if (get_double("ymax", line, &s->grid_y_max)) continue;
    if (get_int("x_cells", line, &s->grid_x_cells)) continue;
    if (get_int("y_cells", line, &s->grid_y_cells)) continue;
    if (get_int("summary_frequency", line, &s->summary_frequency)) continue;
    if (get_int("presteps", line, &s->presteps)) continue;
    if (get_int("ppcg_inner_steps", line, &s->ppcg_inner_steps)) continue;
    if (get_double("epslim", line, &s->eps_lim)) continue;
    if (get_int("max_iters", line, &s->max_iters)) continue;
    if (get_double("eps", line, &s->eps)) continue;
    if (get_int("num_chunks_per_rank", line, &s->num_chunks_per_rank)) continue;
    if (get_int("halo_depth", line, &s->halo_depth)) continue;
    // Keys of this project's solvers.
    if (get_double("rel_eps", line, &g_extra.rel_eps)) continue;
    if (get_int("refresh_interval", line, &g_extra.refresh_interval)) continue;
    if (get_int("stall_window", line, &g_extra.stall_window)) continue;
    if (get_int("fcg_levels", line, &g_extra.fcg_levels)) continue;
    if (get_int("fcg_coarse_sweeps", line, &g_extra.fcg_sweeps)) continue;
    if (get_double("stall_factor", line, &g_extra.stall_factor)) continue;
    if (starts_with("fcg_weights", line)) {
      g_extra.fcg_weights = value_after(line, "fcg_weights");
      continue;
    }

    // Switches.
    if (starts_with("check_result", line)) { s->check_result = true; continue; }
    if (starts_with("errswitch", line)) { s->error_switch = true; continue; }
    if (starts_with("preconditioner_on", line)) { s->preconditioner = true; continue; }
    if (starts_with("use_fortran_kernels", line)) { s->kernel_language = FORTRAN; continue; }
    if (starts_with("use_c_kernels", line)) { s->kernel_language = C; continue; }
    if (starts_with("use_jacobi_pcg", line)) {
      s->solver = JACOBI_PCG_SOLVER;
      strcpy(s->solver_name, "Jacobi-PCG");
      continue;
    }
    if (starts_with("use_mpcg", line)) {
      s->solver = MPCG_SOLVER;
      strcpy(s->solver_name, "MPCG");
      continue;
    }
    if (starts_with("use_fcg", line)) {
      s->solver = FCG_SOLVER;
      strcpy(s->solver_name, "FCG");
      continue;
    }
    if (starts_with("use_jacobi", line)) {
      s->solver = JACOBI_SOLVER;
      strcpy(s->solver_name, "Jacobi");
      continue;
    }
    if (starts_with("use_cg", line)) {
      s->solver = CG_SOLVER;
      strcpy(s->solver_name, "CG");
      continue;
    }
    if (starts_with("use_chebyshev", line)) {
      s->solver = CHEBY_SOLVER;
      strcpy(s->solver_name, "Chebyshev");
      continue;
    }
    if (starts_with("use_ppcg", line)) {
      s->solver = PPCG_SOLVER;
      strcpy(s->solver_name, "PPCG");
      continue;
    }
    if (starts_with("coefficient_density", line)) {
      s->coefficient = CONDUCTIVITY;
      continue;
    }
    if (starts_with("coefficient_inverse_density", line)) {
      s->coefficient = RECIP_CONDUCTIVITY;
      continue;
    }
  }
  // -x / -y on the command line take precedence over the deck.
  if (g_opts.x_cells > 0) s->grid_x_cells = g_opts.x_cells;
  if (g_opts.y_cells > 0) s->grid_y_cells = g_opts.y_cells;
  s->dx = (s->grid_x_max - s->grid_x_min) / (double)s->grid_x_cells;
  s->dy = (s->grid_y_max - s->grid_y_min) / (double)s->grid_y_cells;
}

int read_states(const std::vector<std::string>& lines, Settings* s,
                State** states) {
  int num_states = 0;
  for (const std::string& ln : lines) {
    int id = 0;
    if (get_int("state", ln.c_str(), &id)) num_states = MAX(num_states, id);
  }
  *states = (State*)calloc(num_states > 0 ? num_states : 1, sizeof(State));
  for (const std::string& ln : lines) {
    const char* line = ln.c_str();
    int id = 0;
    if (!get_int("state", line, &id)) continue;
    if (id < 1 || id > num_states) die(__LINE__, __FILE__, "Bad state number %d.\n", id);
    State* st = &(*states)[id - 1];
    if (st->defined) die(__LINE__, __FILE__, "State number %d defined twice.\n", id);
    st->density = atof(value_after(line, "density").c_str());
    st->energy = atof(value_after(line, "energy").c_str());
    if (id > 1) {
      st->x_min = atof(value_after(line, "xmin").c_str()) + s->dx / 100.0;
      st->y_min = atof(value_after(line, "ymin").c_str()) + s->dy / 100.0;
      st->x_max = atof(value_after(line, "xmax").c_str()) - s->dx / 100.0;
      st->y_max = atof(value_after(line, "ymax").c_str()) - s->dy / 100.0;
      const std::string geom = value_after(line, "geometry");
      if (geom == "rectangle") {
        st->geometry = RECTANGULAR;
      } else if (geom == "circular") {
        st->geometry = CIRCULAR;
        st->radius = atof(value_after(line, "radius").c_str());
      } else if (geom == "point") {
        st->geometry = POINT;
      }
    }
    st->defined = true;
  }
  return num_states;
}

}  // namespace

void read_config(Settings* settings, State** states) {
  FILE* fp = fopen(settings->tea_in_filename, "r");
  if (!fp)
    die(__LINE__, __FILE__, "Could not open input file %s\n",
        settings->tea_in_filename);
  const std::vector<std::string> lines = read_lines(fp);
  fclose(fp);

  read_settings(lines, settings);
  settings->num_states = read_states(lines, settings, states);

  initialise_log(settings);
  print_to_log(settings, "Solution Parameters:\n");
  print_to_log(settings, "\tdt_init = %f\n", settings->dt_init);
  print_to_log(settings, "\tend_time = %f\n", settings->end_time);
  print_to_log(settings, "\tend_step = %d\n", settings->end_step);
  print_to_log(settings, "\tgrid_x_min = %f\n", settings->grid_x_min);
  print_to_log(settings, "\tgrid_y_min = %f\n", settings->grid_y_min);
  print_to_log(settings, "\tgrid_x_max = %f\n", settings->grid_x_max);
  print_to_log(settings, "\tgrid_y_max = %f\n", settings->grid_y_max);
  print_to_log(settings, "\tgrid_x_cells = %d\n", settings->grid_x_cells);
  print_to_log(settings, "\tgrid_y_cells = %d\n", settings->grid_y_cells);
  print_to_log(settings, "\tmax_iters = %d\n", settings->max_iters);
  print_to_log(settings, "\teps = %e\n", settings->eps);
  print_to_log(settings, "\thalo_depth = %d\n", settings->halo_depth);
  print_to_log(settings, "\tcheck_result = %d\n", settings->check_result);
  print_to_log(settings, "\tcoefficient = %d\n", settings->coefficient);
  print_to_log(settings, "\tnum_chunks_per_rank = %d\n",
               settings->num_chunks_per_rank);
  print_to_log(settings, "\tsummary_frequency = %d\n",
               settings->summary_frequency);
  for (int ss = 0; ss < settings->num_states; ++ss) {
    const State& st = (*states)[ss];
    print_to_log(settings, "\t\nstate %d\n", ss);
    print_to_log(settings, "\tdensity = %.12E\n", st.density);
    print_to_log(settings, "\tenergy= %.12E\n", st.energy);
    if (ss > 0) {
      print_to_log(settings, "\tx_min = %.12E\n", st.x_min);
      print_to_log(settings, "\ty_min = %.12E\n", st.y_min);
      print_to_log(settings, "\tx_max = %.12E\n", st.x_max);
      print_to_log(settings, "\ty_max = %.12E\n", st.y_max);
      print_to_log(settings, "\tradius = %.12E\n", st.radius);
      print_to_log(settings, "\tgeometry = %d\n", st.geometry);
    }
  }
}
