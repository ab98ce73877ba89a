// The 24 kernel_interface.h entry points of the HIP backend
// (TeaLeaf/kernel_interface.h:13-71; the SYCL backend's counterpart is
// TeaLeaf/c_kernels/sycl/kernel_interface.cpp).  Each one unpacks
// Chunk/Settings and forwards to the C ABI of libteahip.so (tea_hip.h).
//
// This file compiles against either host: this project's
// (exploringsycl_amd/host) or the reference's own TeaLeaf/*.h, with
// csrc/backend first on the include path for chunk_extension.h.  The symbols
// are C++ (the reference host is compiled as C++).
//
// Error behaviour follows the reference: a failed call ends the run through
// the host's die() with the library's message.  Scalar results are added to
// or assigned to the caller's variable as the SYCL backend does: rro and pw
// are added (cg.cpp:133,194); rrn, the field summary, the Jacobi error and
// the 2-norm are assigned (cg.cpp:253, field_summary.cpp:147-150,
// jacobi.cpp:116, solver_methods.cpp:116).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "kernel_interface.h"
#include "tea_hip.h"

// This project's solver extensions exist only with its own host headers.
#if __has_include("kernel_interface_ext.h")
#include "kernel_interface_ext.h"
#define TEA_HAVE_EXT 1
#endif

namespace {

void check(int status, const char* what, int line) {
  if (status != TEA_HIP_OK)
    die(line, __FILE__, "%s failed: %s\n", what, tea_hip_last_error());
}

#define HIP_CALL(expr) check((expr), #expr, __LINE__)

tea_hip_chunk* H(Chunk* chunk) { return chunk->ext->hip; }

double* field_ptr(tea_hip_chunk* h, int field) {
  double* p = nullptr;
  HIP_CALL(tea_hip_field(h, field, &p, nullptr));
  return p;
}

// p, u, r, w and sd are double-buffered by the fused iterations (CG and
// its run-ahead, Jacobi, Chebyshev, PPCG): republish their addresses after
// every call that can swap them.
void refresh_fields(Chunk* chunk) {
  tea_hip_chunk* h = H(chunk);
  chunk->p = field_ptr(h, TEA_FIELD_P);
  chunk->u = field_ptr(h, TEA_FIELD_U);
  chunk->r = field_ptr(h, TEA_FIELD_R);
  chunk->w = field_ptr(h, TEA_FIELD_W);
  chunk->sd = field_ptr(h, TEA_FIELD_SD);
}

// Live handles, for the timing hook below.
std::vector<tea_hip_chunk*> g_handles;

int device_from_env() {
  const char* s = getenv("TEA_HIP_DEVICE");
  return s ? atoi(s) : 0;
}

}  // namespace

void run_kernel_initialise(Chunk* chunk, Settings* settings) {
  tea_hip_chunk* h = nullptr;
  HIP_CALL(tea_hip_chunk_create(device_from_env(), chunk->x, chunk->y,
                                settings->halo_depth, &h));
  chunk->ext->hip = h;
  g_handles.push_back(h);
  const char* fused = getenv("TEA_HIP_FUSED");
  if (fused && atoi(fused) == 0) HIP_CALL(tea_hip_set_fused(h, 0));

  chunk->density0 = field_ptr(h, TEA_FIELD_DENSITY0);
  chunk->density = field_ptr(h, TEA_FIELD_DENSITY);
  chunk->energy0 = field_ptr(h, TEA_FIELD_ENERGY0);
  chunk->energy = field_ptr(h, TEA_FIELD_ENERGY);
  chunk->u = field_ptr(h, TEA_FIELD_U);
  chunk->u0 = field_ptr(h, TEA_FIELD_U0);
  chunk->p = field_ptr(h, TEA_FIELD_P);
  chunk->r = field_ptr(h, TEA_FIELD_R);
  chunk->mi = field_ptr(h, TEA_FIELD_MI);
  chunk->w = field_ptr(h, TEA_FIELD_W);
  chunk->kx = field_ptr(h, TEA_FIELD_KX);
  chunk->ky = field_ptr(h, TEA_FIELD_KY);
  chunk->sd = field_ptr(h, TEA_FIELD_SD);
  chunk->cell_x = field_ptr(h, TEA_FIELD_CELL_X);
  chunk->cell_y = field_ptr(h, TEA_FIELD_CELL_Y);
  chunk->cell_dx = field_ptr(h, TEA_FIELD_CELL_DX);
  chunk->cell_dy = field_ptr(h, TEA_FIELD_CELL_DY);
  chunk->vertex_dx = field_ptr(h, TEA_FIELD_VERTEX_DX);
  chunk->vertex_dy = field_ptr(h, TEA_FIELD_VERTEX_DY);
  chunk->vertex_x = field_ptr(h, TEA_FIELD_VERTEX_X);
  chunk->vertex_y = field_ptr(h, TEA_FIELD_VERTEX_Y);
  chunk->volume = field_ptr(h, TEA_FIELD_VOLUME);
  chunk->x_area = field_ptr(h, TEA_FIELD_X_AREA);
  chunk->y_area = field_ptr(h, TEA_FIELD_Y_AREA);
  chunk->ext->comms_buffer = field_ptr(h, TEA_FIELD_COMMS_BUFFER);

  // Host-side coefficient logs of length max_iters
  // (TeaLeaf/c_kernels/sycl/kernel_initialise.cpp:76-79).
  const int n = settings->max_iters > 0 ? settings->max_iters : 1;
  chunk->cg_alphas = new double[n]();
  chunk->cg_betas = new double[n]();
  chunk->cheby_alphas = new double[n]();
  chunk->cheby_betas = new double[n]();
}

void run_kernel_finalise(Chunk* chunk, Settings* settings) {
  (void)settings;
  for (size_t i = 0; i < g_handles.size(); ++i)
    if (g_handles[i] == H(chunk)) g_handles.erase(g_handles.begin() + i);
  HIP_CALL(tea_hip_chunk_destroy(H(chunk)));
  chunk->ext->hip = nullptr;
  delete[] chunk->cg_alphas;
  delete[] chunk->cg_betas;
  delete[] chunk->cheby_alphas;
  delete[] chunk->cheby_betas;
  chunk->cg_alphas = chunk->cg_betas = nullptr;
  chunk->cheby_alphas = chunk->cheby_betas = nullptr;
}

void run_set_chunk_data(Chunk* chunk, Settings* settings) {
  const double x_min = settings->grid_x_min + settings->dx * (double)chunk->left;
  const double y_min =
      settings->grid_y_min + settings->dy * (double)chunk->bottom;
  HIP_CALL(tea_hip_set_chunk_data(H(chunk), x_min, y_min, settings->dx,
                                  settings->dy));
}

void run_set_chunk_state(Chunk* chunk, Settings* settings, State* states) {
  const int n = settings->num_states;
  if (n < 1) die(__LINE__, __FILE__, "no states defined\n");
  double* buf = new double[7 * (size_t)n];
  int* geometry = new int[n];
  double *density = buf, *energy = buf + n, *x_min = buf + 2 * n,
         *y_min = buf + 3 * n, *x_max = buf + 4 * n, *y_max = buf + 5 * n,
         *radius = buf + 6 * n;
  for (int s = 0; s < n; ++s) {
    density[s] = states[s].density;
    energy[s] = states[s].energy;
    x_min[s] = states[s].x_min;
    y_min[s] = states[s].y_min;
    x_max[s] = states[s].x_max;
    y_max[s] = states[s].y_max;
    radius[s] = states[s].radius;
    geometry[s] = states[s].geometry == RECTANGULAR ? TEA_GEOMETRY_RECTANGULAR
                  : states[s].geometry == CIRCULAR  ? TEA_GEOMETRY_CIRCULAR
                                                    : TEA_GEOMETRY_POINT;
  }
  HIP_CALL(tea_hip_set_chunk_state(H(chunk), n, density, energy, x_min, y_min,
                                   x_max, y_max, radius, geometry));
  delete[] buf;
  delete[] geometry;
}

void run_local_halos(Chunk* chunk, Settings* settings, int depth) {
  int fields[TEA_HALO_NUM];
  for (int s = 0; s < TEA_HALO_NUM; ++s)
    fields[s] = settings->fields_to_exchange[s] ? 1 : 0;
  HIP_CALL(tea_hip_local_halos(H(chunk), depth, chunk->neighbours, fields));
  refresh_fields(chunk);
}

void run_pack_or_unpack(Chunk* chunk, Settings* settings, int depth, int face,
                        bool pack, FieldBufferType field, double* buffer) {
  (void)settings;
  HIP_CALL(tea_hip_pack_or_unpack(H(chunk), depth, face, pack ? 1 : 0, field,
                                  buffer));
  refresh_fields(chunk);
}

void run_store_energy(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_store_energy(H(chunk)));
}

void run_field_summary(Chunk* chunk, Settings* settings, double* vol,
                       double* mass, double* ie, double* temp) {
  (void)settings;
  HIP_CALL(tea_hip_field_summary(H(chunk), vol, mass, ie, temp));
  refresh_fields(chunk);
}

void run_cg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                 double* rro) {
  double v;
  HIP_CALL(tea_hip_cg_init(H(chunk), settings->coefficient, rx, ry, &v));
  refresh_fields(chunk);
  *rro += v;
}

void run_cg_calc_w(Chunk* chunk, Settings* settings, double* pw) {
  (void)settings;
  double v;
  HIP_CALL(tea_hip_cg_calc_w(H(chunk), &v));
  refresh_fields(chunk);
  *pw += v;
}

void run_cg_calc_ur(Chunk* chunk, Settings* settings, double alpha,
                    double* rrn) {
  (void)settings;
  HIP_CALL(tea_hip_cg_calc_ur(H(chunk), alpha, rrn));
  refresh_fields(chunk);
}

void run_cg_calc_p(Chunk* chunk, Settings* settings, double beta) {
  (void)settings;
  HIP_CALL(tea_hip_cg_calc_p(H(chunk), beta));
  refresh_fields(chunk);
}

void run_cheby_init(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_cheby_init(H(chunk), chunk->theta));
  refresh_fields(chunk);
}

void run_cheby_iterate(Chunk* chunk, Settings* settings, double alpha,
                       double beta) {
  (void)settings;
  HIP_CALL(tea_hip_cheby_iterate(H(chunk), alpha, beta));
  refresh_fields(chunk);
}

void run_jacobi_init(Chunk* chunk, Settings* settings, double rx, double ry) {
  HIP_CALL(tea_hip_jacobi_init(H(chunk), settings->coefficient, rx, ry));
  refresh_fields(chunk);
}

void run_jacobi_iterate(Chunk* chunk, Settings* settings, double* error) {
  (void)settings;
  HIP_CALL(tea_hip_jacobi_iterate(H(chunk), error));
  refresh_fields(chunk);
}

void run_ppcg_init(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_ppcg_init(H(chunk), chunk->theta));
  refresh_fields(chunk);
}

void run_ppcg_inner_iteration(Chunk* chunk, Settings* settings, double alpha,
                              double beta) {
  (void)settings;
  HIP_CALL(tea_hip_ppcg_inner_iteration(H(chunk), alpha, beta));
  refresh_fields(chunk);
}

void run_copy_u(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_copy_u(H(chunk)));
  refresh_fields(chunk);
}

void run_calculate_residual(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_calculate_residual(H(chunk)));
  refresh_fields(chunk);
}

void run_calculate_2norm(Chunk* chunk, Settings* settings,
                         FieldBufferType buffer, double* norm) {
  (void)settings;
  HIP_CALL(tea_hip_calculate_2norm(H(chunk), buffer, norm));
  refresh_fields(chunk);
}

void run_finalise(Chunk* chunk, Settings* settings) {
  (void)settings;
  HIP_CALL(tea_hip_finalise(H(chunk)));
  refresh_fields(chunk);
}

#ifdef TEA_HAVE_EXT
// ---- mixed-precision CG (kernel_interface_ext.h) ---------------------------

void run_mpcg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                   double* rr, double* bb) {
  for (int f = 0; f < NUM_FACES; ++f)
    if (chunk->neighbours[f] != EXTERNAL_FACE)
      die(__LINE__, __FILE__, "mixed-precision CG needs all faces external\n");
  double a, b;
  HIP_CALL(tea_hip_mpcg_init(H(chunk), settings->coefficient, rx, ry, &a, &b));
  refresh_fields(chunk);
  *rr += a;
  *bb += b;
}

void run_mpcg_calc_w(Chunk* chunk, Settings* settings, double alpha,
                     int apply_u, double beta, int apply_p, double* pw,
                     double* pr) {
  (void)settings;
  double a, b;
  HIP_CALL(tea_hip_mpcg_calc_w(H(chunk), alpha, apply_u, beta, apply_p, &a, &b));
  *pw += a;
  *pr += b;
}

void run_mpcg_calc_r(Chunk* chunk, Settings* settings, double alpha,
                     double* rr, double* rw) {
  (void)settings;
  double a, b;
  HIP_CALL(tea_hip_mpcg_calc_r(H(chunk), alpha, &a, &b));
  *rr += a;
  *rw += b;
}

void run_mpcg_refresh(Chunk* chunk, Settings* settings, double alpha,
                      int apply_u, double* rr, double* rr_true, double* rw) {
  (void)settings;
  double a, b, w;
  HIP_CALL(tea_hip_mpcg_refresh(H(chunk), alpha, apply_u, &a, &b, &w));
  *rr += a;
  *rr_true += b;
  *rw += w;
}

void run_read_u(Chunk* chunk, Settings* settings, double* host) {
  (void)settings;
  HIP_CALL(tea_hip_field_read(H(chunk), TEA_FIELD_U, host,
                              (size_t)chunk->x * chunk->y));
}

void run_fcg_attach(Chunk* chunk, Settings* settings, const float* params,
                    size_t count, const float* norm, int coarsen, int levels) {
  (void)settings;
  HIP_CALL(tea_hip_fcg_attach(H(chunk), params, count, norm, coarsen, levels));
}

void run_fcg_attach_sweeps(Chunk* chunk, Settings* settings, int sweeps, int coarsen,
                           int levels) {
  (void)settings;
  HIP_CALL(tea_hip_fcg_attach_sweeps(H(chunk), sweeps, coarsen, levels));
}

void run_fcg_init(Chunk* chunk, Settings* settings, double rx, double ry,
                  double* rr, double* bb) {
  for (int f = 0; f < NUM_FACES; ++f)
    if (chunk->neighbours[f] != EXTERNAL_FACE)
      die(__LINE__, __FILE__, "flexible CG needs every face external\n");
  double a = 0.0, b = 0.0;
  HIP_CALL(tea_hip_fcg_init(H(chunk), settings->coefficient, rx, ry, &a, &b));
  refresh_fields(chunk);
  *rr += a;
  *bb += b;
}

void run_fcg_precondition(Chunk* chunk, Settings* settings, int use_net,
                          double* zw) {
  (void)settings;
  double v;
  HIP_CALL(tea_hip_fcg_precondition(H(chunk), use_net, &v));
  *zw += v;
}

void run_fcg_calc_r_precondition(Chunk* chunk, Settings* settings,
                                 double alpha, int use_net, double* rr,
                                 double* zw) {
  (void)settings;
  double a, b;
  HIP_CALL(tea_hip_fcg_calc_r_precondition(H(chunk), alpha, use_net, &a, &b));
  *rr += a;
  *zw += b;
}

void run_fcg_calc_w(Chunk* chunk, Settings* settings, double alpha,
                    int apply_u, double beta, double* pw, double* pr) {
  (void)settings;
  double a, b;
  HIP_CALL(tea_hip_fcg_calc_w(H(chunk), alpha, apply_u, beta, &a, &b));
  *pw += a;
  *pr += b;
}
#endif

// Optional host hook (declared weak by the host): with TEA_HIP_STATS=<file>
// set, "timed_begin" switches on per-kernel event timing and clears the
// counters, and "timed_end" waits for the device and writes, per kernel, the
// launch count and summed device milliseconds of the timed region as JSON.
// "timed_end" always waits for the device so the host's clock covers all
// queued work.
extern "C" void tea_backend_mark(const char* event) {
  const char* path = getenv("TEA_HIP_STATS");
  if (strcmp(event, "timed_begin") == 0) {
    for (tea_hip_chunk* h : g_handles) {
      HIP_CALL(tea_hip_synchronize(h));
      if (path) {
        HIP_CALL(tea_hip_timing_enable(h, 1));
        HIP_CALL(tea_hip_timing_reset(h));
      }
    }
    return;
  }
  if (strcmp(event, "timed_end") != 0) return;
  for (tea_hip_chunk* h : g_handles) HIP_CALL(tea_hip_synchronize(h));
  if (!path) return;
  FILE* fp = fopen(path, "w");
  if (!fp) die(__LINE__, __FILE__, "cannot open %s\n", path);
  fprintf(fp, "{");
  const int nk = tea_hip_num_kernels();
  bool first = true;
  for (int k = 0; k < nk; ++k) {
    long launches = 0;
    double ms = 0.0;
    for (tea_hip_chunk* h : g_handles) {
      long l = 0;
      double t = 0.0;
      HIP_CALL(tea_hip_kernel_stats(h, k, &l, &t));
      launches += l;
      ms += t;
    }
    if (!launches) continue;
    fprintf(fp, "%s\n  \"%s\": {\"launches\": %ld, \"total_ms\": %.6f}",
            first ? "" : ",", tea_hip_kernel_name(k), launches, ms);
    first = false;
  }
  fprintf(fp, "\n}\n");
  fclose(fp);
}
