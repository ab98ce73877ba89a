// kernel_interface.h entry points over the CPU oracle (tea_oracle.c).
// TEST INFRASTRUCTURE ONLY: linked into oracle/build/tealeaf_oracle (this
// project's host) and oracle/_ref/tealeaf_refhost_oracle (the reference's
// own host), never into the product.
//
// Mirrors TeaLeaf/c_kernels/sycl/kernel_interface.cpp: the same kernels in
// the same order for each entry point.  Scalar results are assigned or added
// as that backend does: rro and pw are added (cg.cpp:133,194); rrn, the
// field summary, the Jacobi error and the 2-norm are assigned (cg.cpp:253,
// field_summary.cpp:147-150, jacobi.cpp:116, solver_methods.cpp:116).
#include <stdlib.h>
#include <string.h>

#include "kernel_interface.h"
#include "tea_oracle.h"

#if __has_include("kernel_interface_ext.h")
#include "kernel_interface_ext.h"
#define TEA_HAVE_EXT 1
#endif

// The two-level coarse correction of flexible CG (below).
// Level l coarsens the interior by 8 * 4^l; a level is used while the
// interior divides.  Sizes and buffers of each level.
struct OracleCoarse {
  int levels = 0, sweeps = 0;
  int c[2] = {0, 0}, hc[2] = {0, 0}, wc[2] = {0, 0}, ntiles[2] = {0, 0};
  float* kc[2] = {nullptr, nullptr};
  float* rbar[2] = {nullptr, nullptr};
  float* y[2] = {nullptr, nullptr};
  double* part[2] = {nullptr, nullptr};
  double* cross = nullptr;
};

static void free_coarse(OracleCoarse* q) {
  if (!q) return;
  for (int l = 0; l < 2; ++l) {
    free(q->kc[l]);
    free(q->rbar[l]);
    free(q->y[l]);
    free(q->part[l]);
  }
  free(q->cross);
  delete q;
}

static double* alloc_field(size_t n) {
  double* p = (double*)calloc(n ? n : 1, sizeof(double));
  if (!p) die(__LINE__, __FILE__, "Error allocating buffer\n");
  return p;
}

void run_kernel_initialise(Chunk* c, Settings* s) {
  const size_t x = c->x, y = c->y, xy = x * y;
  c->density0 = alloc_field(xy);
  c->density = alloc_field(xy);
  c->energy0 = alloc_field(xy);
  c->energy = alloc_field(xy);
  c->u = alloc_field(xy);
  c->u0 = alloc_field(xy);
  c->p = alloc_field(xy);
  c->r = alloc_field(xy);
  c->mi = alloc_field(xy);
  c->w = alloc_field(xy);
  c->kx = alloc_field(xy);
  c->ky = alloc_field(xy);
  c->sd = alloc_field(xy);
  c->volume = alloc_field(xy);
  c->x_area = alloc_field((x + 1) * y);
  c->y_area = alloc_field(x * (y + 1));
  c->cell_x = alloc_field(x);
  c->cell_y = alloc_field(y);
  c->cell_dx = alloc_field(x);
  c->cell_dy = alloc_field(y);
  c->vertex_dx = alloc_field(x + 1);
  c->vertex_dy = alloc_field(y + 1);
  c->vertex_x = alloc_field(x + 1);
  c->vertex_y = alloc_field(y + 1);
  c->ext->comms_buffer = alloc_field((x > y ? x : y) * s->halo_depth);
  // The reference host mallocs the extension without clearing it.
  c->ext->mp_p[0] = c->ext->mp_p[1] = nullptr;
  c->ext->mp_r = c->ext->mp_w = c->ext->mp_kx = c->ext->mp_ky = nullptr;
  c->ext->mp_cur = 0;
  c->ext->fcg_z = nullptr;
  c->ext->coarse = nullptr;
  const int n = s->max_iters > 0 ? s->max_iters : 1;
  c->cg_alphas = new double[n]();
  c->cg_betas = new double[n]();
  c->cheby_alphas = new double[n]();
  c->cheby_betas = new double[n]();
}

void run_kernel_finalise(Chunk* c, Settings* s) {
  (void)s;
  double* f[] = {c->density0, c->density, c->energy0, c->energy, c->u,
                 c->u0, c->p, c->r, c->mi, c->w, c->kx, c->ky, c->sd,
                 c->volume, c->x_area, c->y_area, c->cell_x, c->cell_y,
                 c->cell_dx, c->cell_dy, c->vertex_dx, c->vertex_dy,
                 c->vertex_x, c->vertex_y, c->ext->comms_buffer};
  for (double* p : f) free(p);
  float* g[] = {c->ext->mp_p[0], c->ext->mp_p[1], c->ext->mp_r, c->ext->mp_w,
                c->ext->mp_kx, c->ext->mp_ky, c->ext->fcg_z};
  for (float* p : g) free(p);
  free_coarse(c->ext->coarse);
  c->ext->coarse = nullptr;
  delete[] c->cg_alphas;
  delete[] c->cg_betas;
  delete[] c->cheby_alphas;
  delete[] c->cheby_betas;
}

void run_set_chunk_data(Chunk* c, Settings* s) {
  const double x_min = s->grid_x_min + s->dx * (double)c->left;
  const double y_min = s->grid_y_min + s->dy * (double)c->bottom;
  orc_set_chunk_data(c->x, c->y, s->halo_depth, x_min, y_min, s->dx, s->dy,
                     c->vertex_x, c->vertex_y, c->cell_x, c->cell_y, c->volume,
                     c->x_area, c->y_area);
}

void run_set_chunk_state(Chunk* c, Settings* s, State* states) {
  orc_set_chunk_initial_state(c->x, c->y, states[0].energy, states[0].density,
                              c->energy0, c->density);
  for (int ii = 1; ii < s->num_states; ++ii) {
    const State& st = states[ii];
    const int geom = st.geometry == RECTANGULAR ? ORC_RECTANGULAR
                     : st.geometry == CIRCULAR  ? ORC_CIRCULAR
                                                : ORC_POINT;
    orc_set_chunk_state(c->x, c->y, geom, st.x_min, st.y_min, st.x_max,
                        st.y_max, st.radius, st.density, st.energy,
                        c->energy0, c->density, c->u, c->cell_x, c->cell_y,
                        c->vertex_x, c->vertex_y);
  }
}

void run_local_halos(Chunk* c, Settings* s, int depth) {
  // Field order of kernel_interface.cpp:120-125.
  const int order[6] = {FIELD_DENSITY, FIELD_P, FIELD_ENERGY0, FIELD_ENERGY1,
                        FIELD_U, FIELD_SD};
  for (int o = 0; o < 6; ++o) {
    const int f = order[o];
    if (!s->fields_to_exchange[f]) continue;
    double* b = f == FIELD_DENSITY   ? c->density
                : f == FIELD_P       ? c->p
                : f == FIELD_ENERGY0 ? c->energy0
                : f == FIELD_ENERGY1 ? c->energy
                : f == FIELD_U       ? c->u
                                     : c->sd;
    orc_halo_field(c->x, c->y, s->halo_depth, depth, c->neighbours, b);
  }
}

void run_pack_or_unpack(Chunk* c, Settings* s, int depth, int face, bool pack,
                        FieldBufferType field, double* buffer) {
  orc_pack_or_unpack(c->x, c->y, s->halo_depth, depth, face, pack ? 1 : 0,
                     field, buffer);
}

void run_store_energy(Chunk* c, Settings* s) {
  (void)s;
  orc_store_energy(c->x, c->y, c->energy, c->energy0);
}

void run_field_summary(Chunk* c, Settings* s, double* vol, double* mass,
                       double* ie, double* temp) {
  double out[4];
  orc_field_summary(c->x, c->y, s->halo_depth, c->volume, c->density,
                    c->energy0, c->u, out);
  *vol = out[0];
  *mass = out[1];
  *ie = out[2];
  *temp = out[3];
}

void run_cg_init(Chunk* c, Settings* s, double rx, double ry, double* rro) {
  *rro += orc_cg_init(c->x, c->y, s->halo_depth, s->coefficient, rx, ry, c->p,
                      c->r, c->u, c->w, c->kx, c->ky, c->density, c->energy);
}

void run_cg_calc_w(Chunk* c, Settings* s, double* pw) {
  *pw += orc_cg_calc_w(c->x, c->y, s->halo_depth, c->w, c->p, c->kx, c->ky);
}

void run_cg_calc_ur(Chunk* c, Settings* s, double alpha, double* rrn) {
  *rrn = orc_cg_calc_ur(c->x, c->y, s->halo_depth, alpha, c->u, c->r, c->p,
                         c->w);
}

void run_cg_calc_p(Chunk* c, Settings* s, double beta) {
  orc_cg_calc_p(c->x, c->y, s->halo_depth, beta, c->p, c->r);
}

void run_cheby_init(Chunk* c, Settings* s) {
  orc_cheby_init(c->x, c->y, s->halo_depth, c->theta, c->p, c->r, c->u,
                 c->u0, c->w, c->kx, c->ky);
}

void run_cheby_iterate(Chunk* c, Settings* s, double alpha, double beta) {
  orc_cheby_iterate(c->x, c->y, s->halo_depth, alpha, beta, c->p, c->r, c->u,
                    c->u0, c->w, c->kx, c->ky);
}

void run_jacobi_init(Chunk* c, Settings* s, double rx, double ry) {
  orc_jacobi_init(c->x, c->y, s->halo_depth, s->coefficient, rx, ry, c->u,
                  c->u0, c->density, c->energy, c->kx, c->ky);
}

void run_jacobi_iterate(Chunk* c, Settings* s, double* error) {
  *error = orc_jacobi_iterate(c->x, c->y, s->halo_depth, c->u, c->u0, c->r,
                              c->kx, c->ky);
}

void run_ppcg_init(Chunk* c, Settings* s) {
  orc_ppcg_init(c->x, c->y, s->halo_depth, c->theta, c->sd, c->r);
}

void run_ppcg_inner_iteration(Chunk* c, Settings* s, double alpha,
                              double beta) {
  orc_ppcg_inner_iteration(c->x, c->y, s->halo_depth, alpha, beta, c->sd,
                           c->r, c->u, c->kx, c->ky);
}

void run_copy_u(Chunk* c, Settings* s) {
  orc_copy_u(c->x, c->y, s->halo_depth, c->u, c->u0);
}

void run_calculate_residual(Chunk* c, Settings* s) {
  orc_calculate_residual(c->x, c->y, s->halo_depth, c->u, c->u0, c->r, c->kx,
                         c->ky);
}

void run_calculate_2norm(Chunk* c, Settings* s, FieldBufferType buffer,
                         double* norm) {
  *norm = orc_calculate_2norm(c->x, c->y, s->halo_depth, buffer);
}

void run_finalise(Chunk* c, Settings* s) {
  orc_finalise(c->x, c->y, s->halo_depth, c->u, c->density, c->energy);
}

#ifdef TEA_HAVE_EXT
// ---- mixed-precision CG (kernel_interface_ext.h) -----------------------------

void run_mpcg_init(Chunk* c, Settings* s, double rx, double ry, double* rr,
                   double* bb) {
  ChunkExtension* e = c->ext;
  const size_t n = (size_t)c->x * c->y;
  if (!e->mp_r) {
    float** bufs[] = {&e->mp_p[0], &e->mp_p[1], &e->mp_r, &e->mp_w, &e->mp_kx,
                      &e->mp_ky};
    for (float** b : bufs) *b = (float*)calloc(n, sizeof(float));
  }
  const int h = s->halo_depth;
  (void)orc_cg_init(c->x, c->y, h, s->coefficient, rx, ry, c->p, c->r, c->u,
                    c->w, c->kx, c->ky, c->density, c->energy);
  orc_halo_field(c->x, c->y, h, 1, c->neighbours, c->u);
  orc_halo_field(c->x, c->y, h, 1, c->neighbours, c->p);
  orc_copy_u(c->x, c->y, h, c->u, c->u0);
  double out[2];
  orc_mpcg_convert(c->x, c->y, h, c->kx, c->ky, c->p, c->r, c->u0, e->mp_kx,
                   e->mp_ky, e->mp_p[0], e->mp_r, out);
  memcpy(e->mp_p[1], e->mp_p[0], n * sizeof(float));
  e->mp_cur = 0;
  *rr += out[0];
  *bb += out[1];
}

void run_mpcg_calc_w(Chunk* c, Settings* s, double alpha, int apply_u,
                     double beta, int apply_p, double* pw, double* pr) {
  ChunkExtension* e = c->ext;
  *pw += orc_mpcg_calc_w(c->x, c->y, s->halo_depth, alpha, apply_u, beta,
                         apply_p, e->mp_p[e->mp_cur], e->mp_p[e->mp_cur ^ 1],
                         e->mp_r, c->u, e->mp_w, e->mp_kx, e->mp_ky);
  if (apply_p) e->mp_cur ^= 1;
  *pr += orc_dot32(c->x, c->y, s->halo_depth, e->mp_p[e->mp_cur], e->mp_r);
}

void run_mpcg_calc_r(Chunk* c, Settings* s, double alpha, double* rr,
                     double* rw) {
  ChunkExtension* e = c->ext;
  *rr += orc_mpcg_calc_r(c->x, c->y, s->halo_depth, alpha, e->mp_r, e->mp_w);
  *rw += orc_dot32(c->x, c->y, s->halo_depth, e->mp_r, e->mp_w);
}

void run_mpcg_refresh(Chunk* c, Settings* s, double alpha, int apply_u,
                      double* rr, double* rr_true, double* rw) {
  ChunkExtension* e = c->ext;
  double out[3];
  orc_mpcg_refresh(c->x, c->y, s->halo_depth, alpha, apply_u, c->u,
                   e->mp_p[e->mp_cur], c->u0, c->kx, c->ky, e->mp_r, e->mp_w,
                   out);
  *rr += out[0];
  *rr_true += out[1];
  *rw += out[2];
}

// ---- flexible CG: Jacobi, or Jacobi plus the two-level coarse correction
// with Jacobi sweeps as the coarse solve (no network on the CPU) -----------


void run_fcg_attach_sweeps(Chunk* c, Settings* s, int sweeps, int coarsen,
                           int levels) {
  if (sweeps < 1 || coarsen != 8 || levels < 1 || levels > 2)
    die(__LINE__, __FILE__, "coarse sweeps need sweeps >= 1, coarsen 8, 1 or 2 levels\n");
  if (!c->ext->fcg_z)
    c->ext->fcg_z = (float*)calloc((size_t)c->x * c->y, sizeof(float));
  free_coarse(c->ext->coarse);
  OracleCoarse* q = new OracleCoarse();
  q->sweeps = sweeps;
  const int h = s->halo_depth, nx = c->x - 2 * h, ny = c->y - 2 * h;
  for (int l = 0; l < levels; ++l) {
    const int cl = coarsen << (2 * l);
    if (nx % cl != 0 || ny % cl != 0) break;
    q->c[l] = cl;
    q->hc[l] = ny / cl;
    q->wc[l] = nx / cl;
    const size_t cn = (size_t)q->hc[l] * q->wc[l];
    q->ntiles[l] = ((q->wc[l] + 15) / 16) * ((q->hc[l] + 15) / 16);
    q->kc[l] = (float*)calloc(2 * cn, sizeof(float));
    q->rbar[l] = (float*)calloc(cn, sizeof(float));
    q->y[l] = (float*)calloc(cn, sizeof(float));
    q->part[l] = (double*)calloc(2 * (size_t)q->ntiles[l], sizeof(double));
    q->levels = l + 1;
  }
  if (q->levels == 0)
    die(__LINE__, __FILE__, "the interior does not divide into 8 x 8 blocks\n");
  if (q->levels == 2)
    q->cross = (double*)calloc((size_t)q->hc[1] * q->wc[1], sizeof(double));
  c->ext->coarse = q;
}

void run_read_u(Chunk* c, Settings* s, double* host) {
  (void)s;
  memcpy(host, c->u, (size_t)c->x * c->y * sizeof(double));
}

void run_fcg_attach(Chunk* c, Settings* s, const float* params, size_t count,
                    const float* norm, int coarsen, int levels) {
  (void)s; (void)count; (void)norm; (void)coarsen; (void)levels;
  if (params)
    die(__LINE__, __FILE__, "the oracle backend has no network preconditioner\n");
  if (!c->ext->fcg_z)
    c->ext->fcg_z = (float*)calloc((size_t)c->x * c->y, sizeof(float));
  free_coarse(c->ext->coarse);
  c->ext->coarse = nullptr;
}

void run_fcg_init(Chunk* c, Settings* s, double rx, double ry, double* rr,
                  double* bb) {
  run_mpcg_init(c, s, rx, ry, rr, bb);
  OracleCoarse* q = c->ext->coarse;
  if (q)
    for (int l = 0; l < q->levels; ++l)
      orc_fcg_coarse_faces(c->x, c->y, s->halo_depth, q->c[l], c->ext->mp_kx,
                           c->ext->mp_ky, q->kc[l]);
}

void run_fcg_precondition(Chunk* c, Settings* s, int use_net, double* zw) {
  ChunkExtension* e = c->ext;
  OracleCoarse* q = e->coarse;
  const int h = s->halo_depth;
  if (use_net && !q)
    die(__LINE__, __FILE__, "the oracle backend has no network preconditioner\n");
  if (!use_net) {
    *zw += orc_fcg_jacobi(c->x, c->y, h, e->mp_r, e->mp_kx, e->mp_ky, e->mp_w,
                          e->fcg_z);
    return;
  }
  orc_fcg_restrict8(c->x, c->y, h, e->mp_r, q->rbar[0]);
  if (q->levels == 2) orc_fcg_restrict4(q->hc[1], q->wc[1], q->rbar[0], q->rbar[1]);
  for (int l = 0; l < q->levels; ++l)
    orc_fcg_sweeps(q->hc[l], q->wc[l], q->kc[l], q->rbar[l], q->sweeps, q->y[l],
                   q->part[l], l == 0 && q->levels == 2 ? q->cross : nullptr);
  float g[2];
  orc_fcg_gamma(q->levels, q->part[0], q->ntiles[0], q->part[1], q->ntiles[1],
                q->cross, q->y[1], (long)q->hc[1] * q->wc[1],
                (double)q->c[0] * q->c[0], (double)q->c[1] * q->c[1], g);
  *zw += orc_fcg_prolong(c->x, c->y, h, e->mp_r, e->mp_kx, e->mp_ky, e->mp_w,
                         q->levels, q->y[0], q->wc[0], q->c[0], g[0], q->y[1],
                         q->wc[1], q->c[1], g[1], e->fcg_z);
}

void run_fcg_calc_r_precondition(Chunk* c, Settings* s, double alpha,
                                 int use_net, double* rr, double* zw) {
  *rr += orc_mpcg_calc_r(c->x, c->y, s->halo_depth, alpha, c->ext->mp_r,
                         c->ext->mp_w);
  run_fcg_precondition(c, s, use_net, zw);
}

void run_fcg_calc_w(Chunk* c, Settings* s, double alpha, int apply_u,
                    double beta, double* pw, double* pr) {
  ChunkExtension* e = c->ext;
  double out[2];
  orc_fcg_calc_w(c->x, c->y, s->halo_depth, alpha, apply_u, beta,
                 e->mp_p[e->mp_cur], e->mp_p[e->mp_cur ^ 1], e->fcg_z, e->mp_r,
                 c->u, e->mp_w, e->mp_kx, e->mp_ky, out);
  e->mp_cur ^= 1;
  *pw += out[0];
  *pr += out[1];
}
#endif
