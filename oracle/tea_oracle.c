/*
 * CPU oracle for the TeaLeaf CG path: restatement of the reference SYCL
 * backend's FP64 arithmetic.  TEST INFRASTRUCTURE ONLY (see tea_oracle.h).
 *
 * Reference files (relative to the reference repository):
 *   TeaLeaf/c_kernels/sycl/cg.cpp              cg_init_u/_k/_others, calc_w/ur/p
 *   TeaLeaf/c_kernels/sycl/solver_methods.cpp  copy_u, residual, 2norm, finalise
 *   TeaLeaf/c_kernels/sycl/jacobi.cpp, cheby.cpp, ppcg.cpp  the other solvers
 *   TeaLeaf/c_kernels/sycl/local_halos.cpp     update_left/right/top/bottom
 *   TeaLeaf/c_kernels/sycl/field_summary.cpp   field_summary_func (+_reduction)
 *   TeaLeaf/c_kernels/sycl/set_chunk_data.cpp  mesh
 *   TeaLeaf/c_kernels/sycl/set_chunk_state.cpp states
 *   TeaLeaf/c_kernels/sycl/store_energy.cpp    energy := energy0
 *   TeaLeaf/c_kernels/sycl/sycl_shared.hpp:35-79  reduceArray tree
 *   TeaLeaf/shared.h:59-63                     SMVP macro
 *
 * How each function is pinned (checked by tests/test_oracle_*.py):
 *   - whole path: the known answers of TeaLeaf/tea.problems (temperature at
 *     the last step, 1e-5 relative as the reference host checks, and ~1e-10
 *     against the values published to 15 digits), and the Fortran field
 *     summaries of TeaLeaf/Benchmarks/tea_bm_1.out;
 *   - the reference's own host driver (TeaLeaf/ and TeaLeaf/drivers/ sources) compiled
 *     against this oracle reproduces the known answers (oracle/Makefile, _ref);
 *   - orc_smvp / orc_cg_init: symmetry x'Ay = y'Ax, row sums of one, and a
 *     dense numpy solve on a 10x10 grid against the CG result;
 *   - halos: exact mirror identities;
 *   - field summary: closed-form volume/mass/ie/temp for grids on which the
 *     tea.in rectangles fall on cell edges;
 *   - reduction tree: inputs whose pairwise-tree sum differs from the
 *     sequential sum by a value worked out by hand, and exact integer sums;
 *   - orc_pack_or_unpack: each face's strip equals the 2-D slice of the
 *     columns (rows) next to the interior, or of the halo, in row-major
 *     order;
 *   - orc_jacobi_*: jacobi_init's coefficients equal cg_init's, the dense
 *     solution is a fixed point of a sweep, sweeps contract in the max norm,
 *     r is the previous iterate and the error the L1 norm of the update;
 *   - orc_cheby_*: with exact eigenvalue bounds the iteration meets the
 *     Chebyshev error bound |e_k| <= |e_0| / T_k(sigma) at every step, and
 *     w, r are A u and b - A u of the iterate before the step;
 *   - orc_ppcg_*: the inner steps keep r = r0 - A z and meet the same bound
 *     on the residual;
 *   - all three solvers: the known answers of TeaLeaf/tea.problems, and the
 *     reference's own host driving this oracle prints the same iteration
 *     counts, eigenvalue bounds and temperatures as this project's host.
 * No function is left unpinned.
 */
#include "tea_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GROUP 64

static inline int is_interior(long kk, long jj, int x, int y, int h) {
  return kk >= h && kk < x - h && jj >= h && jj < y - h;
}

/* Pairwise tree of sycl_shared.hpp:59-64 / cg.cpp:177-183. */
double orc_tree64(double* a) {
  for (int s = 1; s < GROUP; s *= 2)
    for (int i = 0; i + s < GROUP; i += 2 * s) a[i] = a[i] + a[i + s];
  return a[0];
}

/* reduceArray (sycl_shared.hpp:35-79): repeated 64-wide trees over the
 * partials, zero padded, until one value is left.  `v` is consumed. */
static double reduce_partials(double* v, size_t len) {
  while (len != 1) {
    size_t groups = (len + GROUP - 1) / GROUP;
    for (size_t g = 0; g < groups; ++g) {
      double a[GROUP];
      for (int l = 0; l < GROUP; ++l) {
        size_t i = g * GROUP + l;
        a[l] = i < len ? v[i] : 0.0;
      }
      v[g] = orc_tree64(a); /* g <= g*64: groups are done in order */
    }
    len = groups;
  }
  return v[0];
}

double orc_tree_sum(const double* v, size_t n) {
  if (n == 0) return 0.0;
  size_t groups = (n + GROUP - 1) / GROUP;
  double* part = (double*)malloc(groups * sizeof(double));
  for (size_t g = 0; g < groups; ++g) {
    double a[GROUP];
    for (int l = 0; l < GROUP; ++l) {
      size_t i = g * GROUP + l;
      a[l] = i < n ? v[i] : 0.0;
    }
    part[g] = orc_tree64(a);
  }
  double r = reduce_partials(part, groups);
  free(part);
  return r;
}

/* shared.h:59-63, evaluated in the macro's order with every product and sum
 * rounded separately. */
double orc_smvp(const double* a, const double* kx, const double* ky, long i,
                long x) {
  const double t1 = kx[i + 1] + kx[i];
  const double t2 = ky[i + x] + ky[i];
  const double diag = (1.0 + t1) + t2;
  const double c = diag * a[i];
  const double bx = kx[i + 1] * a[i + 1] + kx[i] * a[i - 1];
  const double by = ky[i + x] * a[i + x] + ky[i] * a[i - x];
  return (c - bx) - by;
}

/* Walk the padded linear range group by group: BODY sets `val` for interior
 * cell (kk, jj) at linear index i; the group's tree goes to part[g]. */
#define GROUP_REDUCE(x, y, h, part, BODY)                                     \
  do {                                                                        \
    const long n_ = (long)(x) * (long)(y);                                    \
    const long groups_ = (n_ + GROUP - 1) / GROUP;                            \
    _Pragma("omp parallel for schedule(static)")                              \
    for (long g = 0; g < groups_; ++g) {                                      \
      double a_[GROUP];                                                       \
      long i0_ = g * GROUP;                                                   \
      long jj = i0_ / (x);                                                    \
      long kk = i0_ - jj * (x);                                               \
      for (int l = 0; l < GROUP; ++l) {                                       \
        const long i = i0_ + l;                                               \
        double val = 0.0;                                                     \
        if (i < n_ && is_interior(kk, jj, (x), (y), (h))) { BODY }            \
        a_[l] = val;                                                          \
        if (++kk == (x)) { kk = 0; ++jj; }                                    \
      }                                                                       \
      (part)[g] = orc_tree64(a_);                                             \
    }                                                                         \
  } while (0)

static double* alloc_partials(int x, int y) {
  long n = (long)x * y;
  return (double*)malloc(((n + GROUP - 1) / GROUP) * sizeof(double));
}

static double finish(double* part, int x, int y) {
  long n = (long)x * y;
  double r = reduce_partials(part, (size_t)((n + GROUP - 1) / GROUP));
  free(part);
  return r;
}

/* set_chunk_data.cpp:8-64 */
void orc_set_chunk_data(int x, int y, int h, double x_min, double y_min,
                        double dx, double dy, double* vertex_x,
                        double* vertex_y, double* cell_x, double* cell_y,
                        double* volume, double* x_area, double* y_area) {
  for (long i = 0; i < x + 1; ++i)
    vertex_x[i] = x_min + dx * ((double)i - (double)h);
  for (long i = 0; i < y + 1; ++i)
    vertex_y[i] = y_min + dy * ((double)i - (double)h);
  for (long i = 0; i < x; ++i) cell_x[i] = 0.5 * (vertex_x[i] + vertex_x[i + 1]);
  for (long i = 0; i < y; ++i) cell_y[i] = 0.5 * (vertex_y[i] + vertex_y[i + 1]);
  const long n = (long)x * y;
  const double vol = dx * dy;
  for (long i = 0; i < n; ++i) {
    volume[i] = vol;
    x_area[i] = dy;
    y_area[i] = dx;
  }
}

/* set_chunk_state.cpp:8-27 */
void orc_set_chunk_initial_state(int x, int y, double energy, double density,
                                 double* energy0, double* density_f) {
  const long n = (long)x * y;
  for (long i = 0; i < n; ++i) {
    energy0[i] = energy;
    density_f[i] = density;
  }
}

/* set_chunk_state.cpp:30-92 */
void orc_set_chunk_state(int x, int y, int geometry, double sx_min,
                         double sy_min, double sx_max, double sy_max,
                         double radius, double s_density, double s_energy,
                         double* energy0, double* density, double* u,
                         const double* cell_x, const double* cell_y,
                         const double* vertex_x, const double* vertex_y) {
  for (long jj = 0; jj < y; ++jj) {
    for (long kk = 0; kk < x; ++kk) {
      const long i = kk + jj * x;
      int apply = 0;
      if (geometry == ORC_RECTANGULAR) {
        apply = vertex_x[kk + 1] >= sx_min && vertex_x[kk] < sx_max &&
                vertex_y[jj + 1] >= sy_min && vertex_y[jj] < sy_max;
      } else if (geometry == ORC_CIRCULAR) {
        const double ddx = cell_x[kk] - sx_min;
        const double ddy = cell_y[jj] - sy_min;
        apply = sqrt(ddx * ddx + ddy * ddy) <= radius;
      } else if (geometry == ORC_POINT) {
        apply = vertex_x[kk] == sx_min && vertex_y[jj] == sy_min;
      }
      if (apply) {
        energy0[i] = s_energy;
        density[i] = s_density;
      }
      if (kk > 0 && kk < x - 1 && jj > 0 && jj < y - 1)
        u[i] = energy0[i] * density[i];
    }
  }
}

/* store_energy.cpp:6-24 */
void orc_store_energy(int x, int y, double* energy, const double* energy0) {
  memcpy(energy, energy0, sizeof(double) * (size_t)x * (size_t)y);
}

/* local_halos.cpp:7-102, applied in kernel_interface.cpp:100-127's face
 * order.  Written with (row, column) indices rather than the reference's
 * flattened ones; the copies are the same. */
void orc_halo_field(int x, int y, int h, int depth, const int neighbours[4],
                    double* b) {
  if (neighbours[ORC_LEFT] < 0)
    for (long jj = 0; jj < y; ++jj)
      for (int k = 0; k < depth; ++k)
        b[jj * x + (h - 1 - k)] = b[jj * x + (h + k)];
  if (neighbours[ORC_RIGHT] < 0)
    for (long jj = 0; jj < y; ++jj)
      for (int k = 0; k < depth; ++k)
        b[jj * x + (x - h + k)] = b[jj * x + (x - h - 1 - k)];
  if (neighbours[ORC_TOP] < 0)
    for (int k = 0; k < depth; ++k)
      for (long kk = 0; kk < x; ++kk)
        b[(long)(y - h + k) * x + kk] = b[(long)(y - h - 1 - k) * x + kk];
  if (neighbours[ORC_BOTTOM] < 0)
    for (int k = 0; k < depth; ++k)
      for (long kk = 0; kk < x; ++kk)
        b[(long)(h - 1 - k) * x + kk] = b[(long)(h + k) * x + kk];
}

/* field_summary.cpp:6-151: four trees over the same index range. */
void orc_field_summary(int x, int y, int h, const double* volume,
                       const double* density, const double* energy0,
                       const double* u, double out[4]) {
  const long n = (long)x * y;
  const long groups = (n + GROUP - 1) / GROUP;
  double* part[4];
  for (int q = 0; q < 4; ++q) part[q] = (double*)malloc(groups * sizeof(double));
#pragma omp parallel for schedule(static)
  for (long g = 0; g < groups; ++g) {
    double a[4][GROUP];
    long i0 = g * GROUP;
    long jj = i0 / x, kk = i0 - jj * x;
    for (int l = 0; l < GROUP; ++l) {
      const long i = i0 + l;
      a[0][l] = a[1][l] = a[2][l] = a[3][l] = 0.0;
      if (i < n && is_interior(kk, jj, x, y, h)) {
        const double cell_vol = volume[i];
        const double cell_mass = cell_vol * density[i];
        a[0][l] = cell_vol;
        a[1][l] = cell_mass;
        a[2][l] = cell_mass * energy0[i];
        a[3][l] = cell_mass * u[i];
      }
      if (++kk == x) { kk = 0; ++jj; }
    }
    for (int q = 0; q < 4; ++q) part[q][g] = orc_tree64(a[q]);
  }
  for (int q = 0; q < 4; ++q) {
    out[q] = reduce_partials(part[q], (size_t)groups);
    free(part[q]);
  }
}

/* cg.cpp:7-134 (cg_init_u, cg_init_k, cg_init_others). */
double orc_cg_init(int x, int y, int h, int coefficient, double rx, double ry,
                   double* p, double* r, double* u, double* w, double* kx,
                   double* ky, const double* density, const double* energy) {
  const long n = (long)x * y;
  /* cg_init_u (cg.cpp:7-37) */
#pragma omp parallel for schedule(static)
  for (long jj = 0; jj < y; ++jj) {
    for (long kk = 0; kk < x; ++kk) {
      const long i = kk + jj * x;
      p[i] = 0.0;
      r[i] = 0.0;
      u[i] = energy[i] * density[i];
      if (jj > 0 && jj < y - 1 && kk > 0 && kk < x - 1)
        w[i] = coefficient == ORC_CONDUCTIVITY ? density[i] : 1.0 / density[i];
    }
  }
  /* cg_init_k (cg.cpp:40-69) */
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - 1; ++jj) {
    for (long kk = h; kk < x - 1; ++kk) {
      const long i = kk + jj * x;
      kx[i] = rx * (w[i - 1] + w[i]) / (2.0 * w[i - 1] * w[i]);
      ky[i] = ry * (w[i - x] + w[i]) / (2.0 * w[i - x] * w[i]);
    }
  }
  (void)n;
  /* cg_init_others (cg.cpp:72-134) */
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const double s = orc_smvp(u, kx, ky, i, x);
    w[i] = s;
    r[i] = u[i] - w[i];
    p[i] = r[i];
    val = r[i] * p[i];
  });
  return finish(part, x, y);
}

/* cg.cpp:137-195 */
double orc_cg_calc_w(int x, int y, int h, double* w, const double* p,
                     const double* kx, const double* ky) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const double s = orc_smvp(p, kx, ky, i, x);
    w[i] = s;
    val = w[i] * p[i];
  });
  return finish(part, x, y);
}

/* cg.cpp:198-254 */
double orc_cg_calc_ur(int x, int y, int h, double alpha, double* u, double* r,
                      const double* p, const double* w) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    u[i] = u[i] + alpha * p[i];
    r[i] = r[i] - alpha * w[i];
    val = r[i] * r[i];
  });
  return finish(part, x, y);
}

/* cg.cpp:257-281 */
void orc_cg_calc_p(int x, int y, int h, double beta, double* p,
                   const double* r) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      p[i] = beta * p[i] + r[i];
    }
}

/* solver_methods.cpp:7-31 */
void orc_copy_u(int x, int y, int h, const double* u, double* u0) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) u0[kk + jj * x] = u[kk + jj * x];
}

/* solver_methods.cpp:34-65 */
void orc_calculate_residual(int x, int y, int h, const double* u,
                            const double* u0, double* r, const double* kx,
                            const double* ky) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      r[i] = u0[i] - orc_smvp(u, kx, ky, i, x);
    }
}

/* solver_methods.cpp:68-117 */
double orc_calculate_2norm(int x, int y, int h, const double* b) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, { val = b[i] * b[i]; });
  return finish(part, x, y);
}

/* solver_methods.cpp:120-145 */
void orc_finalise(int x, int y, int h, const double* u, const double* density,
                  double* energy) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      energy[i] = u[i] / density[i];
    }
}

/* run_pack_or_unpack (kernel_interface.cpp:129-170, pack_halos.cpp:7-184):
 * the face strip of `depth` cells next to the interior (pack) or in the halo
 * (unpack), copied to or from `buffer` in the reference's order.  Faces as
 * the host's neighbours: 0 left, 1 right, 2 bottom, 3 top. */
void orc_pack_or_unpack(int x, int y, int h, int depth, int face, int pack,
                        double* field, double* buffer) {
  const long n = (face == 0 || face == 1) ? (long)y * depth : (long)x * depth;
  for (long i = 0; i < n; ++i) {
    long f;
    const long lines = i / depth;
    if (face == 0)      /* pack_left / unpack_left */
      f = pack ? h + lines * (x - depth) + i : h - depth + lines * (x - depth) + i;
    else if (face == 1) /* pack_right / unpack_right */
      f = pack ? x - h - depth + lines * (x - depth) + i
               : x - h + lines * (x - depth) + i;
    else if (face == 2) /* pack_bottom / unpack_bottom */
      f = pack ? (long)x * h + i : (long)x * (h - depth) + i;
    else                /* pack_top / unpack_top */
      f = pack ? (long)x * (y - h - depth) + i : (long)x * (y - h) + i;
    if (pack)
      buffer[i] = field[f];
    else
      field[f] = buffer[i];
  }
}

/* ---- Jacobi, Chebyshev and PPCG -------------------------------------------
 * Jacobi (jacobi.cpp), Chebyshev (cheby.cpp) and PPCG inner steps (ppcg.cpp),
 * composed as the reference's kernel_interface.cpp:246-328 composes them. */

/* jacobi_init (jacobi.cpp:7-54): u0 = u = energy*density one cell in from
 * the edge of the padded grid; kx, ky from the density as in cg_init_k. */
void orc_jacobi_init(int x, int y, int h, int coefficient, double rx,
                     double ry, double* u, double* u0, const double* density,
                     const double* energy, double* kx, double* ky) {
#pragma omp parallel for schedule(static)
  for (long jj = 0; jj < y; ++jj) {
    for (long kk = 0; kk < x; ++kk) {
      const long i = kk + jj * x;
      if (kk > 0 && kk < x - 1 && jj > 0 && jj < y - 1) {
        u0[i] = energy[i] * density[i];
        u[i] = u0[i];
      }
      if (jj >= h && jj < y - 1 && kk >= h && kk < x - 1) {
        const int cond = coefficient == ORC_CONDUCTIVITY;
        const double dc = cond ? density[i] : 1.0 / density[i];
        const double dl = cond ? density[i - 1] : 1.0 / density[i - 1];
        const double dd = cond ? density[i - x] : 1.0 / density[i - x];
        kx[i] = rx * (dl + dc) / (2.0 * dl * dc);
        ky[i] = ry * (dd + dc) / (2.0 * dd * dc);
      }
    }
  }
}

/* run_jacobi_iterate (kernel_interface.cpp:287-300): jacobi_copy_u
 * (jacobi.cpp:120-137, r = u on every cell) then jacobi_iterate
 * (jacobi.cpp:57-117); returns the tree sum of |u_new - r|. */
double orc_jacobi_iterate(int x, int y, int h, double* u, const double* u0,
                          double* r, const double* kx, const double* ky) {
  const long n = (long)x * y;
  memcpy(r, u, (size_t)n * sizeof(double));
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    u[i] = (u0[i] + (kx[i + 1] * r[i + 1] + kx[i] * r[i - 1]) +
            (ky[i + x] * r[i + x] + ky[i] * r[i - x])) /
           (1.0 + (kx[i] + kx[i + 1]) + (ky[i] + ky[i + x]));
    val = fabs(u[i] - r[i]);
  });
  return finish(part, x, y);
}

/* cheby_init (cheby.cpp:7-43): w = A u, r = u0 - w, p = r / theta. */
void orc_cheby_init(int x, int y, int h, double theta, double* p, double* r,
                    const double* u, const double* u0, double* w,
                    const double* kx, const double* ky) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      w[i] = orc_smvp(u, kx, ky, i, x);
      r[i] = u0[i] - w[i];
      p[i] = r[i] / theta;
    }
}

/* run_cheby_iterate (kernel_interface.cpp:258-271): cheby_iterate
 * (cheby.cpp:73-110: w = A u, r = u0 - w, p = alpha p + beta r) over the
 * whole interior, then cheby_calc_u (cheby.cpp:46-70: u += p). */
void orc_cheby_iterate(int x, int y, int h, double alpha, double beta,
                       double* p, double* r, double* u, const double* u0,
                       double* w, const double* kx, const double* ky) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      w[i] = orc_smvp(u, kx, ky, i, x);
      r[i] = u0[i] - w[i];
      p[i] = alpha * p[i] + beta * r[i];
    }
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      u[i] = u[i] + p[i];
    }
}

/* ppcg_init (ppcg.cpp:7-31): sd = r / theta. */
void orc_ppcg_init(int x, int y, int h, double theta, double* sd,
                   const double* r) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      sd[i] = r[i] / theta;
    }
}

/* run_ppcg_inner_iteration (kernel_interface.cpp:314-328): ppcg_calc_ur
 * (ppcg.cpp:34-66: r -= A sd, u += sd) over the whole interior, then
 * ppcg_calc_sd (ppcg.cpp:69-94: sd = alpha sd + beta r). */
void orc_ppcg_inner_iteration(int x, int y, int h, double alpha, double beta,
                              double* sd, double* r, double* u,
                              const double* kx, const double* ky) {
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      r[i] = r[i] - orc_smvp(sd, kx, ky, i, x);
      u[i] = u[i] + sd[i];
    }
#pragma omp parallel for schedule(static)
  for (long jj = h; jj < y - h; ++jj)
    for (long kk = h; kk < x - h; ++kk) {
      const long i = kk + jj * x;
      sd[i] = alpha * sd[i] + beta * r[i];
    }
}

/* ---- mixed-precision CG (no reference counterpart) ----------------------
 * Restates exploringsycl_amd/csrc/mpcg.hip: w = A p in FP32 in the SMVP
 * operand order; p = beta p + r and r = r - alpha w in FP64 rounded once to
 * FP32; u += alpha p in FP64; dot products of exact FP64 products with the
 * reference's tree.  Pinned by the FP64 path it reduces to (tests compare
 * its converged solution with the FP64 CG's) and bitwise by the GPU tests. */

static const int kAllExternal[4] = {-1, -1, -1, -1};

static float smvp32(const float* a, const float* kx, const float* ky, long i,
                    long x) {
  const float t1 = kx[i + 1] + kx[i];
  const float t2 = ky[i + x] + ky[i];
  const float diag = (1.0f + t1) + t2;
  const float c = diag * a[i];
  const float bx = kx[i + 1] * a[i + 1] + kx[i] * a[i - 1];
  const float by = ky[i + x] * a[i + x] + ky[i] * a[i - x];
  return (c - bx) - by;
}

static void halo1_float(int x, int y, int h, float* b) {
  /* orc_halo_field on a float array: copy through a double image. */
  const long n = (long)x * y;
  double* t = (double*)malloc(n * sizeof(double));
  for (long i = 0; i < n; ++i) t[i] = b[i];
  orc_halo_field(x, y, h, 1, kAllExternal, t);
  for (long i = 0; i < n; ++i) b[i] = (float)t[i];
  free(t);
}

void orc_mpcg_convert(int x, int y, int h, const double* kx, const double* ky,
                      const double* p, const double* r, const double* u0,
                      float* kx32, float* ky32, float* p32, float* r32,
                      double out[2]) {
  const long n = (long)x * y;
  for (long i = 0; i < n; ++i) {
    kx32[i] = (float)kx[i];
    ky32[i] = (float)ky[i];
    p32[i] = (float)p[i];
    r32[i] = (float)r[i];
  }
  double* part0 = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part0, { val = (double)r32[i] * (double)r32[i]; });
  out[0] = finish(part0, x, y);
  double* part1 = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part1, { val = u0[i] * u0[i]; });
  out[1] = finish(part1, x, y);
}

double orc_mpcg_calc_w(int x, int y, int h, double alpha, int apply_u,
                       double beta, int apply_p, const float* p_old,
                       float* p_new, const float* r, double* u, float* w,
                       const float* kx, const float* ky) {
  if (apply_u) {
    for (long jj = h; jj < y - h; ++jj)
      for (long kk = h; kk < x - h; ++kk) {
        const long i = kk + jj * x;
        u[i] = u[i] + alpha * (double)p_old[i];
      }
    orc_halo_field(x, y, h, 1, kAllExternal, u);
  }
  const float* p = p_old;
  if (apply_p) {
    for (long jj = h; jj < y - h; ++jj)
      for (long kk = h; kk < x - h; ++kk) {
        const long i = kk + jj * x;
        p_new[i] = (float)(beta * (double)p_old[i] + (double)r[i]);
      }
    halo1_float(x, y, h, p_new);
    p = p_new;
  }
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const float s = smvp32(p, kx, ky, i, x);
    w[i] = s;
    val = (double)s * (double)p[i];
  });
  return finish(part, x, y);
}

double orc_mpcg_calc_r(int x, int y, int h, double alpha, float* r,
                       const float* w) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const float rn = (float)((double)r[i] - alpha * (double)w[i]);
    r[i] = rn;
    val = (double)rn * (double)rn;
  });
  return finish(part, x, y);
}

double orc_dot32(int x, int y, int h, const float* a, const float* b) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, { val = (double)a[i] * (double)b[i]; });
  return finish(part, x, y);
}

void orc_mpcg_refresh(int x, int y, int h, double alpha, int apply_u,
                      double* u, const float* p, const double* u0,
                      const double* kx, const double* ky, float* r32,
                      const float* w, double out[3]) {
  if (apply_u) {
    for (long jj = h; jj < y - h; ++jj)
      for (long kk = h; kk < x - h; ++kk) {
        const long i = kk + jj * x;
        u[i] = u[i] + alpha * (double)p[i];
      }
    orc_halo_field(x, y, h, 1, kAllExternal, u);
  }
  double* part0 = alloc_partials(x, y);
  double* part1 = alloc_partials(x, y);
  double* part2 = alloc_partials(x, y);
  const long n_ = (long)x * y, groups_ = (n_ + GROUP - 1) / GROUP;
#pragma omp parallel for schedule(static)
  for (long g = 0; g < groups_; ++g) {
    double a0[GROUP], a1[GROUP], a2[GROUP];
    long i0 = g * GROUP, jj = i0 / x, kk = i0 - jj * x;
    for (int l = 0; l < GROUP; ++l) {
      const long i = i0 + l;
      a0[l] = a1[l] = a2[l] = 0.0;
      if (i < n_ && is_interior(kk, jj, x, y, h)) {
        const double rt = u0[i] - orc_smvp(u, kx, ky, i, x);
        const float rf = (float)rt;
        r32[i] = rf;
        a0[l] = (double)rf * (double)rf;
        a1[l] = rt * rt;
        a2[l] = (double)rf * (double)w[i];
      }
      if (++kk == x) { kk = 0; ++jj; }
    }
    part0[g] = orc_tree64(a0);
    part1[g] = orc_tree64(a1);
    part2[g] = orc_tree64(a2);
  }
  out[0] = finish(part0, x, y);
  out[1] = finish(part1, x, y);
  out[2] = finish(part2, x, y);
}

/* ---- flexible CG (no reference counterpart) ---------------------------
 * Restates exploringsycl_amd/csrc/fcg.hip: the Jacobi preconditioner
 * z = r / diag(A) with the diagonal formed in FP64 from the FP32
 * coefficients, and the direction update p = z + beta p_old fused with
 * w = A p and the two dots p.w, p.r. */
double orc_fcg_jacobi(int x, int y, int h, const float* r, const float* kx,
                      const float* ky, const float* w, float* z) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const double d = (1.0 + ((double)kx[i + 1] + (double)kx[i])) +
                     ((double)ky[i + x] + (double)ky[i]);
    const float zv = (float)((double)r[i] / d);
    z[i] = zv;
    val = (double)zv * (double)w[i];
  });
  return finish(part, x, y);
}

void orc_fcg_calc_w(int x, int y, int h, double alpha, int apply_u,
                    double beta, const float* p_old, float* p_new,
                    const float* z, const float* r, double* u, float* w,
                    const float* kx, const float* ky, double out[2]) {
  out[0] = orc_mpcg_calc_w(x, y, h, alpha, apply_u, beta, 1, p_old, p_new, z,
                           u, w, kx, ky);
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, { val = (double)p_new[i] * (double)r[i]; });
  out[1] = finish(part, x, y);
}

/* ---- two-level coarse correction (no counterpart in the reference) -----
 * Restates the coarse pipeline of exploringsycl_amd/csrc/fcg.hip with the
 * non-learned coarse solve: restrict8_kernel (level 0: means of 8 x 8 fine
 * blocks; level 1: means of 4 x 4 level-0 means), coarse_coef_kernel (the
 * Galerkin face coefficients), sweeps_kernel (Jacobi sweeps on the coarse
 * operator, the per-tile partials of y.rbar and y.T y, and the level-1
 * block sums of T y), joint_gamma_kernel and fcg_z1_kernel.  Every sum is
 * taken in the device's order, so the results are bitwise the device's.
 * Coarse cell i of a (hc x wc) level has faces left kc[i], bottom
 * kc[hc wc + i], right kc[i + 1] and top kc[hc wc + i + wc] (zero on the
 * domain edge: the reflective boundary). */

void orc_fcg_coarse_faces(int x, int y, int h, int c, const float* kx,
                          const float* ky, float* kc) {
  const int hc = (y - 2 * h) / c, wc = (x - 2 * h) / c;
  const double cc = (double)c * (double)c;
  for (int Y = 0; Y < hc; ++Y)
    for (int X = 0; X < wc; ++X) {
      const long corner = (long)(h + Y * c) * x + h + (long)X * c;
      double sx = 0.0, sy = 0.0;
      for (int b = 0; b < c; ++b) {
        sx += (double)kx[corner + (long)b * x];
        sy += (double)ky[corner + b];
      }
      const long i = (long)Y * wc + X;
      kc[i] = X == 0 ? 0.0f : (float)(sx / cc);
      kc[(long)hc * wc + i] = Y == 0 ? 0.0f : (float)(sy / cc);
    }
}

void orc_fcg_restrict8(int x, int y, int h, const float* r, float* rbar) {
  const int hc = (y - 2 * h) / 8, wc = (x - 2 * h) / 8;
  for (int Y = 0; Y < hc; ++Y)
    for (int X = 0; X < wc; ++X) {
      double t[2];
      for (int half = 0; half < 2; ++half) {
        double q[8];
        for (int b = 0; b < 8; ++b) {
          const float* p = r + (long)(h + 8 * Y + b) * x + h + 8 * X + 4 * half;
          q[b] = ((double)p[0] + (double)p[1]) + ((double)p[2] + (double)p[3]);
        }
        t[half] = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
      }
      rbar[(long)Y * wc + X] = (float)((t[0] + t[1]) / 64.0);
    }
}

void orc_fcg_restrict4(int hc, int wc, const float* fine, float* coarse) {
  const long wf = 4L * wc;
  for (int Y = 0; Y < hc; ++Y)
    for (int X = 0; X < wc; ++X) {
      double rs[4];
      for (int j = 0; j < 4; ++j) {
        const float* p = fine + (4L * Y + j) * wf + 4L * X;
        rs[j] = ((double)p[0] + (double)p[1]) + ((double)p[2] + (double)p[3]);
      }
      coarse[(long)Y * wc + X] = (float)(((rs[0] + rs[1]) + (rs[2] + rs[3])) / 16.0);
    }
}

/* The four face coefficients of coarse cell (X, Y). */
static void coarse_faces_at(int hc, int wc, const float* kc, int X, int Y,
                            double* kl, double* kr, double* kb, double* kt) {
  const long n = (long)hc * wc, i = (long)Y * wc + X;
  *kl = (double)kc[i];
  *kb = (double)kc[n + i];
  *kr = X + 1 < wc ? (double)kc[i + 1] : 0.0;
  *kt = Y + 1 < hc ? (double)kc[n + i + wc] : 0.0;
}

#define SWEEP_TILE 16

/* Tree of the 256 slots of one 16 x 16 tile, as a 256-thread block's LDS
 * tree (stride 128, 64, ..., 1). */
static double tile_tree(double* s) {
  for (int st = 128; st > 0; st >>= 1)
    for (int t = 0; t < st; ++t) s[t] = s[t] + s[t + st];
  return s[0];
}

void orc_fcg_sweeps(int hc, int wc, const float* kc, const float* rbar,
                    int sweeps, float* y_out, double* part, double* cross) {
  const long n = (long)hc * wc;
  double* ya = (double*)calloc((size_t)n, sizeof(double));
  double* yb = (double*)calloc((size_t)n, sizeof(double));
  double* ty = (double*)calloc((size_t)n, sizeof(double));
  /* Jacobi sweeps from zero (TeaLeaf/c_kernels/sycl/jacobi.cpp:85-91 on the
   * coarse operator), in FP64. */
  for (int s = 0; s < sweeps; ++s) {
    for (int Y = 0; Y < hc; ++Y)
      for (int X = 0; X < wc; ++X) {
        const long i = (long)Y * wc + X;
        double kl, kr, kb, kt;
        coarse_faces_at(hc, wc, kc, X, Y, &kl, &kr, &kb, &kt);
        const double yr = X + 1 < wc ? ya[i + 1] : 0.0;
        const double yl = X > 0 ? ya[i - 1] : 0.0;
        const double yt = Y + 1 < hc ? ya[i + wc] : 0.0;
        const double ybt = Y > 0 ? ya[i - wc] : 0.0;
        yb[i] = ((double)rbar[i] + (kr * yr + kl * yl) + (kt * yt + kb * ybt)) /
                ((1.0 + (kl + kr)) + (kb + kt));
      }
    double* t = ya;
    ya = yb;
    yb = t;
  }
  for (long i = 0; i < n; ++i) y_out[i] = (float)ya[i];
  /* T y with the rounded y, in the operand order of TeaLeaf's SMVP
   * (shared.h:59-63). */
  for (int Y = 0; Y < hc; ++Y)
    for (int X = 0; X < wc; ++X) {
      const long i = (long)Y * wc + X;
      double kl, kr, kb, kt;
      coarse_faces_at(hc, wc, kc, X, Y, &kl, &kr, &kb, &kt);
      const double yc = (double)y_out[i];
      const double yr = X + 1 < wc ? (double)y_out[i + 1] : 0.0;
      const double yl = X > 0 ? (double)y_out[i - 1] : 0.0;
      const double yt = Y + 1 < hc ? (double)y_out[i + wc] : 0.0;
      const double ybt = Y > 0 ? (double)y_out[i - wc] : 0.0;
      ty[i] = (((1.0 + (kr + kl)) + (kt + kb)) * yc - (kr * yr + kl * yl)) -
              (kt * yt + kb * ybt);
    }
  const int ntx = (wc + SWEEP_TILE - 1) / SWEEP_TILE;
  const int nty = (hc + SWEEP_TILE - 1) / SWEEP_TILE;
  const int ntiles = ntx * nty;
  for (int TY = 0; TY < nty; ++TY)
    for (int TX = 0; TX < ntx; ++TX) {
      double sa[256], sb[256];
      for (int t = 0; t < 256; ++t) {
        const int X = TX * SWEEP_TILE + t % SWEEP_TILE;
        const int Y = TY * SWEEP_TILE + t / SWEEP_TILE;
        sa[t] = sb[t] = 0.0;
        if (X < wc && Y < hc) {
          const long i = (long)Y * wc + X;
          const double yc = (double)y_out[i];
          sa[t] = yc * (double)rbar[i];
          sb[t] = yc * ty[i];
        }
      }
      part[TY * ntx + TX] = tile_tree(sa);
      part[ntiles + TY * ntx + TX] = tile_tree(sb);
    }
  if (cross) {
    const int hc2 = hc / 4, wc2 = wc / 4;
    for (int Y = 0; Y < hc2; ++Y)
      for (int X = 0; X < wc2; ++X) {
        double rs[4];
        for (int j = 0; j < 4; ++j) {
          const double* p = ty + (4L * Y + j) * wc + 4L * X;
          rs[j] = (p[0] + p[1]) + (p[2] + p[3]);
        }
        cross[(long)Y * wc2 + X] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
      }
  }
  free(ya);
  free(yb);
  free(ty);
}

/* Sum of v[0..n) (times (double)f[k] if f) in the order of one 1024-thread
 * block: thread t adds k = t, t + 1024, ... in turn, then the LDS tree
 * (stride 512, 256, ..., 1). */
#define GAMMA_THREADS 1024
static double block_sum(const double* v, const float* f, long n) {
  double s[GAMMA_THREADS];
  for (int t = 0; t < GAMMA_THREADS; ++t) {
    double a = 0.0;
    for (long k = t; k < n; k += GAMMA_THREADS) a += f ? v[k] * (double)f[k] : v[k];
    s[t] = a;
  }
  for (int st = GAMMA_THREADS / 2; st > 0; st >>= 1)
    for (int t = 0; t < st; ++t) s[t] = s[t] + s[t + st];
  return s[0];
}

void orc_fcg_gamma(int levels, const double* part0, int n0, const double* part1,
                   int n1, const double* cross, const float* y1, long ncross,
                   double c0sq, double c1sq, float g[2]) {
  const double a0 = block_sum(part0, NULL, n0), b0 = block_sum(part0 + n0, NULL, n0);
  g[0] = g[1] = 0.0f;
  if (levels == 1) {
    const double q = a0 / b0;
    g[0] = (b0 > 0.0 && isfinite(q)) ? (float)q : 0.0f;
    return;
  }
  const double a1 = c0sq * a0, b11 = c0sq * b0;
  const double a2 = c1sq * block_sum(part1, NULL, n1);
  const double b22 = c1sq * block_sum(part1 + n1, NULL, n1);
  const double b12 = c0sq * block_sum(cross, y1, ncross);
  const double det = b11 * b22 - b12 * b12;
  double g1, g2;
  if (b11 > 0.0 && b22 > 0.0 && det > 1e-12 * b11 * b22) {
    g1 = (a1 * b22 - a2 * b12) / det;
    g2 = (a2 * b11 - a1 * b12) / det;
  } else {
    g1 = b11 > 0.0 ? a1 / b11 : 0.0;
    g2 = 0.0;
  }
  g[0] = isfinite(g1) ? (float)g1 : 0.0f;
  g[1] = isfinite(g2) ? (float)g2 : 0.0f;
}

double orc_fcg_prolong(int x, int y, int h, const float* r, const float* kx,
                       const float* ky, const float* w, int levels,
                       const float* y0, int wc0, int c0, float g0,
                       const float* y1, int wc1, int c1, float g1, float* z) {
  double* part = alloc_partials(x, y);
  GROUP_REDUCE(x, y, h, part, {
    const double d = (1.0 + ((double)kx[i + 1] + (double)kx[i])) +
                     ((double)ky[i + x] + (double)ky[i]);
    double v = (double)r[i] / d;
    if (levels > 0) v = v + (double)g0 * (double)y0[((jj - h) / c0) * wc0 + (kk - h) / c0];
    if (levels > 1) v = v + (double)g1 * (double)y1[((jj - h) / c1) * wc1 + (kk - h) / c1];
    const float zv = (float)v;
    z[i] = zv;
    val = (double)zv * (double)w[i];
  });
  return finish(part, x, y);
}
