/*
 * CPU oracle for the TeaLeaf conjugate-gradient path and the reference's
 * other solvers (Jacobi, Chebyshev, PPCG).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * kernel_interface adapter or the host driver) may link or call this code;
 * only tests/, the smoke check and the CPU-baseline leg of bench.py use it,
 * and only as the checker or the timed CPU reference.
 *
 * Every function restates, in plain C on host arrays, the FP64 arithmetic of
 * one SYCL kernel of the reference backend (paths relative to the reference
 * repository, cited per function in tea_oracle.c).  Arrays are row-major with
 * index = kk + jj*x, x and y include the halo (x = x_cells + 2*halo_depth).
 *
 * Reductions follow the reference's tree exactly (sycl_shared.hpp:35-79 and
 * the first-level trees inside cg.cpp / solver_methods.cpp /
 * field_summary.cpp): the padded linear index range [0, 64*ceil(x*y/64)) is
 * cut into groups of 64, every non-interior slot contributes +0.0, each group
 * is summed by the pairwise tree "for s=1,2,..,32: a[2s*l] += a[2s*l+s]", and
 * the partials are summed again by the same tree, zero-padded to 64, until one
 * value is left.  The build compiles this file with -ffp-contract=off so each
 * multiply and add rounds on its own, as written.
 *
 * Scalar results are RETURNED; the kernel_interface adapter
 * (backend_oracle.cpp) adds or assigns them as the reference's SYCL
 * backend does.
 *
 * Parity status: see the header of tea_oracle.c.
 */
#ifndef TEA_ORACLE_H
#define TEA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CONDUCTIVITY = 1, ORC_RECIP_CONDUCTIVITY = 2 };
enum { ORC_RECTANGULAR = 0, ORC_CIRCULAR = 1, ORC_POINT = 2 };
enum { ORC_LEFT = 0, ORC_RIGHT = 1, ORC_BOTTOM = 2, ORC_TOP = 3 };
enum {
  ORC_FIELD_DENSITY = 0, ORC_FIELD_ENERGY0 = 1, ORC_FIELD_ENERGY1 = 2,
  ORC_FIELD_U = 3, ORC_FIELD_P = 4, ORC_FIELD_SD = 5, ORC_NUM_FIELDS = 6
};

/* The reference's reduction tree over an arbitrary array (zero padded). */
double orc_tree_sum(const double* v, size_t n);
/* One 64-wide pairwise tree, in place; returns a[0]. */
double orc_tree64(double* a);

/* Mesh and state setup (set_chunk_data.cpp, set_chunk_state.cpp,
 * store_energy.cpp). */
void orc_set_chunk_data(int x, int y, int h, double x_min, double y_min,
                        double dx, double dy, double* vertex_x,
                        double* vertex_y, double* cell_x, double* cell_y,
                        double* volume, double* x_area, double* y_area);
void orc_set_chunk_initial_state(int x, int y, double energy, double density,
                                 double* energy0, double* density_f);
void orc_set_chunk_state(int x, int y, int geometry, double sx_min,
                         double sy_min, double sx_max, double sy_max,
                         double radius, double s_density, double s_energy,
                         double* energy0, double* density, double* u,
                         const double* cell_x, const double* cell_y,
                         const double* vertex_x, const double* vertex_y);
void orc_store_energy(int x, int y, double* energy, const double* energy0);

/* Reflective halo copies for one buffer, faces in the reference order
 * left, right, top, bottom; a face is updated only if neighbours[face] < 0
 * (EXTERNAL_FACE). */
void orc_halo_field(int x, int y, int h, int depth, const int neighbours[4],
                    double* b);

/* field_summary.cpp: out = {vol, mass, ie, temp}. */
void orc_field_summary(int x, int y, int h, const double* volume,
                       const double* density, const double* energy0,
                       const double* u, double out[4]);

/* cg.cpp: returns rro. */
double orc_cg_init(int x, int y, int h, int coefficient, double rx, double ry,
                   double* p, double* r, double* u, double* w, double* kx,
                   double* ky, const double* density, const double* energy);
/* cg.cpp: returns pw. */
double orc_cg_calc_w(int x, int y, int h, double* w, const double* p,
                     const double* kx, const double* ky);
/* cg.cpp: returns rrn. */
double orc_cg_calc_ur(int x, int y, int h, double alpha, double* u, double* r,
                      const double* p, const double* w);
void orc_cg_calc_p(int x, int y, int h, double beta, double* p,
                   const double* r);

/* solver_methods.cpp */
void orc_copy_u(int x, int y, int h, const double* u, double* u0);
void orc_calculate_residual(int x, int y, int h, const double* u,
                            const double* u0, double* r, const double* kx,
                            const double* ky);
double orc_calculate_2norm(int x, int y, int h, const double* b);
void orc_finalise(int x, int y, int h, const double* u, const double* density,
                  double* energy);

/* pack_halos.cpp: face 0 left, 1 right, 2 bottom, 3 top; pack != 0 copies
 * field -> buffer, else buffer -> field. */
void orc_pack_or_unpack(int x, int y, int h, int depth, int face, int pack,
                        double* field, double* buffer);

/* jacobi.cpp, cheby.cpp, ppcg.cpp, composed as kernel_interface.cpp:246-328
 * composes them.  orc_jacobi_iterate returns the error sum |u_new - u_old|. */
void orc_jacobi_init(int x, int y, int h, int coefficient, double rx,
                     double ry, double* u, double* u0, const double* density,
                     const double* energy, double* kx, double* ky);
double orc_jacobi_iterate(int x, int y, int h, double* u, const double* u0,
                          double* r, const double* kx, const double* ky);
void orc_cheby_init(int x, int y, int h, double theta, double* p, double* r,
                    const double* u, const double* u0, double* w,
                    const double* kx, const double* ky);
void orc_cheby_iterate(int x, int y, int h, double alpha, double beta,
                       double* p, double* r, double* u, const double* u0,
                       double* w, const double* kx, const double* ky);
void orc_ppcg_init(int x, int y, int h, double theta, double* sd,
                   const double* r);
void orc_ppcg_inner_iteration(int x, int y, int h, double alpha, double beta,
                              double* sd, double* r, double* u,
                              const double* kx, const double* ky);

/* Mixed-precision CG (no reference counterpart; restates the arithmetic of
 * exploringsycl_amd/csrc/mpcg.hip): FP32 p, r, w, kx, ky; FP64 u, u0,
 * dot products and scalars.  Halo updates of p and u use orc_halo_field
 * (depth 1, all faces external). */
void orc_mpcg_convert(int x, int y, int h, const double* kx, const double* ky,
                      const double* p, const double* r, const double* u0,
                      float* kx32, float* ky32, float* p32, float* r32,
                      double out[2]);
double orc_mpcg_calc_w(int x, int y, int h, double alpha, int apply_u,
                       double beta, int apply_p, const float* p_old,
                       float* p_new, const float* r, double* u, float* w,
                       const float* kx, const float* ky);
double orc_mpcg_calc_r(int x, int y, int h, double alpha, float* r,
                       const float* w);
void orc_mpcg_refresh(int x, int y, int h, double alpha, int apply_u,
                      double* u, const float* p, const double* u0,
                      const double* kx, const double* ky, float* r32,
                      const float* w, double out[3]);
/* Sum over the interior of a.b (FP64 products of FP32 values) with the
 * reference's reduction tree: p.r and r.w of the mixed-precision CG's
 * flexible recursion (alpha = p.r / p.w, beta = -r.w / p.w). */
double orc_dot32(int x, int y, int h, const float* a, const float* b);

/* Flexible CG (no reference counterpart; restates the arithmetic of
 * exploringsycl_amd/csrc/fcg.hip).  orc_fcg_jacobi: z = r / diag(A),
 * returns z.w.  orc_fcg_calc_w: orc_mpcg_calc_w with p = z + beta p_old,
 * out = {p.w, p.r}. */
double orc_fcg_jacobi(int x, int y, int h, const float* r, const float* kx,
                      const float* ky, const float* w, float* z);
void orc_fcg_calc_w(int x, int y, int h, double alpha, int apply_u,
                    double beta, const float* p_old, float* p_new,
                    const float* z, const float* r, double* u, float* w,
                    const float* kx, const float* ky, double out[2]);

/* Two-level coarse correction of the flexible CG with Jacobi sweeps as the
 * coarse solve (no reference counterpart; restates the coarse pipeline of
 * exploringsycl_amd/csrc/fcg.hip in the device's summation order).
 *   orc_fcg_coarse_faces: Galerkin face coefficients of the level that
 *     coarsens by c (kc: hc*wc left faces, then hc*wc bottom faces).
 *   orc_fcg_restrict8 / orc_fcg_restrict4: block means of r over 8 x 8 fine
 *     cells, and of a level's means over 4 x 4 cells (hc, wc: coarse size).
 *   orc_fcg_sweeps: `sweeps` Jacobi sweeps on T y = rbar from zero; y_out
 *     (FP32), the per-16x16-tile partials of y.rbar (part[0..ntiles)) and
 *     y.T y (part[ntiles..2 ntiles)), and, if cross != NULL, the sums of
 *     T y over 4 x 4 blocks ((hc/4) x (wc/4)).
 *   orc_fcg_gamma: the energy-optimal factors g of one level or jointly of
 *     two (cross: level-0 block sums, y1: level-1 y, ncross cells).
 *   orc_fcg_prolong: z = r / diag(A) + sum_l g_l y_l(block), returns z.w. */
void orc_fcg_coarse_faces(int x, int y, int h, int c, const float* kx,
                          const float* ky, float* kc);
void orc_fcg_restrict8(int x, int y, int h, const float* r, float* rbar);
void orc_fcg_restrict4(int hc, int wc, const float* fine, float* coarse);
void orc_fcg_sweeps(int hc, int wc, const float* kc, const float* rbar,
                    int sweeps, float* y_out, double* part, double* cross);
void orc_fcg_gamma(int levels, const double* part0, int n0, const double* part1,
                   int n1, const double* cross, const float* y1, long ncross,
                   double c0sq, double c1sq, float g[2]);
double orc_fcg_prolong(int x, int y, int h, const float* r, const float* kx,
                       const float* ky, const float* w, int levels,
                       const float* y0, int wc0, int c0, float g0,
                       const float* y1, int wc1, int c1, float g1, float* z);

/* The 5-point operator applied at one cell (shared.h:59-63). */
double orc_smvp(const double* a, const double* kx, const double* ky, long i,
                long x);

#ifdef __cplusplus
}
#endif

#endif
