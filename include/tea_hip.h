/*
 * tea_hip.h — C ABI of the MI355X (gfx950) TeaLeaf CG backend, libteahip.so.
 *
 * This is the thin layer under TeaLeaf's kernel_interface.h: every entry
 * point here replaces one run_* function of the reference's plugin boundary
 * (TeaLeaf/kernel_interface.h:13-71, implemented for SYCL in
 * TeaLeaf/c_kernels/sycl/kernel_interface.cpp) with plain pointers, sizes and
 * an opaque per-chunk handle instead of Chunk* / Settings*.  The C++ adapter
 * exploringsycl_amd/csrc/kernel_interface_hip.cpp maps the 24 run_* symbols
 * onto these calls, so the reference's own host code links against it
 * unchanged.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative TEA_HIP_E* code on
 *    failure; tea_hip_last_error() then describes the failure.  No entry point
 *    exits the process (the adapter turns errors into the host's die()).
 *  - Fields are row-major FP64 device arrays, index = kk + jj*x, with x and y
 *    including the halo (x = x_cells + 2*halo_depth), owned by the handle.
 *  - Scalar results (rro, pw, rrn, norm, field-summary sums) are ASSIGNED to
 *    the out-parameter; callers that accumulate over chunks add them up.
 *    They are computed with the reference's reduction tree
 *    (TeaLeaf/c_kernels/sycl/sycl_shared.hpp:35-79) and are bitwise equal to
 *    the FP64 CPU restatement of the reference kernels.
 *  - Calls on one handle are serialised on one HIP stream.  A call that
 *    returns a scalar returns once the values it needs have arrived: the
 *    final reduction kernel writes them into mapped host memory, which the
 *    call polls, so it may return while that kernel is still retiring.
 *    Raw device access through tea_hip_field therefore needs tea_hip_flush
 *    first: work may still be queued or held back (see tea_hip_set_fused
 *    and tea_hip_local_halos).  A handle must not be used from two host
 *    threads at once.
 */
#ifndef TEA_HIP_H
#define TEA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tea_hip_chunk tea_hip_chunk;

/* Error codes. */
#define TEA_HIP_OK 0
#define TEA_HIP_EINVAL -1  /* bad argument (size, field id, depth, ...) */
#define TEA_HIP_EHIP -2    /* a HIP runtime call failed */
#define TEA_HIP_ENOMEM -3  /* device or pinned allocation failed */

/* Field identifiers, in the order of the FieldBufferType members of
 * TeaLeaf/chunk.h:39-66, plus the backend's comms buffer
 * (TeaLeaf/c_kernels/sycl/chunk_extension.h:9). */
enum {
  TEA_FIELD_DENSITY0 = 0, TEA_FIELD_DENSITY, TEA_FIELD_ENERGY0,
  TEA_FIELD_ENERGY, TEA_FIELD_U, TEA_FIELD_U0, TEA_FIELD_P, TEA_FIELD_R,
  TEA_FIELD_MI, TEA_FIELD_W, TEA_FIELD_KX, TEA_FIELD_KY, TEA_FIELD_SD,
  TEA_FIELD_CELL_X, TEA_FIELD_CELL_Y, TEA_FIELD_CELL_DX, TEA_FIELD_CELL_DY,
  TEA_FIELD_VERTEX_DX, TEA_FIELD_VERTEX_DY, TEA_FIELD_VERTEX_X,
  TEA_FIELD_VERTEX_Y, TEA_FIELD_VOLUME, TEA_FIELD_X_AREA, TEA_FIELD_Y_AREA,
  TEA_FIELD_COMMS_BUFFER, TEA_NUM_FIELDS
};

/* Halo-exchange field slots, as indexed by Settings::fields_to_exchange
 * (TeaLeaf/shared.h:39-44). */
enum {
  TEA_HALO_DENSITY = 0, TEA_HALO_ENERGY0 = 1, TEA_HALO_ENERGY1 = 2,
  TEA_HALO_U = 3, TEA_HALO_P = 4, TEA_HALO_SD = 5, TEA_HALO_NUM = 6
};

/* Faces, as indexed by Chunk::neighbours; a negative neighbour is an
 * external (reflective) face (TeaLeaf/shared.h:30-35). */
enum { TEA_FACE_LEFT = 0, TEA_FACE_RIGHT = 1, TEA_FACE_BOTTOM = 2,
       TEA_FACE_TOP = 3 };

/* Conductivity models (TeaLeaf/shared.h:46-47). */
enum { TEA_CONDUCTIVITY = 1, TEA_RECIP_CONDUCTIVITY = 2 };

/* State geometries (TeaLeaf/settings.h:127-132). */
enum { TEA_GEOMETRY_RECTANGULAR = 0, TEA_GEOMETRY_CIRCULAR = 1,
       TEA_GEOMETRY_POINT = 2 };

/* ---- lifecycle ---------------------------------------------------------- */

/* Creates the device state of one chunk of x*y cells (halo included) on
 * device `device`: all 25 fields, zero-filled, the reduction workspace and
 * the stream.  Replaces run_kernel_initialise
 * (TeaLeaf/kernel_interface.h:18; kernel_initialise.cpp:31-80). */
int tea_hip_chunk_create(int device, int x, int y, int halo_depth,
                         tea_hip_chunk** out);

/* Frees everything the handle owns.  Replaces run_kernel_finalise
 * (TeaLeaf/kernel_interface.h:20; kernel_initialise.cpp:83-126). */
int tea_hip_chunk_destroy(tea_hip_chunk* chunk);

/* Device address and element count of a field.  The addresses of
 * TEA_FIELD_P, TEA_FIELD_U, TEA_FIELD_R, TEA_FIELD_W and TEA_FIELD_SD can
 * change after solver calls (the fused CG, Jacobi, Chebyshev and PPCG
 * iterations and the CG run-ahead write the new p, u, r, w or sd into a
 * second buffer and exchange the two); ask again after each call instead
 * of caching them.  The address is returned as
 * is: deferred updates (see tea_hip_set_fused and tea_hip_local_halos) may
 * not have reached the field yet; call tea_hip_flush before reading or
 * writing through it. */
int tea_hip_field(tea_hip_chunk* chunk, int field, double** dev_ptr,
                  size_t* count);

/* Blocking copies between host memory and a field (count elements from the
 * start of the field).  Pending deferred CG work is applied first. */
int tea_hip_field_write(tea_hip_chunk* chunk, int field, const double* host,
                        size_t count);
int tea_hip_field_read(tea_hip_chunk* chunk, int field, double* host,
                       size_t count);

/* Waits for all work queued on the handle's stream (deferred work stays
 * deferred: see tea_hip_flush). */
int tea_hip_synchronize(tea_hip_chunk* chunk);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  tea_hip_calculate_2norm for a field the caller owns:
 * *norm (a DEVICE pointer, written on `stream`, a hipStream_t or NULL for
 * the null stream) = the sum of b^2 over the interior of the x*y FP64
 * device array `field` (halo_depth cells left out at each edge), with the
 * reference's reduction tree, bit for bit as tea_hip_calculate_2norm.
 * Allocates nothing: its working memory is a fixed device array of 16448
 * doubles per GPU (held until the process ends), so x*y may be at most
 * 16448 * 4096 = 67,371,008 cells (8192^2 with a halo of 2); larger grids
 * are rejected with TEA_HIP_EINVAL.  Calls that overlap on one GPU share
 * that array and race.  Asynchronous: the result is ready when `stream`
 * is. */
int tea_hip_calculate_2norm_noalloc(int device, const double* field, int x, int y,
                                    int halo_depth, double* norm, void* stream);

/* Text of the last error on the calling thread ("" if none). */
const char* tea_hip_last_error(void);

/* ---- initialisation kernels -------------------------------------------- */

/* Mesh: vertices, cell centres, volume and face areas.  x_min/y_min are the
 * chunk's lower-left coordinates.  Replaces run_set_chunk_data
 * (TeaLeaf/kernel_interface.h:14; set_chunk_data.cpp:8-64). */
int tea_hip_set_chunk_data(tea_hip_chunk* chunk, double x_min, double y_min,
                           double dx, double dy);

/* Applies state 0 to every cell, then states 1..num_states-1 in order, and
 * sets u = energy0*density.  Arrays are indexed by state; bounds are the
 * already shrunk ones the host parser produces.  Replaces
 * run_set_chunk_state (TeaLeaf/kernel_interface.h:16;
 * set_chunk_state.cpp:8-92). */
int tea_hip_set_chunk_state(tea_hip_chunk* chunk, int num_states,
                            const double* density, const double* energy,
                            const double* x_min, const double* y_min,
                            const double* x_max, const double* y_max,
                            const double* radius, const int* geometry);

/* energy := energy0 on every cell.  Replaces run_store_energy
 * (TeaLeaf/kernel_interface.h:27; store_energy.cpp:6-24). */
int tea_hip_store_energy(tea_hip_chunk* chunk);

/* ---- solver-wide kernels ----------------------------------------------- */

/* Reflective halo update of depth `depth` (1 <= depth <= halo_depth) for
 * every field whose fields_to_exchange slot is non-zero, on every face whose
 * neighbour is negative; one launch covers all fields and faces and gives the
 * same result as the reference's left, right, top, bottom order.  Replaces
 * run_local_halos (TeaLeaf/kernel_interface.h:23;
 * kernel_interface.cpp:100-127, local_halos.cpp:7-102).
 * In fused mode some updates are held back and folded into the pass that
 * next reads the field: the depth-1 U/P update after tea_hip_cg_calc_p
 * (into the next tea_hip_cg_calc_w), the SD update after a held PPCG inner
 * step or the U update after a held Chebyshev step, and a depth-1 SD or U
 * update requested with nothing held (the next two-step pass reflects the
 * ring of the field it stages).  Every other entry point applies them
 * first, and tea_hip_flush applies them at once. */
int tea_hip_local_halos(tea_hip_chunk* chunk, int depth,
                        const int neighbours[4],
                        const int fields_to_exchange[6]);

/* Remote-halo strip of one field (one of the handle's fields, by device
 * address): pack != 0 copies the `depth` rows or columns next to the
 * interior on `face` (TEA_FACE_*) into host_buffer; pack == 0 copies
 * host_buffer into that face's halo cells.  The strip has y*depth values
 * for the left and right faces (row by row) and x*depth for the bottom and
 * top, in the reference's order, staged through the comms buffer.  Blocks
 * until the copy is done.  Replaces run_pack_or_unpack
 * (TeaLeaf/kernel_interface.h:25-27; kernel_interface.cpp:129-170,
 * pack_halos.cpp:7-184). */
int tea_hip_pack_or_unpack(tea_hip_chunk* chunk, int depth, int face,
                           int pack, const double* field, double* host_buffer);

/* Interior sums vol = sum V, mass = sum V*rho, ie = sum V*rho*energy0,
 * temp = sum V*rho*u.  Replaces run_field_summary
 * (TeaLeaf/kernel_interface.h:29-31; field_summary.cpp:6-151). */
int tea_hip_field_summary(tea_hip_chunk* chunk, double* vol, double* mass,
                          double* ie, double* temp);

/* ---- CG solver kernels -------------------------------------------------- */

/* p = r = 0, u = energy*density, conductivity w, face coefficients kx, ky,
 * then w = A u, r = u - w, p = r and *rro = r.p.  Replaces run_cg_init
 * (TeaLeaf/kernel_interface.h:34-36; cg.cpp:7-134). */
int tea_hip_cg_init(tea_hip_chunk* chunk, int coefficient, double rx,
                    double ry, double* rro);

/* w = A p on the interior and *pw = p.w.  Replaces run_cg_calc_w
 * (TeaLeaf/kernel_interface.h:37-38; cg.cpp:137-195).
 * Run-ahead (fused mode, x even, all reductions polled): after this call
 * the device forms alpha = rro / pw itself (rro: the value the last
 * tea_hip_cg_init or tea_hip_cg_calc_ur returned) and already runs
 * r - alpha w into a spare buffer; after tea_hip_cg_calc_ur it forms beta
 * and runs the next w = A p likewise.  The next call takes that work over
 * when its alpha (and, for tea_hip_cg_calc_w, the deferred beta, u update
 * and U/P halo update) is bitwise what the device used; any other call
 * drops it.  Results are bitwise those without it (TEA_HIP_RUN_AHEAD=0). */
int tea_hip_cg_calc_w(tea_hip_chunk* chunk, double* pw);

/* u += alpha p, r -= alpha w, *rrn = r.r.  Replaces run_cg_calc_ur
 * (TeaLeaf/kernel_interface.h:39-40; cg.cpp:198-254). */
int tea_hip_cg_calc_ur(tea_hip_chunk* chunk, double alpha, double* rrn);

/* p = beta p + r.  Replaces run_cg_calc_p (TeaLeaf/kernel_interface.h:41-42;
 * cg.cpp:257-281). */
int tea_hip_cg_calc_p(tea_hip_chunk* chunk, double beta);

/* ---- shared solver kernels --------------------------------------------- */

/* u0 = u on the interior.  Replaces run_copy_u
 * (TeaLeaf/kernel_interface.h:63-64; solver_methods.cpp:7-31). */
int tea_hip_copy_u(tea_hip_chunk* chunk);

/* r = u0 - A u on the interior.  Replaces run_calculate_residual
 * (TeaLeaf/kernel_interface.h:65-66; solver_methods.cpp:34-65). */
int tea_hip_calculate_residual(tea_hip_chunk* chunk);

/* *norm = sum over the interior of b^2, where b is one of the handle's
 * fields given by its device address (as returned by tea_hip_field).
 * Replaces run_calculate_2norm (TeaLeaf/kernel_interface.h:67-68;
 * solver_methods.cpp:68-117). */
int tea_hip_calculate_2norm(tea_hip_chunk* chunk, const double* buffer,
                            double* norm);

/* energy = u / density on the interior.  Replaces run_finalise
 * (TeaLeaf/kernel_interface.h:69-70; solver_methods.cpp:120-145). */
int tea_hip_finalise(tea_hip_chunk* chunk);

/* ---- Jacobi, Chebyshev and PPCG solver kernels -------------------------- */

/* u0 = u = energy*density one cell in from the edge of the padded grid, and
 * the face coefficients kx, ky from the density (conductivity model
 * `coefficient`).  Replaces run_jacobi_init (TeaLeaf/kernel_interface.h:
 * 52-53; jacobi.cpp:7-54). */
int tea_hip_jacobi_init(tea_hip_chunk* chunk, int coefficient, double rx,
                        double ry);

/* One Jacobi sweep: r = u on every cell, then on the interior
 * u = (u0 + kx' r_right + kx r_left + ky' r_top + ky r_bottom) /
 *     (1 + kx + kx' + ky + ky'),
 * and *error = the sum over the interior of |u - r|.  The sweep writes the
 * new u into r's buffer and exchanges the two fields' addresses.  Replaces
 * run_jacobi_iterate (TeaLeaf/kernel_interface.h:54-55;
 * kernel_interface.cpp:287-300, jacobi.cpp:57-137). */
int tea_hip_jacobi_iterate(tea_hip_chunk* chunk, double* error);

/* w = A u, r = u0 - w, p = r / theta on the interior.  Replaces
 * run_cheby_init (TeaLeaf/kernel_interface.h:46-47; cheby.cpp:7-43). */
int tea_hip_cheby_init(tea_hip_chunk* chunk, double theta);

/* One Chebyshev step on the interior: w = A u, r = u0 - w,
 * p = alpha p + beta r, u = u + p (one pass; the new u goes to a second
 * buffer whose address replaces u's).  In the fused mode the step is held
 * back until the next call: if that is the depth-1 U halo update and then
 * the next step, both steps run as one pass (u and p then move to second
 * buffers); any other entry point runs it first.  Replaces run_cheby_iterate
 * (TeaLeaf/kernel_interface.h:48-49; kernel_interface.cpp:258-271,
 * cheby.cpp:46-110). */
int tea_hip_cheby_iterate(tea_hip_chunk* chunk, double alpha, double beta);

/* sd = r / theta on the interior.  Replaces run_ppcg_init
 * (TeaLeaf/kernel_interface.h:58-59; ppcg.cpp:7-31). */
int tea_hip_ppcg_init(tea_hip_chunk* chunk, double theta);

/* One inner Chebyshev step of PPCG on the interior: r = r - A sd,
 * u = u + sd, sd = alpha sd + beta r (one pass; the new sd goes to a second
 * buffer whose address replaces sd's).  In the fused mode the step is held
 * back until the next call: if that is the depth-1 SD halo update and then
 * the next inner step, both steps run as one pass (sd and r then move to
 * second buffers); any other entry point runs it first.  Replaces
 * run_ppcg_inner_iteration
 * (TeaLeaf/kernel_interface.h:60-61; kernel_interface.cpp:314-328,
 * ppcg.cpp:34-94). */
int tea_hip_ppcg_inner_iteration(tea_hip_chunk* chunk, double alpha,
                                 double beta);

/* ---- mixed-precision CG --------------------------------------------------
 * No counterpart in the reference (its CG is FP64 throughout).  p, r, w and
 * the face coefficients are stored in FP32; u, u0, the dot products, alpha
 * and beta stay FP64, and tea_hip_mpcg_refresh replaces the recursive
 * residual by the FP64 true residual u0 - A u.  One chunk per GPU: all four
 * faces must be external.  A solve is: init; then per iteration calc_w
 * (applying the previous alpha to u and beta to p) and calc_r; refresh every
 * few iterations and at convergence.  The dots returned serve the flexible
 * CG(1) recursion with the identity preconditioner (Notay 2000):
 * alpha = p.r / p.w and beta = -r.w / p.w, which equal r.r / p.w and
 * r'.r' / r.r in exact arithmetic but keep consecutive directions
 * A-orthogonal when w = A p is formed in FP32.  Dot products use the same
 * reduction tree as the FP64 path. */

/* FP64 initialisation as tea_hip_cg_init, the depth-1 U/P halos and
 * u0 = u (tea_hip_copy_u), then the FP32 copies; *rr = r.r of the FP32
 * residual, *bb = u0.u0 (the right-hand side's squared norm). */
int tea_hip_mpcg_init(tea_hip_chunk* chunk, int coefficient, double rx,
                      double ry, double* rr, double* bb);

/* If apply_u: u += alpha p_old (FP64).  If apply_p: p = beta p_old + r.
 * Then w = A p (FP32), *pw = p.w and *pr = p.r.  With four cells per lane
 * (x a multiple of 4) two calls in turn hold their u update back and the
 * third makes all three in one pass over u, rounded as if made one by one
 * (TEA_HIP_HOLD_U=0 makes each at once); every other entry point except
 * tea_hip_mpcg_calc_r and the FCG steps makes held updates first, and
 * tea_hip_flush does before u is read through tea_hip_field's address. */
int tea_hip_mpcg_calc_w(tea_hip_chunk* chunk, double alpha, int apply_u,
                        double beta, int apply_p, double* pw, double* pr);

/* r -= alpha w, *rr = r.r and *rw = r.w (the updated r). */
int tea_hip_mpcg_calc_r(tea_hip_chunk* chunk, double alpha, double* rr,
                        double* rw);

/* If apply_u: u += alpha p.  Then r = u0 - A u in FP64 (stored as FP32);
 * *rr = r.r of the stored FP32 residual, *rr_true = the same of the FP64
 * residual, *rw = r.w of the stored residual and the current w.  u is then
 * current (halo ring included). */
int tea_hip_mpcg_refresh(tea_hip_chunk* chunk, double alpha, int apply_u,
                         double* rr, double* rr_true, double* rw);

/* Copies one of the FP32 arrays (TEA_MP_*) to host memory. */
enum { TEA_MP_P = 0, TEA_MP_R = 1, TEA_MP_W = 2, TEA_MP_KX = 3, TEA_MP_KY = 4 };
int tea_hip_mp_read(tea_hip_chunk* chunk, int which, float* host,
                    size_t count);

/* ---- learned preconditioner network --------------------------------------
 * No counterpart in the reference.  A 6-layer 3x3 convolutional network,
 * 32 hidden channels, ReLU after the first five layers, inputs (residual,
 * kx feature, ky feature), one output channel; 38,177 FP32 master
 * parameters; BF16 weights and activations with FP32 accumulation on MFMA.
 * Parameter order: W1 (32,3,3,3), b1, W2..W5 (32,32,3,3) with b2..b5,
 * W6 (1,32,3,3), b6, each W indexed [out][in][ky][kx].  Inputs and outputs
 * are host arrays: input [batch][height][width][3], output and target
 * [batch][height][width].  width must be a multiple of 32; training needs
 * width <= 128. */
typedef struct tea_nn tea_nn;

int tea_nn_num_params(void);
int tea_nn_create(int device, int max_batch, int height, int width,
                  tea_nn** out);
int tea_nn_destroy(tea_nn* net);
/* Sets the FP32 master parameters (and clears the Adam state). */
int tea_nn_set_params(tea_nn* net, const float* params, size_t count);
int tea_nn_get_params(tea_nn* net, float* params, size_t count);
/* Network output for `batch` inputs. */
int tea_nn_forward(tea_nn* net, int batch, const float* input, float* output);
/* Mean squared error against `target` and its gradient with respect to the
 * master parameters (grad may be NULL). */
int tea_nn_loss_grad(tea_nn* net, int batch, const float* input,
                     const float* target, float* loss, float* grad);
/* tea_nn_loss_grad followed by one Adam step (beta 0.9/0.999, eps 1e-8). */
int tea_nn_train_step(tea_nn* net, int batch, const float* input,
                      const float* target, float lr, float* loss);
/* One Adam step (Kingma & Ba 2015, Algorithm 1, with bias correction:
 * beta 0.9/0.999, eps 1e-8 added to the square root of the corrected second
 * moment) on the master parameters with the host gradient `grad`
 * (tea_nn_num_params() values), then the BF16 copies are repacked.
 * tea_nn_train_step is tea_nn_loss_grad followed by this step. */
int tea_nn_adam_step(tea_nn* net, const float* grad, float lr);
/* The Adam moments m and v (count = tea_nn_num_params() each) and the
 * number of steps taken since the last tea_nn_set_params. */
int tea_nn_get_adam_state(tea_nn* net, float* m, float* v, size_t count,
                          long long* step);
/* Copies a training set of `count` samples to device memory. */
int tea_nn_load_dataset(tea_nn* net, int count, const float* inputs,
                        const float* targets);
/* One Adam pass over the resident samples order[0..count) in batches of
 * `batch`; *mean_loss is the mean batch loss. */
int tea_nn_train_epoch(tea_nn* net, const int* order, int count, int batch,
                       float lr, float* mean_loss);

/* ---- flexible CG with the learned preconditioner ------------------------
 * No counterpart in the reference.  FP32 p, r, w, z; FP64 u, dots, scalars;
 * FP64 true-residual refresh with tea_hip_mpcg_refresh.  Preconditioners:
 * Jacobi, z = r / diag(A); learned, z = r / diag(A) + sum over levels of
 * g P y, y = D^-1/2 sigma N(D^-1/2 R r / sigma, f), with R the block mean of
 * the level, P its piecewise-constant prolongation, T = P^T A P / c^2 the
 * Galerkin coarse operator, D its diagonal, sigma the local RMS of
 * D^-1/2 R r (13 x 13 window, at least 1e-3 of its RMS over the grid), N
 * the network, f the couplings of D^-1/2 T D^-1/2 to the left and bottom
 * neighbours normalised by feature_norm = {mean_x, std_x, mean_y, std_y},
 * and g = (y.Rr) / (y.T y).  An iteration: calc_w (p = z + beta p_old, w = A p,
 * p.w and p.r, previous alpha applied to u), alpha = p.r / p.w,
 * tea_hip_mpcg_calc_r, precondition (z = M(r), z.w), beta = -z.w / p.w. */

/* Allocates the FCG buffers; with params != NULL also loads the network
 * (tea_nn_num_params() values) on up to `levels` (1 or 2) coarse levels,
 * level l coarsening by coarsen * 4^l: a level is used while the interior
 * divides by its factor and its coarse width is a multiple of 32 (the first
 * level must be).  coarsen must be a power of two. */
int tea_hip_fcg_attach(tea_hip_chunk* chunk, const float* params,
                       size_t count, const float* feature_norm, int coarsen,
                       int levels);
/* The same two-level preconditioner with a non-learned coarse solve:
 * y_l = `sweeps` Jacobi sweeps on T_l y = R_l r from y = 0 (the iteration of
 * TeaLeaf/c_kernels/sycl/jacobi.cpp:85-91 on the coarse operator, FP64),
 * with the energy-optimal factors g.  sweeps = 6 is what the network is
 * trained to do; sweeps = 1 is y = D^-1 R r.  coarsen must be 8; level l
 * coarsens by 8 * 4^l while the interior divides (no width condition).
 * Every sum is taken in a fixed order: the results do not depend on the
 * launch and are reproducible bit for bit. */
int tea_hip_fcg_attach_sweeps(tea_hip_chunk* chunk, int sweeps, int coarsen,
                              int levels);
/* The number of coarse levels in use (0: Jacobi only). */
int tea_hip_fcg_levels(tea_hip_chunk* chunk, int* levels);
/* tea_hip_mpcg_init plus the coarse operators (and the network's input
 * features). */
int tea_hip_fcg_init(tea_hip_chunk* chunk, int coefficient, double rx,
                     double ry, double* rr, double* bb);
/* z = M(r) with the coarse correction (use_net != 0) or Jacobi alone;
 * *zw = z.w. */
int tea_hip_fcg_precondition(tea_hip_chunk* chunk, int use_net, double* zw);
/* tea_hip_mpcg_calc_r then tea_hip_fcg_precondition with one wait:
 * r -= alpha w, *rr = r.r, z = M(r), *zw = z.w. */
int tea_hip_fcg_calc_r_precondition(tea_hip_chunk* chunk, double alpha,
                                    int use_net, double* rr, double* zw);
/* If apply_u: u += alpha p_old.  p = z + beta p_old; w = A p;
 * *pw = p.w, *pr = p.r. */
int tea_hip_fcg_calc_w(tea_hip_chunk* chunk, double alpha, int apply_u,
                       double beta, double* pw, double* pr);
/* Copies z to host memory. */
int tea_hip_fcg_read_z(tea_hip_chunk* chunk, float* host, size_t count);

/* ---- execution mode and measurement ------------------------------------ */

/* Mode 1 (default): CG calls are fused.  tea_hip_cg_calc_ur updates r and
 * returns r.r at once but defers u += alpha p; tea_hip_cg_calc_p and the
 * depth-1 U/P halo update that follows it are deferred too, and the next
 * tea_hip_cg_calc_w applies all of them in the same pass that computes
 * w = A p (88 bytes per cell per iteration instead of 104).  Every other
 * entry point applies pending work first, so results are bitwise those of
 * mode 0.  Mode 0: one launch per reference kernel. */
int tea_hip_set_fused(tea_hip_chunk* chunk, int fused);

/* Applies every deferred update now (CG work, held Chebyshev or PPCG
 * steps, held halo updates) and waits for the stream. */
int tea_hip_flush(tea_hip_chunk* chunk);

/* Kernel timing with HIP events on the handle's stream (off by default). */
int tea_hip_timing_enable(tea_hip_chunk* chunk, int enable);
/* Launch count and summed device time of kernel `kernel_id`
 * (0 <= kernel_id < tea_hip_num_kernels()). */
int tea_hip_kernel_stats(tea_hip_chunk* chunk, int kernel_id,
                         long* launches, double* total_ms);
int tea_hip_timing_reset(tea_hip_chunk* chunk);
int tea_hip_num_kernels(void);
const char* tea_hip_kernel_name(int kernel_id);

/* Measured streaming ceiling (no counterpart in the reference): copies
 * `bytes` (a multiple of 16) between two device buffers `reps` times with
 * 16-byte nontemporal loads and stores and sets *gbs = read + write bytes
 * per second / 1e9 of the average copy, timed with HIP events.  Allocates
 * and frees its buffers. */
int tea_hip_copy_bandwidth(int device, size_t bytes, int reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif
