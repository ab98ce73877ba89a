// libteahip.so: mesh/state initialisation, reflective halos and the field
// summary.

#include <hip/hip_runtime.h>

#include "tea_hip_impl.h"

namespace tea {
namespace {

// ---- mesh (TeaLeaf/c_kernels/sycl/set_chunk_data.cpp:8-64) ---------------

__global__ __launch_bounds__(kBlock) void set_vertices_kernel(
    int x, int y, int h, double x_min, double y_min, double dx, double dy,
    double* __restrict__ vertex_x, double* __restrict__ vertex_y,
    double* __restrict__ cell_x, double* __restrict__ cell_y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < x + 1) vertex_x[i] = x_min + dx * ((double)i - (double)h);
  if (i < y + 1) vertex_y[i] = y_min + dy * ((double)i - (double)h);
}

// Cell centres need both neighbouring vertices, so they follow in a second
// launch (the reference also orders the two kernels on its queue).
__global__ __launch_bounds__(kBlock) void set_cells_kernel(
    Geo g, double dx, double dy, const double* __restrict__ vertex_x,
    const double* __restrict__ vertex_y, double* __restrict__ cell_x,
    double* __restrict__ cell_y, double* __restrict__ volume,
    double* __restrict__ x_area, double* __restrict__ y_area) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < g.x) cell_x[i] = 0.5 * (vertex_x[i] + vertex_x[i + 1]);
  if (i < g.y) cell_y[i] = 0.5 * (vertex_y[i] + vertex_y[i + 1]);
  if (i < g.n) {
    volume[i] = dx * dy;
    x_area[i] = dy;
    y_area[i] = dx;
  }
}

// ---- states (TeaLeaf/c_kernels/sycl/set_chunk_state.cpp:8-92) ------------

__global__ __launch_bounds__(kBlock) void initial_state_kernel(
    int n, double energy, double density, double* __restrict__ energy0,
    double* __restrict__ density_f) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    energy0[i] = energy;
    density_f[i] = density;
  }
}

struct StateArgs {
  int geometry;
  double x_min, y_min, x_max, y_max, radius, density, energy;
};

__global__ __launch_bounds__(kBlock) void state_kernel(
    Geo g, StateArgs s, double* __restrict__ energy0,
    double* __restrict__ density, double* __restrict__ u,
    const double* __restrict__ cell_x, const double* __restrict__ cell_y,
    const double* __restrict__ vertex_x, const double* __restrict__ vertex_y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= g.n) return;
  const int jj = (int)g.dx.div((uint32_t)i);
  const int kk = i - jj * g.x;
  bool apply = false;
  if (s.geometry == TEA_GEOMETRY_RECTANGULAR) {
    apply = vertex_x[kk + 1] >= s.x_min && vertex_x[kk] < s.x_max &&
            vertex_y[jj + 1] >= s.y_min && vertex_y[jj] < s.y_max;
  } else if (s.geometry == TEA_GEOMETRY_CIRCULAR) {
    const double ddx = cell_x[kk] - s.x_min;
    const double ddy = cell_y[jj] - s.y_min;
    apply = sqrt(ddx * ddx + ddy * ddy) <= s.radius;
  } else if (s.geometry == TEA_GEOMETRY_POINT) {
    apply = vertex_x[kk] == s.x_min && vertex_y[jj] == s.y_min;
  }
  double e0 = energy0[i], rho = density[i];
  if (apply) {
    e0 = s.energy;
    rho = s.density;
    energy0[i] = e0;
    density[i] = rho;
  }
  if (kk > 0 && kk < g.x - 1 && jj > 0 && jj < g.y - 1) u[i] = e0 * rho;
}

__global__ __launch_bounds__(kBlock) void copy_kernel(
    int n, double* __restrict__ dst, const double* __restrict__ src) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---- halos (TeaLeaf/c_kernels/sycl/local_halos.cpp:7-102) ----------------
//
// The reference runs left, right, top, bottom as four launches per field, so
// top/bottom copy the corner columns that left/right have just written.  One
// launch gives the same final values: top/bottom threads read their source
// cell through the left/right mirror (the value left/right would have put
// there), and left/right threads skip the rows that top/bottom overwrite.
// No thread then reads a cell another thread writes.  Requires
// depth <= halo_depth and interior extent >= depth (checked by the caller).

struct HaloFields {
  double* f[TEA_HALO_NUM];
};

__global__ __launch_bounds__(kBlock) void halo_kernel(Geo g, int depth,
                                                     int ext_l, int ext_r,
                                                     int ext_b, int ext_t,
                                                     HaloFields hf) {
  double* __restrict__ b = hf.f[blockIdx.y];
  const int x = g.x, y = g.y, h = g.h;
  const int lr = y * depth;  // items per left/right face
  const int tb = x * depth;  // items per top/bottom face
  int item = blockIdx.x * kBlock + threadIdx.x;
  if (item < 2 * lr) {
    const bool right = item >= lr;
    if (right ? !ext_r : !ext_l) return;
    if (right) item -= lr;
    const int jj = item / depth, k = item - (item / depth) * depth;
    if (ext_t && jj >= y - h && jj < y - h + depth) return;
    if (ext_b && jj >= h - depth && jj < h) return;
    const long row = (long)jj * x;
    if (!right)
      b[row + (h - 1 - k)] = b[row + (h + k)];
    else
      b[row + (x - h + k)] = b[row + (x - h - 1 - k)];
    return;
  }
  item -= 2 * lr;
  if (item >= 2 * tb) return;
  const bool top = item >= tb;
  if (top ? !ext_t : !ext_b) return;
  if (top) item -= tb;
  const int k = item / x, kk = item - (item / x) * x;
  int src_col = kk;
  if (ext_l && kk >= h - depth && kk < h) src_col = 2 * h - 1 - kk;
  if (ext_r && kk >= x - h && kk < x - h + depth) src_col = 2 * (x - h) - 1 - kk;
  if (top)
    b[(long)(y - h + k) * x + kk] = b[(long)(y - h - 1 - k) * x + src_col];
  else
    b[(long)(h - 1 - k) * x + kk] = b[(long)(h + k) * x + src_col];
}

// ---- remote-halo strips (TeaLeaf/c_kernels/sycl/pack_halos.cpp:7-184) -----
//
// Item i of a face's strip of n = depth * (y or x) items: the left/right
// strips run row by row, depth cells each, the bottom/top strips are depth
// whole rows.  Packing reads the cells next to the interior, unpacking
// writes the halo cells, exactly at the reference's offsets.
__global__ __launch_bounds__(kBlock) void strip_kernel(Geo g, int depth,
                                                      int face, int pack,
                                                      double* __restrict__ field,
                                                      double* __restrict__ strip) {
  const int x = g.x, y = g.y, h = g.h;
  const long n = (face == TEA_FACE_LEFT || face == TEA_FACE_RIGHT)
                     ? (long)y * depth : (long)x * depth;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const long lines = i / depth;
  long f;
  if (face == TEA_FACE_LEFT)
    f = (pack ? h : h - depth) + lines * (x - depth) + i;
  else if (face == TEA_FACE_RIGHT)
    f = (pack ? x - h - depth : x - h) + lines * (x - depth) + i;
  else if (face == TEA_FACE_BOTTOM)
    f = (long)x * (pack ? h : h - depth) + i;
  else
    f = (long)x * (pack ? y - h - depth : y - h) + i;
  if (pack)
    strip[i] = field[f];
  else
    field[f] = strip[i];
}

// ---- field summary (TeaLeaf/c_kernels/sycl/field_summary.cpp:6-151) ------

struct SummaryBody {
  const double* __restrict__ volume;
  const double* __restrict__ density;
  const double* __restrict__ energy0;
  const double* __restrict__ u;
  __device__ void operator()(int i, int, int, double (&out)[4]) const {
    const double cell_vol = volume[i];
    const double cell_mass = cell_vol * density[i];
    out[0] = cell_vol;
    out[1] = cell_mass;
    out[2] = cell_mass * energy0[i];
    out[3] = cell_mass * u[i];
  }
};

}  // namespace
}  // namespace tea

using namespace tea;

extern "C" {

int tea_hip_set_chunk_data(tea_hip_chunk* c, double x_min, double y_min,
                           double dx, double dy) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(apply_pending(c));
  c->faces_ok_u = c->faces_ok_p = false;
  const Geo g = c->geo;
  double** f = c->field;
  {
    Timed t(c, K_SET_CHUNK_DATA);
    const int nv = (g.x > g.y ? g.x : g.y) + 1;
    hipLaunchKernelGGL(set_vertices_kernel, flat_grid(nv), dim3(kBlock), 0,
                       c->stream, g.x, g.y, g.h, x_min, y_min, dx, dy,
                       f[TEA_FIELD_VERTEX_X], f[TEA_FIELD_VERTEX_Y],
                       f[TEA_FIELD_CELL_X], f[TEA_FIELD_CELL_Y]);
    hipLaunchKernelGGL(set_cells_kernel, flat_grid(g.n), dim3(kBlock), 0,
                       c->stream, g, dx, dy, f[TEA_FIELD_VERTEX_X],
                       f[TEA_FIELD_VERTEX_Y], f[TEA_FIELD_CELL_X],
                       f[TEA_FIELD_CELL_Y], f[TEA_FIELD_VOLUME],
                       f[TEA_FIELD_X_AREA], f[TEA_FIELD_Y_AREA]);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

int tea_hip_set_chunk_state(tea_hip_chunk* c, int num_states,
                            const double* density, const double* energy,
                            const double* x_min, const double* y_min,
                            const double* x_max, const double* y_max,
                            const double* radius, const int* geometry) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(num_states >= 1, "need at least one state");
  TEA_CHECK_ARG(density && energy, "null state arrays");
  TEA_CHECK_ARG(num_states == 1 || (x_min && y_min && x_max && y_max &&
                                    radius && geometry),
                "null state geometry arrays");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(apply_pending(c));
  c->faces_ok_u = c->faces_ok_p = false;
  const Geo g = c->geo;
  double** f = c->field;
  {
    Timed t(c, K_SET_CHUNK_STATE);
    hipLaunchKernelGGL(initial_state_kernel, flat_grid(g.n), dim3(kBlock), 0,
                       c->stream, g.n, energy[0], density[0],
                       f[TEA_FIELD_ENERGY0], f[TEA_FIELD_DENSITY]);
    for (int s = 1; s < num_states; ++s) {
      StateArgs a{geometry[s], x_min[s],  y_min[s],   x_max[s],
                  y_max[s],    radius[s], density[s], energy[s]};
      hipLaunchKernelGGL(state_kernel, flat_grid(g.n), dim3(kBlock), 0,
                         c->stream, g, a, f[TEA_FIELD_ENERGY0],
                         f[TEA_FIELD_DENSITY], f[TEA_FIELD_U],
                         f[TEA_FIELD_CELL_X], f[TEA_FIELD_CELL_Y],
                         f[TEA_FIELD_VERTEX_X], f[TEA_FIELD_VERTEX_Y]);
    }
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

int tea_hip_store_energy(tea_hip_chunk* c) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(apply_pending(c));
  {
    Timed t(c, K_STORE_ENERGY);
    hipLaunchKernelGGL(copy_kernel, flat_grid(c->geo.n), dim3(kBlock), 0,
                       c->stream, c->geo.n, c->field[TEA_FIELD_ENERGY],
                       c->field[TEA_FIELD_ENERGY0]);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

}  // extern "C"

namespace tea {

// Launches one reflective-halo pass; used by tea_hip_local_halos and by the
// flush of deferred CG halos.
int launch_halos(tea_hip_chunk* c, int depth, const int neighbours[4],
                 const int fields[6]) {
  const Geo g = c->geo;
  TEA_CHECK_ARG(depth >= 1 && depth <= g.h, "halo depth out of range");
  TEA_CHECK_ARG(g.x - 2 * g.h >= depth && g.y - 2 * g.h >= depth,
                "interior smaller than the halo depth");
  static const int slot_field[TEA_HALO_NUM] = {
      TEA_FIELD_DENSITY, TEA_FIELD_ENERGY0, TEA_FIELD_ENERGY,
      TEA_FIELD_U,       TEA_FIELD_P,       TEA_FIELD_SD};
  HaloFields hf{};
  int nf = 0;
  for (int s = 0; s < TEA_HALO_NUM; ++s)
    if (fields[s]) hf.f[nf++] = c->field[slot_field[s]];
  const int ext_l = neighbours[TEA_FACE_LEFT] < 0;
  const int ext_r = neighbours[TEA_FACE_RIGHT] < 0;
  const int ext_b = neighbours[TEA_FACE_BOTTOM] < 0;
  const int ext_t = neighbours[TEA_FACE_TOP] < 0;
  if (nf == 0 || !(ext_l || ext_r || ext_b || ext_t)) return TEA_HIP_OK;
  if (ext_l && ext_r && ext_b && ext_t) {
    c->faces_ok_u |= fields[TEA_HALO_U] != 0;
    c->faces_ok_p |= fields[TEA_HALO_P] != 0;
  }
  const long items = 2L * g.y * depth + 2L * g.x * depth;
  {
    Timed t(c, K_LOCAL_HALOS);
    hipLaunchKernelGGL(halo_kernel,
                       dim3((unsigned)((items + kBlock - 1) / kBlock), nf),
                       dim3(kBlock), 0, c->stream, g, depth, ext_l, ext_r,
                       ext_b, ext_t, hf);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

}  // namespace tea

extern "C" {

int tea_hip_local_halos(tea_hip_chunk* c, int depth, const int neighbours[4],
                        const int fields_to_exchange[6]) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(neighbours && fields_to_exchange, "null arrays");
  TEA_HIP_TRY(hipSetDevice(c->device));
  bool all_external = true;
  for (int f = 0; f < 4; ++f) all_external &= neighbours[f] < 0;
  // The depth-1 SD (U) update between two PPCG inner steps (Chebyshev
  // steps) is folded into the pass that runs both steps; the held-back
  // first step is flushed instead if anything else comes between them.
  if (c->held && !c->held_halo && depth == 1 && all_external) {
    const int slot = c->held == kHeldPpcg ? TEA_HALO_SD : TEA_HALO_U;
    bool only_slot = true;
    for (int s = 0; s < TEA_HALO_NUM; ++s)
      only_slot &= (fields_to_exchange[s] != 0) == (s == slot);
    if (only_slot) {
      c->held_halo = true;
      return TEA_HIP_OK;
    }
  }
  // A depth-1 SD (U) update with no step held is held back too: the next
  // PPCG inner step (Chebyshev step) is then held, and the pass that runs
  // it with the one after reflects the ring of the field it stages.  Any
  // other entry point runs it first.
  if (c->fused && c->pairs && !c->held && c->pre_halo < 0 && depth == 1 &&
      all_external) {
    int only = -1, count = 0;
    for (int s = 0; s < TEA_HALO_NUM; ++s)
      if (fields_to_exchange[s]) {
        only = s;
        ++count;
      }
    if (count == 1 && (only == TEA_HALO_SD || only == TEA_HALO_U) &&
        !(only == TEA_HALO_U && (c->pend_u || c->pend_halo))) {
      c->pre_halo = only;
      return TEA_HIP_OK;
    }
  }
  // The depth-1 update that follows cg_calc_p is folded into the next fused
  // cg_calc_w when all faces are external: of U and P with u += alpha p
  // pending too (CG), of P alone without (the PPCG outer step).
  if (c->fused && !c->held && c->pend_p && !c->pend_halo && depth == 1 &&
      all_external) {
    bool only_up = true, only_p = true;
    for (int s = 0; s < TEA_HALO_NUM; ++s) {
      const bool on = fields_to_exchange[s] != 0;
      only_up &= on == (s == TEA_HALO_U || s == TEA_HALO_P);
      only_p &= on == (s == TEA_HALO_P);
    }
    if (c->pend_u ? (only_up && c->faces_ok_u && c->faces_ok_p)
                  : (only_p && c->faces_ok_p)) {
      c->pend_halo = true;
      return TEA_HIP_OK;
    }
  }
  // An update of fields other than U and P leaves deferred CG work on u and
  // p deferred (a held step, which it may read, is run first).
  if (fields_to_exchange[TEA_HALO_U] || fields_to_exchange[TEA_HALO_P] ||
      c->held || c->pre_halo >= 0)
    TEA_RETURN_IF(apply_pending(c));
  return launch_halos(c, depth, neighbours, fields_to_exchange);
}

int tea_hip_pack_or_unpack(tea_hip_chunk* c, int depth, int face, int pack,
                           const double* field, double* host_buffer) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(depth >= 1 && depth <= c->geo.h, "depth must be 1..halo_depth");
  TEA_CHECK_ARG(face >= TEA_FACE_LEFT && face <= TEA_FACE_TOP, "bad face");
  TEA_HIP_TRY(hipSetDevice(c->device));
  int fid = -1;
  for (int f = 0; f < TEA_NUM_FIELDS; ++f)
    if (field == c->field[f] && c->field_len[f] >= (size_t)c->geo.n) fid = f;
  TEA_CHECK_ARG(fid >= 0, "field is not an x*y field of this chunk");
  TEA_RETURN_IF(apply_pending(c));  // may move the field to its second buffer
  const long n = (face == TEA_FACE_LEFT || face == TEA_FACE_RIGHT)
                     ? (long)c->geo.y * depth : (long)c->geo.x * depth;
  TEA_CHECK_ARG(host_buffer != nullptr || n == 0, "null host buffer");
  TEA_CHECK_ARG((size_t)n <= c->field_len[TEA_FIELD_COMMS_BUFFER],
                "strip longer than the comms buffer");
  double* strip = c->field[TEA_FIELD_COMMS_BUFFER];
  const size_t bytes = (size_t)n * sizeof(double);
  if (!pack)
    TEA_HIP_TRY(hipMemcpyAsync(strip, host_buffer, bytes, hipMemcpyHostToDevice,
                               c->stream));
  {
    Timed t(c, K_LOCAL_HALOS);
    hipLaunchKernelGGL(strip_kernel, flat_grid(n), dim3(kBlock), 0, c->stream,
                       c->geo, depth, face, pack, c->field[fid], strip);
  }
  TEA_HIP_TRY(hipGetLastError());
  if (pack)
    TEA_HIP_TRY(hipMemcpyAsync(host_buffer, strip, bytes, hipMemcpyDeviceToHost,
                               c->stream));
  if (!pack && fid == TEA_FIELD_P) c->p_alt_synced = false;
  if (!pack) c->faces_ok_u = c->faces_ok_p = false;
  return sync_chunk(c);
}

int tea_hip_field_summary(tea_hip_chunk* c, double* vol, double* mass,
                          double* ie, double* temp) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(vol && mass && ie && temp, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  SummaryBody body{f[TEA_FIELD_VOLUME], f[TEA_FIELD_DENSITY],
                   f[TEA_FIELD_ENERGY0], f[TEA_FIELD_U]};
  {
    Timed t(c, K_FIELD_SUMMARY);
    hipLaunchKernelGGL((span_kernel<4, SummaryBody>), span_grid(c),
                       dim3(kBlock), 0, c->stream, c->geo, body, c->partials,
                       c->nspan, c->ngroup == 1 ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  double out[4];
  TEA_RETURN_IF(finish_reduction(c, 4, out));
  *vol = out[0];
  *mass = out[1];
  *ie = out[2];
  *temp = out[3];
  return TEA_HIP_OK;
}

}  // extern "C"
