// libteahip.so: chunk handle, field access, timing and the final pass of
// every reduction.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tea_hip_impl.h"

namespace tea {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

const char* const kKernelNames[K_NUM] = {
    "set_chunk_data", "set_chunk_state", "store_energy", "local_halos",
    "field_summary",  "cg_init_u",       "cg_init_k",    "cg_init_others",
    "cg_calc_w",      "cg_calc_ur",      "cg_calc_p",    "cg_fused_w",
    "cg_fused_r",     "cg_apply_pending", "copy_u",      "calculate_residual",
    "calculate_2norm", "finalise",       "mpcg_convert", "mpcg_w",
    "mpcg_r",         "mpcg_apply_u",    "mpcg_refresh",  "fcg_w",
    "fcg_jacobi",     "fcg_restrict",    "fcg_net",       "fcg_prolong",
    "span_finish",    "jacobi_init",     "jacobi_copy_u", "jacobi_iterate",
    "jacobi_fused",   "cheby_init",      "cheby_iterate", "cheby_calc_u",
    "cheby_fused",    "ppcg_init",       "ppcg_calc_ur",  "ppcg_calc_sd",
    "ppcg_fused",     "ppcg_two_step",   "cheby_two_step", "fcg_sweeps",
    "cg_ra_finish"};

Timed::Timed(tea_hip_chunk* c_, int kid_, hipStream_t s)
    : c(c_), kid(kid_), stream(s ? s : c_->stream) {
  if (!c->timing || !c->time_kernel[kid]) return;
  if (c->event_pool.size() < 2) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      c->event_pool.push_back(e);
    }
  }
  start = c->event_pool.back();
  c->event_pool.pop_back();
  stop = c->event_pool.back();
  c->event_pool.pop_back();
  (void)hipEventRecord(start, stream);
}

Timed::~Timed() {
  if (!start) return;
  (void)hipEventRecord(stop, stream);
  c->inflight.push_back({kid, start, stop});
}

int resolve_timing(tea_hip_chunk* c) {
  for (auto& p : c->inflight) {
    float ms = 0.0f;
    TEA_HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
    c->launches[p.kid] += 1;
    c->total_ms[p.kid] += ms;
    c->event_pool.push_back(p.start);
    c->event_pool.push_back(p.stop);
  }
  c->inflight.clear();
  return TEA_HIP_OK;
}

int sync_chunk(tea_hip_chunk* c) {
  TEA_HIP_TRY(hipStreamSynchronize(c->stream));
  c->poll_len = 0;
  return resolve_timing(c);
}

int resolve_ready(tea_hip_chunk* c) {
  size_t k = 0;
  for (; k < c->inflight.size(); ++k) {
    const auto& p = c->inflight[k];
    const hipError_t e = hipEventQuery(p.stop);
    if (e == hipErrorNotReady) break;
    TEA_HIP_TRY(e);
    float ms = 0.0f;
    TEA_HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
    c->launches[p.kid] += 1;
    c->total_ms[p.kid] += ms;
    c->event_pool.push_back(p.start);
    c->event_pool.push_back(p.stop);
  }
  c->inflight.erase(c->inflight.begin(), c->inflight.begin() + k);
  return TEA_HIP_OK;
}

// A signalling NaN.  Every partial is a sum, and IEEE 754 arithmetic never
// returns a signalling NaN (a NaN operand comes out quiet), so no partial
// can equal it; the wait below would still end when the stream goes idle.
constexpr uint64_t kPollSentinel = 0x7ff4c0ffee5e4d1dull;

// Spins until none of the nr x len values at v (arrays stride apart) holds
// the sentinel.  The stream is queried only after about 2^22 spins
// (milliseconds, longer than a pass over an 8192^2 chunk) and then every
// 2^16 spins, a few queries per pass even over the largest chunks: a query
// every few hundred spins slowed the kernels it overlapped.  An idle stream
// with a sentinel left means a kernel wrote fewer values than were waited
// for: an error naming the value.
static int wait_values(tea_hip_chunk* c, const volatile uint64_t* v, size_t stride, int nr,
                       int len) {
  int q = 0, i = 0;
  for (uint64_t spins = 1;; ++spins) {
    while (q < nr) {
      while (i < len && v[q * stride + i] != kPollSentinel) ++i;
      if (i < len) break;
      ++q;
      i = 0;
    }
    if (q == nr) break;
    if (spins >= (1ull << 22) && (spins & 0xffff) == 0) {
      const hipError_t e = hipStreamQuery(c->stream);
      if (e == hipSuccess) {
        std::atomic_thread_fence(std::memory_order_acquire);
        for (int qq = 0; qq < nr; ++qq)
          for (int ii = 0; ii < len; ++ii)
            if (v[qq * stride + ii] == kPollSentinel) {
              set_error("reduction " + std::to_string(qq) + " left partial " +
                        std::to_string(ii) + " unwritten");
              return TEA_HIP_EINVAL;
            }
        break;
      }
      if (e != hipErrorNotReady) TEA_HIP_TRY(e);
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return resolve_ready(c);
}

// The host's view of the partials is coherent, mapped memory: the finisher's
// stores reach it while the kernel is still retiring, so the host can form
// the next scalar and launch the next kernel before the stream reports
// completion.
static int wait_partials(tea_hip_chunk* c) {
  if (c->poll_len == 0) return sync_chunk(c);
  const int len = c->poll_len, nr = c->poll_nr;
  c->poll_len = 0;
  return wait_values(c, reinterpret_cast<const volatile uint64_t*>(c->host_partials),
                     (size_t)c->nspan, nr, len);
}

void arm_host_values(double* host, int count) {
  uint64_t* v = reinterpret_cast<uint64_t*>(host);
  for (int i = 0; i < count; ++i) v[i] = kPollSentinel;
}

int wait_host_values(tea_hip_chunk* c, const double* host, int count) {
  return wait_values(c, reinterpret_cast<const volatile uint64_t*>(host), 0, 1, count);
}

// Final passes of the reference's reduceArray (TeaLeaf/c_kernels/sycl/
// sycl_shared.hpp:43-75) over one launch's level-2 partials, on the host:
// the kernels write one partial per 4096-cell span straight into pinned host
// memory, so after the stream is synchronised the remaining passes (e.g.
// 4105 -> 65 -> 2 -> 1 at 4096^2) are a few thousand host additions, with
// no extra launch.  Each pass cuts the array into zero-padded groups of 64
// and sums each group with the pairwise tree.
static double host_tree_passes(double* v, size_t len) {
  while (len != 1) {
    const size_t groups = (len + kWave - 1) / kWave;
    for (size_t g = 0; g < groups; ++g) {
      double a[kWave];
      for (int l = 0; l < kWave; ++l) {
        const size_t i = g * kWave + l;
        a[l] = i < len ? v[i] : 0.0;
      }
      for (int st = 1; st < kWave; st *= 2)
        for (int i = 0; i + st < kWave; i += 2 * st) a[i] = a[i] + a[i + st];
      v[g] = a[0];  // g <= 64g: earlier groups never read it again
    }
    len = groups;
  }
  return v[0];
}

// Level-2 partial of each span from the 32 wave sums of its 8 blocks (zero
// for blocks past the grid, as the reference pads its level-2 input with
// zeros), then the next tree pass.  A 512-thread block covers 64 spans, one
// group of that pass: thread t takes 512-cell block 512 b + t and its four
// contiguous wave sums (so the loads coalesce), forms (w0 + w1) + (w2 + w3),
// lanes 8j..8j+7 then combine the span's 8 blocks pairwise (steps s = 8, 16,
// 32 of the level-2 tree), and wave 0 runs the group's tree over the 64 span
// partials (zero past nspan).  With a single span there is no further pass
// and its partial is the sum.
constexpr int kFinishBlock = kSubPerSpan * kWave;  // 512 threads = 64 spans

// The host's final tree passes (host_tree_passes) over len <= 64^2 values
// in one wave; buf is LDS of 64 doubles.  Lane 0 returns the sum.
__device__ inline double wave_tree_passes(const double* part, int len, double* buf) {
  const int lane = threadIdx.x & (kWave - 1);
  if (len == 1) return part[0];
  const int groups = (len + kWave - 1) / kWave;
  for (int g = 0; g < groups; ++g) {
    const int i = g * kWave + lane;
    const double v = wave_tree(i < len ? part[i] : 0.0);
    if (lane == 0) buf[g] = v;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (groups == 1) return buf[0];
  return wave_tree(lane < groups ? buf[lane] : 0.0);
}

template <int NR, bool TAIL = false>
__global__ __launch_bounds__(kFinishBlock) void span_finish_kernel(
    const double* __restrict__ sub, int nblk, double* partials, int nspan,
    RaTail tail = RaTail{}) {
  __shared__ double lds[NR][kWave];
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int blk = blockIdx.x * kFinishBlock + t;
  const int span = blk / kSubPerSpan;
  // Every sum's loads first (a round per sum was waited for sum by sum).
  const bool in = blk < nblk && span < nspan;
  double w[NR][kWaves];
#pragma unroll
  for (int q = 0; q < NR; ++q)
#pragma unroll
    for (int k = 0; k < kWaves; ++k)
      w[q][k] = sub[(size_t)q * kWaves * nblk + (in ? kWaves * blk + k : 0)];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    if (!in) w[q][0] = w[q][1] = w[q][2] = w[q][3] = 0.0;
    double v = (w[q][0] + w[q][1]) + (w[q][2] + w[q][3]);
    v = v + __shfl_xor(v, 1);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 4);
    if ((t & (kSubPerSpan - 1)) == 0) lds[q][t / kSubPerSpan] = v;
  }
  __syncthreads();
  if (t >= kWave) return;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    if (nspan == 1) {
      if (blockIdx.x == 0 && lane == 0) partials[q * nspan] = lds[q][0];
      continue;
    }
    const double g3 = wave_tree(lds[q][lane]);
    if (lane == 0) partials[q * nspan + blockIdx.x] = g3;
  }
  if (!TAIL) return;
  // The last block to store its partial (device-scope release before the
  // count, acquire after it) sums them all.
  unsigned prev = 0;
  if (lane == 0) {
    __threadfence();
    prev = atomicAdd(tail.count, 1u);
  }
  prev = __shfl(prev, 0);
  if (prev != gridDim.x - 1) return;
  __threadfence();
  const int len = nspan == 1 ? 1 : (int)gridDim.x;
  const double sum = wave_tree_passes(partials, len, &lds[0][0]);
  if (lane != 0) return;
  const double r0 = tail.from_dev ? tail.ra[2] : tail.rro;
  double q;
  if (tail.mode == 0) {
    q = r0 / sum;
    tail.ra[0] = q;
  } else {
    q = sum / r0;
    tail.ra[1] = q;
    tail.ra[2] = sum;
  }
  tail.out[0] = sum;
  tail.out[1] = q;
  *tail.count = 0u;
}

int launch_span_finish(tea_hip_chunk* c, int nr, int* len, int slot,
                       bool poll) {
  TEA_CHECK_ARG(nr >= 1 && nr <= kSubMaxNr, "span_finish: bad reduction count");
  TEA_CHECK_ARG(slot >= 0 && slot + nr <= 4, "span_finish: bad partials slot");
  const dim3 grid((unsigned)((c->nspan + kWave - 1) / kWave));
  double* dst = c->partials + (size_t)slot * c->nspan;
  c->poll_len = 0;
  if (poll && c->poll && slot == 0) {
    const int n = c->nspan == 1 ? 1 : (c->nspan + kWave - 1) / kWave;
    uint64_t* v = reinterpret_cast<uint64_t*>(c->host_partials);
    for (int q = 0; q < nr; ++q)
      for (int i = 0; i < n; ++i) v[(size_t)q * c->nspan + i] = kPollSentinel;
    c->poll_nr = nr;
    c->poll_len = n;
  }
  {
    Timed t(c, K_SPAN_FINISH);
    if (nr == 1)
      hipLaunchKernelGGL(span_finish_kernel<1>, grid, dim3(kFinishBlock), 0,
                         c->stream, c->sub, c->nblk, dst, c->nspan);
    else if (nr == 2)
      hipLaunchKernelGGL(span_finish_kernel<2>, grid, dim3(kFinishBlock), 0,
                         c->stream, c->sub, c->nblk, dst, c->nspan);
    else
      hipLaunchKernelGGL(span_finish_kernel<3>, grid, dim3(kFinishBlock), 0,
                         c->stream, c->sub, c->nblk, dst, c->nspan);
  }
  TEA_HIP_TRY(hipGetLastError());
  *len = c->nspan == 1 ? 1 : (c->nspan + kWave - 1) / kWave;
  return TEA_HIP_OK;
}

int launch_span_finish_ra(tea_hip_chunk* c, double* dst, const RaTail& tail) {
  const int grid = (c->nspan + kWave - 1) / kWave;
  TEA_CHECK_ARG(grid <= kWave * kWave, "span_finish: too many partials for the run-ahead");
  {
    Timed t(c, K_CG_RA_FINISH);
    hipLaunchKernelGGL((span_finish_kernel<1, true>), dim3((unsigned)grid), dim3(kFinishBlock),
                       0, c->stream, c->sub, c->nblk, dst, c->nspan, tail);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// dst = src in 16-byte pieces, one per thread, nontemporal both ways: the
// streaming ceiling for tea_hip_copy_bandwidth.
__global__ __launch_bounds__(kBlock) void stream_copy_kernel(const double* __restrict__ src,
                                                             double* __restrict__ dst,
                                                             size_t n2) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n2) return;
  const dv2 v = __builtin_nontemporal_load(reinterpret_cast<const dv2*>(src) + i);
  __builtin_nontemporal_store(v, reinterpret_cast<dv2*>(dst) + i);
}

int finish_reduction_n(tea_hip_chunk* c, int nr, int len, double* out) {
  TEA_RETURN_IF(wait_partials(c));
  for (int q = 0; q < nr; ++q)
    out[q] = host_tree_passes(c->host_partials + (size_t)q * c->nspan,
                              (size_t)len);
  return TEA_HIP_OK;
}

int finish_reduction_lens(tea_hip_chunk* c, int nr, const int* lens, double* out) {
  TEA_RETURN_IF(sync_chunk(c));
  for (int q = 0; q < nr; ++q)
    out[q] = host_tree_passes(c->host_partials + (size_t)q * c->nspan,
                              (size_t)lens[q]);
  return TEA_HIP_OK;
}

int finish_reduction(tea_hip_chunk* c, int nr, double* out) {
  return finish_reduction_n(c, nr, c->nspan, out);
}

}  // namespace tea

using namespace tea;

#define TEA_CHECK_CHUNK(c) TEA_CHECK_ARG((c) != nullptr, "null chunk handle")

extern "C" {

const char* tea_hip_last_error(void) { return g_last_error.c_str(); }

int tea_hip_chunk_create(int device, int x, int y, int halo_depth,
                         tea_hip_chunk** out) {
  TEA_CHECK_ARG(out != nullptr, "null output handle");
  *out = nullptr;
  TEA_CHECK_ARG(halo_depth >= 1, "halo_depth must be >= 1");
  TEA_CHECK_ARG(x > 2 * halo_depth && y > 2 * halo_depth,
                "x and y must exceed 2*halo_depth");
  const long n = (long)x * (long)y;
  TEA_CHECK_ARG(n < (1L << 31) - kSpan, "chunk too large (x*y >= 2^31)");
  int ndev = 0;
  TEA_HIP_TRY(hipGetDeviceCount(&ndev));
  TEA_CHECK_ARG(device >= 0 && device < ndev, "no such HIP device");
  TEA_HIP_TRY(hipSetDevice(device));

  tea_hip_chunk* c = new tea_hip_chunk();
  c->device = device;
  c->geo.x = x;
  c->geo.y = y;
  c->geo.h = halo_depth;
  c->geo.n = (int)n;
  c->geo.dx = FastDiv((uint32_t)x);
  c->ngroup = (int)((n + kWave - 1) / kWave);
  c->nspan = (int)((n + kSpan - 1) / kSpan);
  if (const char* b = getenv("TEA_HIP_BATCH")) c->batch = atoi(b);
  c->pairs = (x % 2 == 0) && y >= 3;
  if (const char* pv = getenv("TEA_HIP_PAIRS")) c->pairs = c->pairs && atoi(pv) != 0;
  c->quads = c->pairs && x % 4 == 0;
  if (const char* qv = getenv("TEA_HIP_QUADS")) c->quads = c->quads && atoi(qv) != 0;
  if (const char* hv = getenv("TEA_HIP_HOLD_U")) c->hold_ok = atoi(hv) != 0;
  c->nblk = (int)((n + kPairBlock - 1) / kPairBlock);
  if (const char* pb = getenv("TEA_HIP_PAIR_BLOCKS")) c->pair_blocks = atoi(pb);
  for (int k = 0; k < K_NUM; ++k) c->time_kernel[k] = true;
  if (const char* only = getenv("TEA_HIP_STATS_ONLY")) {
    for (int k = 0; k < K_NUM; ++k) {
      const std::string name = kKernelNames[k];
      const std::string list = std::string(",") + only + ",";
      c->time_kernel[k] = list.find("," + name + ",") != std::string::npos;
    }
  }
  if (const char* pv = getenv("TEA_HIP_POLL")) c->poll = atoi(pv) != 0;
  if (const char* rv = getenv("TEA_HIP_RUN_AHEAD")) c->ra_enabled = atoi(rv) != 0;
  if (const char* sv = getenv("TEA_HIP_STACK")) c->stack = atoi(sv);
  if (const char* sp = getenv("TEA_HIP_SPIN"))
    if (atoi(sp) != 0) (void)hipSetDeviceFlags(hipDeviceScheduleSpin);

  auto fail = [&](int code) {
    tea_hip_chunk_destroy(c);
    return code;
  };

  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreateWithFlags failed");
    return fail(TEA_HIP_EHIP);
  }

  const size_t xy = (size_t)n;
  const size_t X = (size_t)x, Y = (size_t)y;
  for (int f = 0; f < TEA_NUM_FIELDS; ++f) c->field_len[f] = xy;
  c->field_len[TEA_FIELD_X_AREA] = (X + 1) * Y;
  c->field_len[TEA_FIELD_Y_AREA] = X * (Y + 1);
  c->field_len[TEA_FIELD_CELL_X] = X;
  c->field_len[TEA_FIELD_CELL_Y] = Y;
  c->field_len[TEA_FIELD_CELL_DX] = X;
  c->field_len[TEA_FIELD_CELL_DY] = Y;
  c->field_len[TEA_FIELD_VERTEX_DX] = X + 1;
  c->field_len[TEA_FIELD_VERTEX_DY] = Y + 1;
  c->field_len[TEA_FIELD_VERTEX_X] = X + 1;
  c->field_len[TEA_FIELD_VERTEX_Y] = Y + 1;
  c->field_len[TEA_FIELD_COMMS_BUFFER] = std::max(X, Y) * (size_t)halo_depth;

  for (int f = 0; f < TEA_NUM_FIELDS; ++f) {
    const size_t bytes = c->field_len[f] * sizeof(double);
    if (hipMalloc(&c->field[f], bytes) != hipSuccess) {
      c->field[f] = nullptr;
      set_error("hipMalloc of a field failed");
      return fail(TEA_HIP_ENOMEM);
    }
    if (hipMemsetAsync(c->field[f], 0, bytes, c->stream) != hipSuccess) {
      set_error("hipMemsetAsync failed");
      return fail(TEA_HIP_EHIP);
    }
  }
  if (hipMalloc(&c->p_alt, xy * sizeof(double)) != hipSuccess) {
    c->p_alt = nullptr;
    set_error("hipMalloc of the second p buffer failed");
    return fail(TEA_HIP_ENOMEM);
  }
  if (hipMalloc(&c->sub, (size_t)kSubMaxNr * kWaves * c->nblk * sizeof(double)) != hipSuccess) {
    c->sub = nullptr;
    set_error("hipMalloc of the subtree sums failed");
    return fail(TEA_HIP_ENOMEM);
  }
  const size_t pbytes = 4 * (size_t)c->nspan * sizeof(double);
  if (hipHostMalloc(&c->host_partials, pbytes,
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    c->host_partials = nullptr;
    set_error("hipHostMalloc of the reduction partials failed");
    return fail(TEA_HIP_ENOMEM);
  }
  std::memset(c->host_partials, 0, pbytes);
  if (hipHostGetDevicePointer((void**)&c->partials, c->host_partials, 0) !=
      hipSuccess) {
    set_error("hipHostGetDevicePointer failed");
    return fail(TEA_HIP_EHIP);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    set_error("hipStreamSynchronize failed");
    return fail(TEA_HIP_EHIP);
  }
  *out = c;
  return TEA_HIP_OK;
}

int tea_hip_chunk_destroy(tea_hip_chunk* c) {
  if (!c) return TEA_HIP_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->inflight) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (int f = 0; f < TEA_NUM_FIELDS; ++f)
    if (c->field[f]) (void)hipFree(c->field[f]);
  if (c->p_alt) (void)hipFree(c->p_alt);
  for (double* b : {c->r_alt, c->u_alt, c->w_alt, c->ra_dev})
    if (b) (void)hipFree(b);
  if (c->ra_host) (void)hipHostFree(c->ra_host);
  if (c->spare) (void)hipFree(c->spare);
  if (c->spare2) (void)hipFree(c->spare2);
  if (c->sub) (void)hipFree(c->sub);
  for (float* b : {c->mp_p[0], c->mp_p[1], c->mp_p[2], c->mp_r, c->mp_w, c->mp_kx, c->mp_ky,
                   c->fcg_z})
    if (b) (void)hipFree(b);
  if (c->fcg_xsum) (void)hipFree(c->fcg_xsum);
  if (c->fcg_g) (void)hipFree(c->fcg_g);
  if (c->fcg_side) {
    (void)hipStreamSynchronize(c->fcg_side);
    (void)hipStreamDestroy(c->fcg_side);
  }
  if (c->fcg_fork) (void)hipEventDestroy(c->fcg_fork);
  if (c->fcg_join) (void)hipEventDestroy(c->fcg_join);
  for (FcgLevel& L : c->fcg_lv) {
    for (float* b : {L.feat, L.rbar, L.kc, L.dis, L.sig, L.y})
      if (b) (void)hipFree(b);
    if (L.rms) (void)hipFree(L.rms);
    if (L.tpart) (void)hipFree(L.tpart);
    if (L.net) tea_nn_destroy(L.net);
  }
  if (c->host_partials) (void)hipHostFree(c->host_partials);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TEA_HIP_OK;
}

int tea_hip_field(tea_hip_chunk* c, int field, double** dev_ptr,
                  size_t* count) {
  TEA_CHECK_CHUNK(c);
  TEA_CHECK_ARG(field >= 0 && field < TEA_NUM_FIELDS, "bad field id");
  if (dev_ptr) *dev_ptr = c->field[field];
  if (count) *count = c->field_len[field];
  return TEA_HIP_OK;
}

int tea_hip_field_write(tea_hip_chunk* c, int field, const double* host,
                        size_t count) {
  TEA_CHECK_CHUNK(c);
  TEA_CHECK_ARG(field >= 0 && field < TEA_NUM_FIELDS, "bad field id");
  TEA_CHECK_ARG(count <= c->field_len[field], "count exceeds field length");
  TEA_CHECK_ARG(host != nullptr || count == 0, "null host buffer");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_RETURN_IF(apply_pending(c));
  TEA_HIP_TRY(hipMemcpyAsync(c->field[field], host, count * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  if (field == TEA_FIELD_P) c->p_alt_synced = false;
  c->faces_ok_u = c->faces_ok_p = false;
  return sync_chunk(c);
}

int tea_hip_field_read(tea_hip_chunk* c, int field, double* host,
                       size_t count) {
  TEA_CHECK_CHUNK(c);
  TEA_CHECK_ARG(field >= 0 && field < TEA_NUM_FIELDS, "bad field id");
  TEA_CHECK_ARG(count <= c->field_len[field], "count exceeds field length");
  TEA_CHECK_ARG(host != nullptr || count == 0, "null host buffer");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_RETURN_IF(apply_pending(c));
  TEA_HIP_TRY(hipMemcpyAsync(host, c->field[field], count * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  return sync_chunk(c);
}

int tea_hip_synchronize(tea_hip_chunk* c) {
  TEA_CHECK_CHUNK(c);
  TEA_HIP_TRY(hipSetDevice(c->device));
  return sync_chunk(c);
}

int tea_hip_set_fused(tea_hip_chunk* c, int fused) {
  TEA_CHECK_CHUNK(c);
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_RETURN_IF(apply_pending(c));
  c->fused = fused ? 1 : 0;
  return TEA_HIP_OK;
}

int tea_hip_flush(tea_hip_chunk* c) {
  TEA_CHECK_CHUNK(c);
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_RETURN_IF(apply_pending(c));
  return sync_chunk(c);
}

int tea_hip_timing_enable(tea_hip_chunk* c, int enable) {
  TEA_CHECK_CHUNK(c);
  TEA_RETURN_IF(sync_chunk(c));
  c->timing = enable != 0;
  return TEA_HIP_OK;
}

int tea_hip_kernel_stats(tea_hip_chunk* c, int kernel_id, long* launches,
                         double* total_ms) {
  TEA_CHECK_CHUNK(c);
  TEA_CHECK_ARG(kernel_id >= 0 && kernel_id < K_NUM, "bad kernel id");
  TEA_RETURN_IF(sync_chunk(c));
  if (launches) *launches = c->launches[kernel_id];
  if (total_ms) *total_ms = c->total_ms[kernel_id];
  return TEA_HIP_OK;
}

int tea_hip_timing_reset(tea_hip_chunk* c) {
  TEA_CHECK_CHUNK(c);
  TEA_RETURN_IF(sync_chunk(c));
  for (int k = 0; k < K_NUM; ++k) {
    c->launches[k] = 0;
    c->total_ms[k] = 0.0;
  }
  return TEA_HIP_OK;
}

int tea_hip_copy_bandwidth(int device, size_t bytes, int reps, double* gbs) {
  TEA_CHECK_ARG(gbs != nullptr, "null output");
  TEA_CHECK_ARG(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
  TEA_CHECK_ARG(reps >= 1, "reps must be positive");
  TEA_HIP_TRY(hipSetDevice(device));
  const size_t n2 = bytes / 16;
  double *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess) {
    set_error("hipMalloc of the copy source failed");
    return TEA_HIP_ENOMEM;
  }
  if (hipMalloc(&dst, bytes) != hipSuccess) {
    (void)hipFree(src);
    set_error("hipMalloc of the copy destination failed");
    return TEA_HIP_ENOMEM;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int status = TEA_HIP_OK;
  float ms = 0.0f;
  const dim3 grid((unsigned)((n2 + kBlock - 1) / kBlock));
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(src, 0, bytes, st) != hipSuccess) {
    set_error("copy bandwidth: stream, event or memset failed");
    status = TEA_HIP_EHIP;
  } else {
    for (int k = 0; k < 2; ++k)  // warm-up
      hipLaunchKernelGGL(stream_copy_kernel, grid, dim3(kBlock), 0, st, src, dst, n2);
    (void)hipEventRecord(e0, st);
    for (int k = 0; k < reps; ++k)
      hipLaunchKernelGGL(stream_copy_kernel, grid, dim3(kBlock), 0, st, src, dst, n2);
    (void)hipEventRecord(e1, st);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
      set_error("copy bandwidth: launch or timing failed");
      status = TEA_HIP_EHIP;
    } else {
      *gbs = 2.0 * (double)bytes / ((double)ms * 1e-3 / reps) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(src);
  (void)hipFree(dst);
  return status;
}

int tea_hip_num_kernels(void) { return K_NUM; }

const char* tea_hip_kernel_name(int kernel_id) {
  if (kernel_id < 0 || kernel_id >= K_NUM) return "";
  return kKernelNames[kernel_id];
}

}  // extern "C"
