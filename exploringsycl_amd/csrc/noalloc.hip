// libteahip.so: the 2-norm of a field without any allocation (no
// counterpart in the reference), for callers that cannot allocate device
// memory: code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc.
//
// tea_hip_calculate_2norm_noalloc is tea_hip_calculate_2norm for a field
// the caller owns: the sum of squares over the interior with the
// reference's reduction tree (TeaLeaf/c_kernels/sycl/sycl_shared.hpp:35-79),
// bit for bit.  Its only working memory, the span partials of the tree, is
// the file-scope device array `scratch` below: each GPU holds its own copy
// until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): the address
// on each device comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is used.  Every call on one
// GPU shares its array, so calls that overlap on one GPU race; every span
// partial is written before it is read, so nothing needs clearing between
// calls.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tea_hip_impl.h"

namespace tea {
namespace {

// Span partials: one per 4096 cells, at most 16448 (8192^2 cells with a
// halo of 2 need 16400).
constexpr int kNoAllocMaxSpans = 16448;
constexpr int kMaxGroups = (kNoAllocMaxSpans + kWave - 1) / kWave;  // 257
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocMaxSpans];

struct SquareBody {
  const double* b;
  __device__ void operator()(int i, int, int, double (&out)[1]) const { out[0] = b[i] * b[i]; }
};

// The tree's passes over the len span partials, as the host finishes them
// after tea_hip_calculate_2norm (zero-padded groups of 64, each reduced by
// the reference's pairwise tree, until one value is left), in one block;
// the result goes to out[0].
__global__ __launch_bounds__(kBlock) void tree_passes_kernel(const double* __restrict__ part,
                                                             int len, double* __restrict__ out) {
  __shared__ double buf[2][kMaxGroups];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const double* src = part;
  int n = len, cur = 0;
  while (n != 1) {
    const int groups = (n + kWave - 1) / kWave;
    for (int g = wv; g < groups; g += kWaves) {
      const int i = g * kWave + lane;
      const double v = wave_tree(i < n ? src[i] : 0.0);
      if (lane == 0) buf[cur][g] = v;
    }
    __syncthreads();
    src = buf[cur];
    cur ^= 1;
    n = groups;
  }
  if (threadIdx.x == 0) out[0] = src[0];
}

// The device address of `scratch` on each device, from hipGetSymbolAddress.
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int scratch_on(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    TEA_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) {
      set_error("hipGetSymbolAddress returned no address for the scratch array");
      return TEA_HIP_EHIP;
    }
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return TEA_HIP_OK;
}

}  // namespace
}  // namespace tea

using namespace tea;

extern "C" {

int tea_hip_calculate_2norm_noalloc(int device, const double* field, int x, int y,
                                    int halo_depth, double* norm, void* stream) {
  TEA_CHECK_ARG(field != nullptr && norm != nullptr, "null field or norm");
  TEA_CHECK_ARG(device >= 0 && device < kMaxDevices, "device out of range");
  TEA_CHECK_ARG(halo_depth >= 1 && x > 2 * halo_depth && y > 2 * halo_depth,
                "x and y must exceed 2*halo_depth >= 2");
  const long n = (long)x * (long)y;
  const long nspan = (n + kSpan - 1) / kSpan;
  if (nspan > kNoAllocMaxSpans) {
    set_error("tea_hip_calculate_2norm_noalloc: " + std::to_string(n) +
              " cells exceed the scratch array (at most " +
              std::to_string((long)kNoAllocMaxSpans * kSpan) + " cells)");
    return TEA_HIP_EINVAL;
  }
  int prev = 0;
  TEA_HIP_TRY(hipGetDevice(&prev));
  TEA_HIP_TRY(hipSetDevice(device));
  double* part = nullptr;
  int status = scratch_on(device, &part);
  if (status == TEA_HIP_OK) {
    Geo g;
    g.x = x;
    g.y = y;
    g.h = halo_depth;
    g.n = (int)n;
    g.dx = FastDiv((uint32_t)x);
    const int single_group = n <= kWave ? 1 : 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL((span_kernel<1, SquareBody>), dim3((unsigned)nspan), dim3(kBlock), 0,
                       st, g, SquareBody{field}, part, (int)nspan, single_group);
    hipLaunchKernelGGL(tree_passes_kernel, dim3(1), dim3(kBlock), 0, st, part, (int)nspan,
                       norm);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("tea_hip_calculate_2norm_noalloc: ") + hipGetErrorString(e));
      status = TEA_HIP_EHIP;
    }
  }
  (void)hipSetDevice(prev);
  return status;
}

}  // extern "C"
