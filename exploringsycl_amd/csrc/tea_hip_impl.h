// Internal declarations of libteahip.so: the chunk handle, launch helpers
// and the device-side reduction machinery shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "tea_hip.h"

namespace tea {

// ---------------------------------------------------------------------------
// Errors

void set_error(const std::string& msg);

#define TEA_HIP_TRY(call)                                                     \
  do {                                                                        \
    hipError_t err_ = (call);                                                 \
    if (err_ != hipSuccess) {                                                 \
      ::tea::set_error(std::string(#call) + " failed: " +                     \
                       hipGetErrorString(err_) + " (" + __FILE__ + ":" +      \
                       std::to_string(__LINE__) + ")");                       \
      return TEA_HIP_EHIP;                                                    \
    }                                                                         \
  } while (0)

#define TEA_CHECK_ARG(cond, msg)                                              \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::tea::set_error(std::string(__func__) + ": " + (msg));                 \
      return TEA_HIP_EINVAL;                                                  \
    }                                                                         \
  } while (0)

#define TEA_RETURN_IF(status)                                                 \
  do {                                                                        \
    int st_ = (status);                                                       \
    if (st_ != TEA_HIP_OK) return st_;                                        \
  } while (0)

// ---------------------------------------------------------------------------
// Geometry of a chunk and fast division by its row length.

// Division by a runtime constant d (1 <= d < 2^31) for n < 2^31 with one
// multiply-high: q = (umulhi(n, m) + n) >> s, s = ceil(log2 d),
// m = floor(2^32 (2^s - d) / d) + 1 (Granlund & Montgomery 1994, sec. 4).
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t divisor) : d(divisor) {
    s = 0;
    while ((uint64_t(1) << s) < d) ++s;
    m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t t = __umulhi(n, m);
#else
    const uint32_t t = uint32_t((uint64_t(n) * m) >> 32);
#endif
    return uint32_t((uint64_t(t) + n) >> s);
  }
};

struct Geo {
  int x, y, h;
  int n;      // x*y
  FastDiv dx; // division by x
};

__device__ inline bool interior(const Geo& g, int kk, int jj) {
  return kk >= g.h && kk < g.x - g.h && jj >= g.h && jj < g.y - g.h;
}

// ---------------------------------------------------------------------------
// Launch shape.  A "span" is 4096 consecutive linear cells = 64 groups of 64:
// exactly one level-2 partial of the reference's reduction tree.  A span is
// processed by one 256-thread block (4 wave64s), 16 cells per thread, with
// cell index base + k*256 + threadIdx.x, so wave w's k-th cell set is group
// 4k+w of the span.

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kItems = 16;
constexpr int kSpan = kBlock * kItems;  // 4096
constexpr int kNumXcd = 8;              // MI355X: 8 XCDs, one L2 each

// Blocks are dispatched round-robin over the XCDs (block b -> XCD b % 8).
// Renumber them so each XCD walks one contiguous range of spans: the rows a
// stencil reads above and below were then fetched by the same XCD, into the
// same L2, moments earlier.
__device__ inline int xcd_span(int b, int nb) {
  const int xcd = b % kNumXcd, idx = b / kNumXcd;
  const int per = nb / kNumXcd, rem = nb % kNumXcd;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Cross-lane moves on the VALU's data-parallel-primitive (DPP) path instead
// of LDS permutes.  CTRL is a DPP control: 0x130 wave_shl:1 (lane l takes
// lane l+1), 0x138 wave_shr:1 (lane l takes lane l-1), 0xB1 and 0x4E
// quad_perm [1,0,3,2] and [2,3,0,1] (lane l takes lane l^1, l^2), 0x141
// row_half_mirror (lane i of each 8 takes lane 7-i), 0x140 row_mirror (lane
// i of each 16 takes lane 15-i).  A lane without a source keeps its value.
template <int CTRL>
__device__ inline float dpp(float v) {
  const int b = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false));
}
template <class T>
__device__ inline T from_next_lane(T v) { return dpp<0x130>(v); }
template <class T>
__device__ inline T from_prev_lane(T v) { return dpp<0x138>(v); }

// The value of lane `l` (a constant), in every lane.
__device__ inline double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Pairwise tree of 64 lanes (lane l <- lane l^m for m = 1..32).  After the
// step with mask m every lane of an aligned block of 2m lanes holds the same
// value, the tree sum of that block, so the result equals the reference's
// "a[2s*l] += a[2s*l+s]" tree bit for bit (IEEE addition commutes exactly).
__device__ inline double wave_tree(double v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v = v + __shfl_xor(v, m);
  return v;
}

// Level-1 trees of the 16 groups a thread's wave holds (value k of lane l is
// cell l of group k), with the values transposed across lanes as the tree
// goes so each butterfly step moves half as many values as the last: after
// steps m = 1, 2, 4, 8 lane l holds one group, k = bitreverse4(l & 15), and
// steps 16 and 32 finish it.  Each step still adds the two aligned halves of
// a block, so every sum is the reference tree's.  Returns, in lanes 0..15,
// the level-1 partial of group bitreverse4(lane).
__device__ inline double wave_tree16(double (&v)[kItems]) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int step = 0; step < 4; ++step) {
    const int m = 1 << step;
    const int half = kItems >> (step + 1);
    const bool upper = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      const double send = upper ? v[i] : v[half + i];
      const double keep = upper ? v[half + i] : v[i];
      v[i] = keep + __shfl_xor(send, m);
    }
  }
  double s = v[0];
  s = s + __shfl_xor(s, 16);
  s = s + __shfl_xor(s, 32);
  return s;
}

__device__ inline int bitrev4(int l) {
  return ((l & 1) << 3) | ((l & 2) << 1) | ((l & 4) >> 1) | ((l & 8) >> 3);
}

// Reduces the 16 values per thread of one span (NR independent sums) to the
// span's level-2 partials and stores them at partials[q*nspan + span].
// `single_group` (the whole field is one group of 64) stores the level-1
// value instead, as the reference's reduceArray then runs no pass at all.
template <int NR>
__device__ inline void span_reduce(double (&v)[NR][kItems], double* partials,
                                   int nspan, int span, bool single_group) {
  __shared__ double lds[NR][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const double t = wave_tree16(v[q]);
    if (lane < kItems) lds[q][bitrev4(lane) * kWaves + wv] = t;
  }
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double level1 = lds[q][lane];
      const double t = single_group ? lds[q][0] : wave_tree(level1);
      if (lane == 0) partials[q * nspan + span] = t;
    }
  }
}


// ---- two cells per lane ---------------------------------------------------
//
// Kernels over cell pairs (2m, 2m+1), m = lane index within a 512-cell slot:
// with x even every pair starts on an aligned boundary, so each operand is
// one 8-byte (FP32) or 16-byte (FP64) load, and the left/right neighbours
// come from the adjacent lanes (lane l-1 holds 2m-1, lane l+1 holds 2m+2);
// only lanes 0 and 63 load theirs.  A wave covers 128 consecutive cells =
// two groups of the reduction tree (lanes 0-31 and 32-63); a 256-thread
// block covers a 4096-cell span in 8 slots of 512 cells.  The first tree
// level is the in-lane pair sum, then the butterfly runs over each 32-lane
// half.

constexpr int kSlots = kSpan / (2 * kBlock);  // 8

__device__ inline double2 ld2(const double* p, int i) {
  return *reinterpret_cast<const double2*>(p + i);
}

__device__ inline float2 ld2(const float* p, int i) {
  return *reinterpret_cast<const float2*>(p + i);
}

// Stores a cell pair with one vector store, or the interior member alone.
template <class T>
__device__ inline void st2(T* p, int i, bool in0, bool in1, T v0, T v1) {
  if (in0 && in1) {
    typedef T vec2 __attribute__((ext_vector_type(2)));
    *reinterpret_cast<vec2*>(p + i) = vec2{v0, v1};
  } else {
    if (in0) p[i] = v0;
    if (in1) p[i + 1] = v1;
  }
}

__device__ inline int bitrev3(int l) {
  return ((l & 1) << 2) | (l & 2) | ((l & 4) >> 2);
}

// Level-1 trees of the 8 slot pair-sums of a lane, transposed across lanes
// as in wave_tree16 (steps 1, 2, 4), then steps 8 and 16 within each 32-lane
// half; stores each group's partial in lds[q][group-in-span] and finishes
// the span's level-2 trees in wave 0 (partials[q * nspan + span]).
template <int NR>
__device__ inline void span_reduce_pairs_n(double (&v)[NR][kSlots], double* partials,
                                           int nspan, int span, bool single_group) {
  __shared__ double lds[NR][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
#pragma unroll
    for (int step = 0; step < 3; ++step) {
      const int m = 1 << step;
      const int half = kSlots >> (step + 1);
      const bool upper = (lane & m) != 0;
#pragma unroll
      for (int i = 0; i < half; ++i) {
        const double send = upper ? v[q][i] : v[q][half + i];
        const double keep = upper ? v[q][half + i] : v[q][i];
        v[q][i] = keep + __shfl_xor(send, m);
      }
    }
    double t = v[q][0];
    t = t + __shfl_xor(t, 8);
    t = t + __shfl_xor(t, 16);
    if ((lane & 31) < kSlots)
      lds[q][bitrev3(lane & 7) * 8 + wv * 2 + (lane >> 5)] = t;
  }
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double r = single_group ? lds[q][0] : wave_tree(lds[q][lane]);
      if (lane == 0) partials[q * nspan + span] = r;
    }
  }
}
// Streaming (nontemporal) access for data read or written once per pass
// and not read again before the next iteration (u, p_new, and the residual
// pass's r); stencil operands (p, r, kx, ky) use ld2 so the neighbouring rows
// can reuse them from L2, and w, which the residual pass reads next, is
// stored normally so that it can stay in the 256 MB Infinity Cache (storing
// it streaming made that pass 1-12% slower).
typedef double dvec2 __attribute__((ext_vector_type(2)));

__device__ inline double2 ld2_nt(const double* p, int i) {
  const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(p + i));
  return double2{v.x, v.y};
}

// Stores a cell pair with one 16-byte streaming store, or the interior
// member alone.
__device__ inline void st2_nt(double* p, int i, bool in0, bool in1, double v0,
                              double v1) {
  if (in0 && in1) {
    __builtin_nontemporal_store(dvec2{v0, v1}, reinterpret_cast<dvec2*>(p + i));
  } else {
    if (in0) __builtin_nontemporal_store(v0, p + i);
    if (in1) __builtin_nontemporal_store(v1, p + i + 1);
  }
}

// The FP32 pair store of the mixed-precision kernels.
typedef float fvec2 __attribute__((ext_vector_type(2)));

__device__ inline void st2_nt(float* p, int i, bool in0, bool in1, float v0,
                              float v1) {
  if (in0 && in1) {
    __builtin_nontemporal_store(fvec2{v0, v1}, reinterpret_cast<fvec2*>(p + i));
  } else {
    if (in0) __builtin_nontemporal_store(v0, p + i);
    if (in1) __builtin_nontemporal_store(v1, p + i + 1);
  }
}

// ---- one pair per thread ---------------------------------------------------
//
// A 256-thread block over 512 consecutive cells = 8 groups of the reduction
// tree, one cell pair per lane.  The reference's level-2 tree over a span's
// 64 level-1 partials (sycl_shared.hpp:59-64: a[2s*l] += a[2s*l + s] for
// s = 1, 2, 4, ...) sums aligned runs of 8 leaves before it combines them, so
// a block's 8-leaf subtree is an exact intermediate of that tree: blocks
// store it in sub[q * nblk + block], and span_finish_kernel combines the 8
// subtrees of each span into its level-2 partial.
//
// These kernels keep the launch order of blocks (no xcd_span renumbering):
// the hardware deals consecutive blocks round-robin to the 8 XCDs, so the
// whole chip streams through memory as one front, and rows re-read above and
// below a block were fetched moments earlier (measured at 4096^2: 8% faster
// than one contiguous range per XCD for 512-cell blocks).

constexpr int kPairBlock = 2 * kBlock;              // 512 cells
constexpr int kSubPerSpan = kSpan / kPairBlock;     // 8
constexpr int kSubMaxNr = 3;                        // dot products per launch

// From each lane's in-lane pair sum (tree step 1): steps 2..32 within each
// 32-lane half (one group each), then the wave's two groups; lane 0 stores
// that sum, g(2w) + g(2w+1), in sub[q * 4 * nblk + 4 * blk + w].  The
// block's four wave sums are combined by span_finish_kernel (tree steps 2
// and 4 of the 8-leaf subtree), so the streaming kernels need no LDS and no
// barrier.
template <int NR>
__device__ inline void wave_subtree_pairs(double (&v)[NR], double* sub,
                                          int nblk, int blk, int wv = -1) {
  const int lane = threadIdx.x & (kWave - 1);
  if (wv < 0) wv = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    // Tree steps 2, 4, 8 and 16 on the DPP path: after each, every lane of
    // an aligned block holds the block's sum (the mirrors pair whole blocks
    // of equal lanes).  Then the four 16-lane sums: (r0 + r1) and (r2 + r3)
    // are the wave's two groups.
    double t = v[q];
    t = t + dpp<0xB1>(t);
    t = t + dpp<0x4E>(t);
    t = t + dpp<0x141>(t);
    t = t + dpp<0x140>(t);
    const double g0 = lane_value(t, 0) + lane_value(t, 16);
    const double g1 = lane_value(t, 32) + lane_value(t, 48);
    if (lane == 0) sub[(size_t)q * kWaves * nblk + kWaves * blk + wv] = g0 + g1;
  }
}

// ---- four cells per lane ---------------------------------------------------
//
// A 128-thread block over one 512-cell pair block, four cells per lane (x a
// multiple of 4, so a quad never straddles a row and every quad, and the
// quads one row up and down, are 16-byte aligned).  The block covers the
// cells of a pair block so that, as there, the rows above and below a block
// at x = 4100 lie 8 blocks back and ahead: on the same XCD, whose L2 has
// them (with 1024-cell blocks they lie 4 blocks away, on another XCD: the
// prolongation, which reads ky one row up, took 12-22% longer at 4096^2).  The in-lane sum ((v0 + v1) +
// (v2 + v3)) is tree steps 1 and 2; the four DPP steps of wave_subtree_pairs
// then finish a 64-cell group in each 16-lane row, so wave w (of 2) holds
// four groups, g0..g3, and stores the sums pair waves 2w and 2w + 1 would:
// g0 + g1 and g2 + g3, in sub[q * 4 * nblk + 4 * block + 2 * w] and the slot
// after it.  span_finish_kernel completes the spans from there, bit for bit
// as for the pair kernels.

constexpr int kQuadThreads = kPairBlock / 4;  // 128
constexpr int kQuadBlock = kPairBlock;        // 512 cells

__device__ inline float4 ld4(const float* p, int i) {
  return *reinterpret_cast<const float4*>(p + i);
}

__device__ inline float q_at(const float4& v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

typedef float fvec4 __attribute__((ext_vector_type(4)));

// Stores a quad with one 16-byte store, or its interior members alone.
__device__ inline void st4(float* p, int i, const bool (&in)[4], const float (&v)[4]) {
  if (in[0] && in[1] && in[2] && in[3]) {
    *reinterpret_cast<fvec4*>(p + i) = fvec4{v[0], v[1], v[2], v[3]};
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c]) p[i + c] = v[c];
  }
}

__device__ inline void st4(float* p, int i, const bool (&in)[4], fvec4 v) {
  if (in[0] && in[1] && in[2] && in[3]) {
    *reinterpret_cast<fvec4*>(p + i) = v;
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c]) p[i + c] = v[c];
  }
}

// The same with streaming stores.
__device__ inline void st4_nt(float* p, int i, const bool (&in)[4], fvec4 v) {
  if (in[0] && in[1] && in[2] && in[3]) {
    __builtin_nontemporal_store(v, reinterpret_cast<fvec4*>(p + i));
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c]) __builtin_nontemporal_store(v[c], p + i + c);
  }
}

__device__ inline void st4_nt(float* p, int i, const bool (&in)[4], const float (&v)[4]) {
  if (in[0] && in[1] && in[2] && in[3]) {
    __builtin_nontemporal_store(fvec4{v[0], v[1], v[2], v[3]}, reinterpret_cast<fvec4*>(p + i));
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c]) __builtin_nontemporal_store(v[c], p + i + c);
  }
}

template <int NR>
__device__ inline void wave_subtree_quads(double (&v)[NR], double* sub, int nblk, int blk,
                                          int wv) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    double t = v[q];
    t = t + dpp<0xB1>(t);
    t = t + dpp<0x4E>(t);
    t = t + dpp<0x141>(t);
    t = t + dpp<0x140>(t);
    const double g01 = lane_value(t, 0) + lane_value(t, 16);
    const double g23 = lane_value(t, 32) + lane_value(t, 48);
    if (lane == 0) {
      double* d = sub + (size_t)q * kWaves * nblk + kWaves * blk + 2 * wv;
      *reinterpret_cast<dvec2*>(d) = dvec2{g01, g23};
    }
  }
}

// The cell pair (e, e + 1) of a pair-per-thread lane, e = 512 b + 2 m for
// lane m of block b (x even: a pair never straddles a row).
struct PairAt {
  int e;          // first cell
  bool valid;     // e < x*y
  bool in0, in1;  // interior cells
  int ic;         // e if valid, else a safe even index (x + 2)
  int ir;         // e if either cell is interior, else x + 2
};

__device__ inline PairAt pair_at(const Geo& g, int e) {
  PairAt q;
  q.e = e;
  q.valid = e < g.n;
  const int row = q.valid ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * g.x;
  q.in0 = q.valid && interior(g, kk, row);
  q.in1 = q.valid && interior(g, kk + 1, row);
  q.ic = q.valid ? e : g.x + 2;
  q.ir = (q.in0 || q.in1) ? e : g.x + 2;
  return q;
}

// One block per span: each thread evaluates `body` on its 16 cells (only
// interior cells; everything else contributes +0.0, as in the reference) and
// the block stores the span's NR level-2 partials.
template <int NR, class Body>
__global__ __launch_bounds__(kBlock) void span_kernel(Geo g, Body body,
                                                      double* partials,
                                                      int nspan,
                                                      int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  double v[NR][kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int e = base + k * kBlock + threadIdx.x;
    double out[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) out[q] = 0.0;
    if (e < g.n) {
      const int jj = (int)g.dx.div((uint32_t)e);
      const int kk = e - jj * g.x;
      if (interior(g, kk, jj)) body(e, kk, jj, out);
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) v[q][k] = out[q];
  }
  span_reduce<NR>(v, partials, nspan, span, single_group != 0);
}

// One thread per cell of the padded x*y range.
template <class Body>
__global__ __launch_bounds__(kBlock) void each_kernel(Geo g, Body body) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.n) return;
  const int jj = (int)g.dx.div((uint32_t)e);
  const int kk = e - jj * g.x;
  body(e, kk, jj);
}

// ---------------------------------------------------------------------------
// Kernel identifiers for timing.

enum KernelId {
  K_SET_CHUNK_DATA = 0,
  K_SET_CHUNK_STATE,
  K_STORE_ENERGY,
  K_LOCAL_HALOS,
  K_FIELD_SUMMARY,
  K_CG_INIT_U,
  K_CG_INIT_K,
  K_CG_INIT_OTHERS,
  K_CG_CALC_W,
  K_CG_CALC_UR,
  K_CG_CALC_P,
  K_CG_FUSED_W,
  K_CG_FUSED_R,
  K_CG_APPLY_PENDING,
  K_COPY_U,
  K_CALC_RESIDUAL,
  K_CALC_2NORM,
  K_FINALISE,
  K_MP_CONVERT,
  K_MP_W,
  K_MP_R,
  K_MP_APPLY_U,
  K_MP_REFRESH,
  K_FCG_W,
  K_FCG_JACOBI,
  K_FCG_RESTRICT,
  K_FCG_NET,
  K_FCG_PROLONG,
  K_SPAN_FINISH,
  K_JACOBI_INIT,
  K_JACOBI_COPY_U,
  K_JACOBI_ITERATE,
  K_JACOBI_FUSED,
  K_CHEBY_INIT,
  K_CHEBY_ITERATE,
  K_CHEBY_CALC_U,
  K_CHEBY_FUSED,
  K_PPCG_INIT,
  K_PPCG_CALC_UR,
  K_PPCG_CALC_SD,
  K_PPCG_FUSED,
  K_PPCG_TWO_STEP,
  K_CHEBY_TWO_STEP,
  K_FCG_SWEEPS,
  K_CG_RA_FINISH,
  K_NUM
};

extern const char* const kKernelNames[K_NUM];

}  // namespace tea

namespace tea {

// Arguments of the mixed-precision stencil kernel (mpcg.hip), shared with
// flexible CG (fcg.hip).
struct MpArgs {
  const float* p_old;
  float* p_new;
  const float* r;     // added to beta p_old (the residual, or z for FCG)
  double* u;
  float* w;
  const float* kx;
  const float* ky;
  const float* rdot;  // second dot product p.rdot (DOT2 only)
  double alpha, beta;
  int apply_u, apply_p;
  // Held updates applied first (mp_w4_kernel, UM 3): u += alpha_h p_h,
  // then u += alpha_h2 p_h2; p_h is p_new's buffer, read before it is
  // overwritten.
  const float* p_h;
  const float* p_h2;
  double alpha_h, alpha_h2;
};

template <int kBatch, bool DOT2>
__global__ void mp_w_kernel(Geo g, MpArgs a, double* partials, int nspan,
                            int single_group);
template <bool RW>
__global__ void mp_r_kernel(Geo g, float* r, const float* w, double alpha,
                            double* partials, int nspan, int single_group);
template <bool DOT2, int SB>
__global__ void mp_w2_kernel(Geo g, MpArgs a, double* partials, int nspan,
                             int single_group);
template <bool DOT2>
__global__ void mp_w1_kernel(Geo g, MpArgs a, double* sub, int nblk);
// UM: 0 leaves u alone, 1 adds alpha p_old, 3 adds alpha_h p_h, alpha_h2
// p_h2 and then alpha p_old.
template <bool DOT2, int UM, int KS>
__global__ void mp_w4_kernel(Geo g, MpArgs a, double* sub, int nblk, int S);

}  // namespace tea

struct tea_nn;

namespace tea {
namespace nn {
typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Stores one 32-pixel row of a layer's MFMA output (bias, ReLU, BF16) at
// px, the row's pixel of lane col: lane (hf, col) holds channels 8 g + 4 hf
// + 0..3, g = 0..3, of pixel col.  Lanes col and col + 32 swap half of
// theirs, so that each holds 16 consecutive channels (32 bytes at 16 hf),
// stored with two 16-byte stores (four 8-byte stores at a 64-byte stride
// each wrote the first layer's output at 1.5 TB/s).  Every lane takes part
// in the exchange; only lanes with ok store.
__device__ inline void store_out_row(bf16* px, const f32x16& acc, const float (&bco)[16],
                                     int hf, bool ok = true) {
  union Pack {
    bf16x4 v;
    unsigned u[2];
  };
  Pack o[4];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float vv = acc[4 * g + q] + bco[4 * g + q];
      o[g].v[q] = (bf16)(vv > 0.0f ? vv : 0.0f);
    }
  // hf 0 keeps channels 0-3 and 8-11 and takes 4-7 and 12-15 (the
  // partner's g 0, 1); hf 1 keeps 20-23 and 28-31 and takes 16-19 and
  // 24-27 (the partner's g 2, 3).
  Pack in[2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int w = 0; w < 2; ++w) in[j].u[w] = __shfl_xor(hf ? o[j].u[w] : o[2 + j].u[w], 32);
  union Pack8 {
    bf16x8 v;
    bf16x4 h[2];
  };
  Pack8 lo, hi;
  lo.h[0] = hf ? in[0].v : o[0].v;
  lo.h[1] = hf ? o[2].v : in[0].v;
  hi.h[0] = hf ? in[1].v : o[1].v;
  hi.h[1] = hf ? o[3].v : in[1].v;
  if (!ok) return;
  *reinterpret_cast<bf16x8*>(px + 16 * hf) = lo.v;
  *reinterpret_cast<bf16x8*>(px + 16 * hf + 8) = hi.v;
}

// The network's device copies of its parameters, in the layouts of its
// forward kernels: layer 1 wf1[co][tap][8] (3 channels used) and the hidden
// layers wfh[layer][co][tap][ci] in BF16, the output layer w6[tap][ci]
// (BF16 values in FP32) and the biases bias[layer * 32 + co], bias[160].
struct PackedWeights {
  const bf16* wf1;
  const bf16* wfh;
  const float* w6;
  const float* bias;
};
PackedWeights packed_weights(const tea_nn* n);
int forward_on_stream(tea_nn* n, hipStream_t st);
// The hidden layers alone, from layer `first` on (0-based; the caller has
// written layer first - 1's activations at hidden(n, first - 1)); the last
// layer's activations (padded (H + 2) x (W + 2) pixels of 32 BF16 channels,
// zero ring) at last_hidden.
int forward_hidden_on_stream(tea_nn* n, hipStream_t st, int first = 0);
const bf16* last_hidden(tea_nn* n);
bf16* hidden(tea_nn* n, int l);
bf16* input_tensor(tea_nn* n);
const float* output(tea_nn* n);
}  // namespace nn
}  // namespace tea

// ---------------------------------------------------------------------------
// The chunk handle.

// The mixed-precision p buffers: a ring of three, so that a w-step can
// hold the u updates of two steps back (their p are the two buffers it
// does not write).
constexpr int kMpRing = 3;
inline int mp_next(int i) { return i + 1 == kMpRing ? 0 : i + 1; }

// One coarse level of the learned preconditioner.
struct FcgLevel {
  int c = 0, hc = 0, wc = 0;  // coarsening factor, coarse rows and columns
  tea_nn* net = nullptr;      // the network, sized hc x wc
  float* feat = nullptr;      // [2][hc][wc] normalised features
  float* rbar = nullptr;      // [hc][wc] block means of r
  float* kc = nullptr;        // [2][hc][wc] coarse left / bottom face coefficients
  float* dis = nullptr;       // [hc][wc] D^-1/2 of the coarse operator
  float* sig = nullptr;       // [hc][wc] local input scale of the network
  double* rms = nullptr;      // per-restriction-block sums of (D^-1/2 rbar)^2
  int nrms = 0;
  // The coarse solve's output y (network or Jacobi sweeps) and the
  // per-16x16-tile partials of y.rbar and y.T y (2 * ntiles).
  float* y = nullptr;
  double* tpart = nullptr;
  int ntiles = 0;
};
constexpr int kFcgMaxLevels = 2;

struct tea_hip_chunk {
  int device = 0;
  tea::Geo geo{};
  int nspan = 0;   // spans covering x*y (level-2 partial count)
  int ngroup = 0;  // groups of 64 covering x*y (level-1 partial count)
  hipStream_t stream = nullptr;

  double* field[TEA_NUM_FIELDS] = {};
  size_t field_len[TEA_NUM_FIELDS] = {};

  // Second p buffer for the fused iteration (p is read as a stencil while
  // the new p is written).  p_alt_synced: its non-interior cells equal p's.
  double* p_alt = nullptr;
  bool p_alt_synced = false;

  // Reduction workspace: up to 4 x nspan level-2 partials in pinned,
  // coherent host memory; kernels write through the device view.
  double* partials = nullptr;       // device view
  double* host_partials = nullptr;  // host view of the same memory
  // Wave sums of the pair-per-thread kernels: kSubMaxNr x 4 x nblk, device.
  int nblk = 0;                     // 512-cell blocks covering x*y
  double* sub = nullptr;
  int pair_blocks = 3;              // pair-per-thread kernels: bit 0 fused_w,
                                    // bit 1 fused_r

  // Spare FP64 field buffer (x*y, allocated by the first Chebyshev or PPCG
  // call): the fused Chebyshev and PPCG steps write the new u or sd into it
  // and exchange it with the field.
  double* spare = nullptr;
  // Second spare buffer (x*y, allocated by the first two-step pass), which
  // takes the new r (PPCG) or p (Chebyshev) there: neighbouring tiles still
  // read the old values.
  double* spare2 = nullptr;

  // A PPCG inner step or Chebyshev step held back (fused mode) so that it
  // can run in one pass with the next one: which (kHeld*), its
  // coefficients, and whether the depth-1 halo update that follows it (of
  // SD, or U) has been requested and held back too.
  int held = 0;
  bool held_halo = false;
  double held_alpha = 0.0, held_beta = 0.0;
  // A depth-1 reflective update of SD (PPCG) or U (Chebyshev), all faces
  // external, requested with no step held and held back so that the next
  // two-step pass reflects the ring of the field it stages instead (the
  // halo slot, or -1).
  int pre_halo = -1;

  // Deferred CG work (fused mode).
  int fused = 1;
  int batch = 0;   // load batch of the fused stencil kernel (0: default)
  bool pairs = false;  // two-cells-per-lane kernels (x even)
  bool quads = false;  // four-cells-per-lane mixed-precision kernels (x % 4 == 0)
  bool pend_u = false, pend_p = false, pend_halo = false;
  double pend_alpha = 0.0, pend_beta = 0.0;

  // Mixed-precision CG: FP32 p (a ring of three buffers, mp_cur is
  // current; each w-step writes the next), r, w and
  // face coefficients; allocated by the first tea_hip_mpcg_init.
  float* mp_p[kMpRing] = {nullptr, nullptr, nullptr};
  int mp_cur = 0;
  // Up to two updates u += hold_alpha[i] p, p in mp_p[hold_buf[i]] (oldest
  // first), reported done to the host but held back (TEA_HIP_HOLD_U=0 turns
  // this off): the four-cells-per-lane w-step that overwrites the oldest's
  // p buffer reads both held p first and makes the held updates and its own
  // in one pass over u; any other entry point makes them first.
  bool hold_ok = true;
  int hold = 0;
  int hold_buf[2] = {0, 0};
  double hold_alpha[2] = {0.0, 0.0};
  float* mp_r = nullptr;
  float* mp_w = nullptr;
  float* mp_kx = nullptr;
  float* mp_ky = nullptr;

  // Flexible CG: the preconditioned residual z (FP32) and, for the learned
  // preconditioner, up to kFcgMaxLevels coarse levels (coarsening c, 4c),
  // each with its network instance, coarse operator and work arrays.
  float* fcg_z = nullptr;
  float fcg_norm[4] = {0.0f, 1.0f, 0.0f, 1.0f};
  int fcg_levels = 0;
  FcgLevel fcg_lv[kFcgMaxLevels];
  // Coarse solve by fcg_sweeps Jacobi sweeps instead of the network (0:
  // network); the level-0 sums of T y over the level-1 blocks (the cross
  // term of two levels), and the factors g of the levels.
  int fcg_sweeps = 0;
  double* fcg_xsum = nullptr;
  float* fcg_g = nullptr;
  // With two levels, the second level's chain runs on a side stream
  // (default priority) between a fork after the first level's restriction
  // and a join before the joint factors.
  hipStream_t fcg_side = nullptr;
  hipEvent_t fcg_fork = nullptr, fcg_join = nullptr;

  // Run-ahead of the fused CG iteration (cg.hip): the device forms alpha and
  // beta from its own partials and runs the next pass of the iteration into
  // spare buffers before the host asks for it; the host's call commits it
  // when its scalar is bitwise the device's.
  int ra = 0;                 // kRaNone, kRaR (r-step in flight), kRaWR
  bool ra_enabled = true;     // TEA_HIP_RUN_AHEAD=0 turns it off
  // Blocks one row apart per workgroup of the stencil passes mp_w4_kernel,
  // fused_w1_kernel and jacobi_pair_kernel: 1, 2 (the default) or 4
  // (TEA_HIP_STACK).  (fcg_z4_kernel, whose only row above is ky's, took
  // 57.4 us with two against 56.0 us with one.)
  int stack = 2;
  bool ra_dirty = false;      // a discarded run-ahead may still be writing ra_host
  int ra_fail = 0;            // consecutive mismatches (3 turn it off)
  int ra_set = 0;             // host slot set of the step in flight
  long ra_commits = 0;        // run-ahead passes taken over (statistics)
  double ra_rro = 0.0;        // r.r the host was last given (its rro)
  bool ra_rro_known = false;
  double ra_alpha = 0.0;      // alpha of the r-step in flight (the device's)
  double ra_beta = 0.0;       // beta of the w-step in flight (the device's)
  double ra_w_alpha = 0.0;    // alpha whose u update the w-step in flight applies
  double* r_alt = nullptr;    // spare r, u and w (x*y each)
  double* u_alt = nullptr;
  double* w_alt = nullptr;
  double* ra_dev = nullptr;   // device: [0] alpha, [1] beta, [2] r.r, then the
                              // finisher's partials and its blocks-done count
  unsigned* ra_count = nullptr;
  double* ra_host = nullptr;  // pinned: 2 sets x {w-step, r-step} x {sum, ratio}
  double* ra_host_dev = nullptr;  // device view of ra_host

  // Waiting for a reduction by its values: launch_span_finish(..., poll)
  // fills the first poll_len partials of arrays 0 .. poll_nr-1 with a
  // sentinel, and the finish waits until the finisher has overwritten them
  // all (or the stream is idle).
  bool poll = true;
  int poll_nr = 0, poll_len = 0;

  // A depth-1 reflective update also copies the cells of its face lines
  // beyond the ring (e.g. column h-1 at rows below h-1) from other cells
  // outside the interior, which no kernel of the iteration changes.  A
  // U/P update is deferred into a fused pass, which writes only the ring,
  // once a real update of all external faces has made those cells what the
  // update would leave; host writes and initialisations clear the flags.
  bool faces_ok_u = false, faces_ok_p = false;

  // Timing.
  bool timing = false;
  struct Pending {
    int kid;
    hipEvent_t start, stop;
  };
  std::vector<Pending> inflight;
  std::vector<hipEvent_t> event_pool;
  bool time_kernel[tea::K_NUM];     // kernels timed (all, or TEA_HIP_STATS_ONLY)
  long launches[tea::K_NUM] = {};
  double total_ms[tea::K_NUM] = {};
};

namespace tea {

// Records start/stop events around one kernel launch when timing is on.
struct Timed {
  tea_hip_chunk* c;
  int kid;
  hipStream_t stream;  // the chunk's stream unless given
  hipEvent_t start = nullptr, stop = nullptr;
  Timed(tea_hip_chunk* c_, int kid_, hipStream_t s = nullptr);
  ~Timed();
};

// Resolves the recorded events (after the stream has been synchronised).
int resolve_timing(tea_hip_chunk* c);

// Synchronises the stream and resolves timing.
int sync_chunk(tea_hip_chunk* c);

// Resolves the recorded events that have completed, oldest first, without
// waiting.
int resolve_ready(tea_hip_chunk* c);

// Waits for the stream, then finishes nr reductions on the host: the
// reference tree's remaining passes over each array of nspan level-2
// partials; results go to out[0..nr).
int finish_reduction(tea_hip_chunk* c, int nr, double* out);

// Combines the 8 subtree sums of each span in c->sub (nr dot products) into
// the span's level-2 partials and runs the next tree pass over them, leaving
// *len partials per dot product in c->partials.
// The partials go to arrays slot .. slot + nr - 1 of c->partials.  With
// poll (slot 0), the next finish waits for the values themselves instead of
// the stream; only for a finish that follows with no other kernel writing
// the partials in between.
// The run-ahead's scalar, formed by the last block of the finisher once
// every block's partial is in: the host's final tree passes (zero-padded
// groups of 64, each by the reference's pairwise tree) over the grid's
// partials give the sum; mode 0 (after a w-step) alpha = rro / sum into
// ra[0]; mode 1 (after an r-step) beta = sum / rro into ra[1] and sum into
// ra[2].  rro is ra[2] when from_dev, else the value.  The sum and the
// scalar also go to out[0], out[1] (pinned host memory the host polls).
// count is the blocks-done counter, zero between launches.
struct RaTail {
  double rro;
  int from_dev, mode;
  double* ra;
  double* out;
  unsigned* count;
};
// The finisher of one reduction (array 0 of sub) into dst, device memory
// (ceil(nspan / 64) partials, 1 with a single span), with the run-ahead's
// scalar (at most 64^2 partials); not waited for.
int launch_span_finish_ra(tea_hip_chunk* c, double* dst, const RaTail& tail);
int launch_span_finish(tea_hip_chunk* c, int nr, int* len, int slot = 0,
                       bool poll = false);

// finish_reduction over the first len partials of each array (len is what
// launch_span_finish leaves).
int finish_reduction_n(tea_hip_chunk* c, int nr, int len, double* out);
// Values in pinned, mapped host memory that a kernel of the stream writes:
// arm them with the wait sentinel before the launch, then wait until the
// kernel has overwritten all count of them.
void arm_host_values(double* host, int count);
int wait_host_values(tea_hip_chunk* c, const double* host, int count);

// The same with a length per array.
int finish_reduction_lens(tea_hip_chunk* c, int nr, const int* lens, double* out);

// The pair-per-thread kernels apply: x even (pairs) and more than one group
// (the reference then runs its level-2 pass; for a single group it stops at
// level 1, which the span-per-block kernels reproduce).
inline bool use_pair_blocks(const tea_hip_chunk* c, int kernel_bit) {
  return c->pairs && (c->pair_blocks & kernel_bit) && c->ngroup > 1;
}

// The four-cells-per-lane variants of the pair-per-thread kernels (same
// results bit for bit) apply where those do and x is a multiple of 4.
inline bool use_quad_blocks(const tea_hip_chunk* c, int kernel_bit) {
  return c->quads && use_pair_blocks(c, kernel_bit);
}

// Kinds of held-back step.
enum { kHeldNone = 0, kHeldPpcg = 1, kHeldCheby = 2 };

// Applies the deferred CG work (fused mode: u += alpha p, p = beta p + r and
// their halo ring), then runs a held-back step.  Deferred CG work is always
// older than a held step: a step is held with at most u += alpha p pending,
// and the CG entry points run a held step before they defer new work.
int apply_pending(tea_hip_chunk* c);
// The deferred CG work alone (cg.hip).
int apply_cg_pending(tea_hip_chunk* c);
// A held-back halo update before a step, alone (solvers.hip).
int apply_pre_halo(tea_hip_chunk* c);
// Runs what is held back, oldest first: deferred CG work, a held-back halo
// update before a step (pre_halo), a held-back PPCG inner step or Chebyshev
// step and its held-back halo update (solvers.hip).
int flush_held(tea_hip_chunk* c);
// Mixed-precision residual update r -= alpha w with the partials of r.r
// (and, with rw, of r.w as a second array) written to c->partials
// (mpcg.hip); *len is the number left to finish in each.  Without finish
// (four-cells-per-lane kernel only) the subtree sums stay in c->sub for the
// caller's finisher.
int launch_mp_r(tea_hip_chunk* c, double alpha, int* len, bool rw, bool finish = true);
// The four-cells-per-lane w-step for the arguments a (p_old in
// mp_p[mp_cur], p_new in mp_p[mp_next(mp_cur)]), timed as kernel kid, with
// u += alpha p_old held back over two steps: held while fewer than two
// are, made with the two held updates when the oldest one's p is this
// step's p_new buffer; held updates that cannot ride along run first on
// their own (mpcg.hip).
int launch_mp_w4(tea_hip_chunk* c, MpArgs a, int kid);
// Makes the held-back u updates on their own, oldest first (mpcg.hip).
int apply_mp_hold(tea_hip_chunk* c);

// One reflective-halo launch (all flagged fields and external faces).
int launch_halos(tea_hip_chunk* c, int depth, const int neighbours[4],
                 const int fields[6]);

// Drops a run-ahead of the fused CG iteration (cg.hip): it wrote only
// spare buffers and device scalars, so nothing is undone.
int ra_discard(tea_hip_chunk* c);

// Checks the handle and makes its device current, leaving a run-ahead in
// flight (the CG entry points that can take it over); applies a held-back
// mixed-precision u update.
#define TEA_ENTER_RA(c)                                   \
  TEA_CHECK_ARG((c) != nullptr, "null chunk handle");     \
  TEA_HIP_TRY(hipSetDevice((c)->device));                 \
  if ((c)->hold) TEA_RETURN_IF(::tea::apply_mp_hold(c))

// Checks the handle and makes its device current (entry points); drops a
// run-ahead of the CG iteration.
#define TEA_ENTER_HELD(c)                                 \
  TEA_ENTER_RA(c);                                        \
  if ((c)->ra) TEA_RETURN_IF(::tea::ra_discard(c))

// Entry points other than the PPCG inner step and the halo update first run
// a held-back halo update and PPCG or Chebyshev step (and drop a run-ahead)
// and apply a held-back mixed-precision u update.
#define TEA_FLUSH_HELD(c)                                                  \
  do {                                                                     \
    if ((c)->hold) TEA_RETURN_IF(::tea::apply_mp_hold(c));                 \
    TEA_FLUSH_HELD_KEEP_U(c);                                              \
  } while (0)

// TEA_FLUSH_HELD leaving a held-back mixed-precision u update held: for the
// mixed-precision and FCG steps, which touch neither u nor the held p.
#define TEA_FLUSH_HELD_KEEP_U(c)                                           \
  do {                                                                     \
    if ((c)->ra) TEA_RETURN_IF(::tea::ra_discard(c));                      \
    if ((c)->held || (c)->pre_halo >= 0) TEA_RETURN_IF(::tea::flush_held(c)); \
  } while (0)

#define TEA_ENTER(c)   \
  TEA_ENTER_HELD(c);   \
  TEA_FLUSH_HELD(c)

inline dim3 span_grid(const tea_hip_chunk* c) { return dim3(c->nspan); }
inline dim3 pair_grid(const tea_hip_chunk* c) { return dim3(c->nblk); }
inline dim3 quad_grid(const tea_hip_chunk* c) { return dim3(c->nblk); }
inline dim3 flat_grid(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// Stacked workgroups of the 512-cell-block stencil kernels: with S = x / 512
// blocks per row, block c + S t holds the cells one row above block c + S
// (t - 1), so a workgroup of KS blocks takes c + S t for t = KS j .. KS j +
// KS - 1 and their rows above and below are mostly each other's own rows,
// loaded on the same CU (L1 hits, not L2 requests).  Each block's cells,
// lanes and wave sums are those of one block per workgroup.  stack_factor is
// c->stack where it applies (1, 2 or 4; 1 for rows shorter than a block).
inline int stack_factor(const tea_hip_chunk* c) {
  const int S = c->geo.x / kPairBlock;
  return (c->stack == 2 || c->stack == 4) && S >= 1 ? c->stack : 1;
}
inline int stack_rows(const tea_hip_chunk* c) { return c->geo.x / kPairBlock; }
inline dim3 stacked_grid(const tea_hip_chunk* c, int ks) {
  if (ks == 1) return dim3(c->nblk);
  const int S = stack_rows(c), tn = (c->nblk + S - 1) / S;
  return dim3((unsigned)(S * ((tn + ks - 1) / ks)));
}
// The block and thread index of the calling thread in a workgroup of KS
// blocks of T threads (blk >= nblk for the unused blocks of the last row).
template <int KS, int T>
__device__ inline int stacked_block(int S, int* tid) {
  *tid = KS == 1 ? (int)threadIdx.x : (int)threadIdx.x % T;
  if (KS == 1) return (int)blockIdx.x;
  return (int)(blockIdx.x % S) + S * ((int)(blockIdx.x / S) * KS + (int)threadIdx.x / T);
}

}  // namespace tea
