// libteahip.so: the learned preconditioner's network (no counterpart in the
// reference).  Six 3x3 convolutions, 32 hidden channels, ReLU after the
// first five; inputs (residual, kx feature, ky feature), one output channel.
// 38,177 parameters, FP32 master copy; BF16 weights and activations with
// FP32 accumulation on the MFMA units (v_mfma_f32_32x32x16_bf16).
//
// Tensors are NHWC with a one-pixel zero border: [B][H+2][W+2][C], C = 8
// for the input (3 used) and 32 for hidden activations, so a pixel's 8 or
// 32 channels are one or four 16-byte vectors.  W must be a multiple of 32.
//
// A 32->32 convolution is an implicit GEMM per row tile of 32 pixels:
// D[co][pixel] = sum over k = (tap, ci) of Wt[co][k] * In[pixel + tap][ci],
// K = 9*32 = 288 = 18 MFMA steps of 16.  A (32 x 16) is the weights (lane l:
// co = l%32, k = step slice + 8*(l/32) + 0..7), B (16 x 32) the input patch
// (lane l: pixel l%32, same k), D (32 x 32): lane l holds pixel l%32 and
// channels 8*(v/4) + 4*(l/32) + v%4 for v = 0..15 (layout checked on the
// device).  The backward-data pass is the same kernel with the weights
// flipped and transposed, times the ReLU mask of the layer's input.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "tea_hip_impl.h"

namespace tea {
namespace nn {

constexpr int kC = 32;   // hidden channels
constexpr int kC0 = 8;   // input channels in memory (3 used)
constexpr int kIn = 3;   // real input channels
constexpr int kTaps = 9;
constexpr int kFwdRows = 4;  // output rows per block of the multi-row forward kernel

// FP32 master parameters, in the order W1 (32,3,3,3), b1 (32), W2..W5
// (32,32,3,3) with b2..b5 (32), W6 (1,32,3,3), b6 (1); W[co][ci][ky][kx].
constexpr int kW1 = kC * kIn * kTaps;  // 864
constexpr int kWh = kC * kC * kTaps;   // 9216
constexpr int kW6 = kC * kTaps;        // 288
constexpr int kNumParams = (kW1 + kC) + 4 * (kWh + kC) + (kW6 + 1);  // 38177
constexpr int kWgradMaxW = 128;  // widest row the weight-gradient kernel stages

__host__ __device__ inline int w_offset(int layer) {  // layer 0..5
  return layer == 0 ? 0 : (kW1 + kC) + (layer - 1) * (kWh + kC);
}
__host__ __device__ inline int b_offset(int layer) {
  return layer == 0 ? kW1 : layer < 5 ? w_offset(layer) + kWh : w_offset(5) + kW6;
}

// ---- forward / backward-data convolution on MFMA ---------------------------

// k slice of MFMA step t for lane half hf: CIN = 32 -> tap t/2, channels
// 16*(t%2) + 8*hf; CIN = 8 -> tap 2t + hf (tap 9 is padding), channels 0..7.
template <int CIN>
__device__ inline void kmap(int t, int hf, int& tap, int& ci) {
  if (CIN == kC) {
    tap = t >> 1;
    ci = ((t & 1) << 4) + 8 * hf;
  } else {
    tap = 2 * t + hf;
    ci = 0;
  }
}

template <int CIN>
constexpr int nsteps() {
  return CIN == kC ? 18 : 5;
}

// MODE 0: out = relu(D + bias).  MODE 1: out = D * (mask > 0).
// Grid (H, B), 256 threads; each wave walks 32-pixel tiles of row y, so the
// weights it holds in registers serve several tiles.
template <int CIN, int MODE>
__global__ __launch_bounds__(256) void conv_mfma_kernel(
    const bf16* __restrict__ in, const bf16* __restrict__ wt,
    const float* __restrict__ bias, const bf16* __restrict__ mask,
    bf16* __restrict__ out, int H, int W, int B) {
  constexpr int NS = nsteps<CIN>();
  (void)B;
  const int y = blockIdx.x, b = blockIdx.y;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int hf = lane >> 5, col = lane & 31;
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  bf16x8 A[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    int tap, ci;
    kmap<CIN>(t, hf, tap, ci);
    if (tap < kTaps)
      A[t] = *reinterpret_cast<const bf16x8*>(wt + ((size_t)col * kTaps + tap) * CIN + ci);
    else
      A[t] = bf16x8{};
  }
  float bco[16];
  if (MODE == 0) {
#pragma unroll
    for (int v = 0; v < 16; ++v) bco[v] = bias[8 * (v >> 2) + 4 * hf + (v & 3)];
  }
  const bf16* inb = in + (size_t)b * plane * CIN;
  for (int tile = wv; tile < W / 32; tile += 4) {
    const int xo = tile * 32 + col;
    bf16x8 Bv[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      int tap, ci;
      kmap<CIN>(t, hf, tap, ci);
      if (tap < kTaps) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        Bv[t] = *reinterpret_cast<const bf16x8*>(
            inb + ((size_t)(y + 1 + dy) * Wp + (xo + 1 + dx)) * CIN + ci);
      } else {
        Bv[t] = bf16x8{};
      }
    }
    f32x16 acc = {};
#pragma unroll
    for (int t = 0; t < NS; ++t)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[t], Bv[t], acc, 0, 0, 0);
    const size_t pix = (size_t)b * plane + (size_t)(y + 1) * Wp + (xo + 1);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int co0 = 8 * g + 4 * hf;
      bf16x4 o;
      if (MODE == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float vv = acc[4 * g + q] + bco[4 * g + q];
          o[q] = (bf16)(vv > 0.0f ? vv : 0.0f);
        }
      } else {
        const bf16x4 m = *reinterpret_cast<const bf16x4*>(mask + pix * kC + co0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          o[q] = (bf16)((float)m[q] > 0.0f ? acc[4 * g + q] : 0.0f);
      }
      *reinterpret_cast<bf16x4*>(out + pix * kC + co0) = o;
    }
  }
}

// Stages NROWS padded input rows (from row y0) of a 130-pixel segment
// (from padded column x0) into LDS, CPC 16-byte chunks per pixel at a pitch
// of kPitch chunks: every thread issues all its loads before its first LDS
// store (a load-store loop waited for each load in turn: 13 global-memory
// latencies per thread of the four-row kernel).
template <int NROWS, int CPC, int kPx, int kPitch>
__device__ inline void stage_rows(bf16x8* sm, const bf16* __restrict__ inb, int y0, int Wp,
                                  int x0) {
  constexpr int kTotal = NROWS * kPx * CPC;
  constexpr int kIt = (kTotal + 255) / 256;
  constexpr int CIN = CPC * 8;
  bf16x8 v[kIt];
  int slot[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const int q = threadIdx.x + 256 * k;
    const int qq = q < kTotal ? q : 0;
    const int row = qq / (kPx * CPC), rem = qq - row * (kPx * CPC);
    const int px = rem / CPC, ch = rem - px * CPC;
    slot[k] = row * (kPx * kPitch) + px * kPitch + ch;
    v[k] = *reinterpret_cast<const bf16x8*>(inb + ((size_t)(y0 + row) * Wp + (x0 + px)) * CIN +
                                            ch * 8);
  }
#pragma unroll
  for (int k = 0; k < kIt; ++k)
    if (threadIdx.x + 256 * k < kTotal) sm[slot[k]] = v[k];
}

// The (segment, row group, sample) of this block in a (W / 128, H / R, B)
// grid, numbered so that each XCD takes one contiguous range of row groups:
// the hardware deals blocks round-robin to the 8 XCDs, which put the row
// groups above and below a block on other XCDs, so the input rows two
// blocks share were fetched into two L2s (1.27x the activations per hidden
// layer at 512^2 with four rows per block).
__device__ inline void xcd_tile(int& seg, int& grp, int& b) {
  const int nb = gridDim.x * gridDim.y * gridDim.z;
  const int L = xcd_span(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
  seg = L % gridDim.x;
  const int rest = L / gridDim.x;
  grp = rest % gridDim.y;
  b = rest / gridDim.y;
}

// Forward layer for rows of 128 pixels or more: a block takes a 128-pixel
// segment of row y, stages input rows y-1..y+1 over the segment plus its
// 1-pixel ring in LDS once (130 pixels x CIN channels x 3 rows, in 16-byte
// chunks; a pixel's chunks are padded to kPitch so that the 16 lanes of a
// 16-byte LDS read, which take consecutive pixels, hit distinct banks),
// and each wave computes one 32-pixel tile from LDS.  Same arithmetic and
// output as conv_mfma_kernel<CIN, 0>.  Grid (W / 128, H, B).
template <int CIN>
__global__ __launch_bounds__(256) void conv_fwd_lds_kernel(
    const bf16* __restrict__ in, const bf16* __restrict__ wt,
    const float* __restrict__ bias, bf16* __restrict__ out, int H, int W) {
  constexpr int NS = nsteps<CIN>();
  constexpr int CPC = CIN / 8;                   // 16-byte chunks per pixel
  constexpr int kPitch = CPC == 1 ? 1 : CPC + 1;  // padded chunks per pixel
  constexpr int kPx = 130;
  __shared__ bf16x8 sm[3][kPx * kPitch];
  int seg, y, b;
  xcd_tile(seg, y, b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int hf = lane >> 5, col = lane & 31;
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  const bf16* inb = in + (size_t)b * plane * CIN;
  const int x0 = seg * 128;  // padded column of the segment's left ring pixel
  // Weights and biases first: loaded after the staged rows' LDS stores, they
  // were a second global-memory latency before the first MFMA.
  bf16x8 A[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    int tap, ci;
    kmap<CIN>(t, hf, tap, ci);
    if (tap < kTaps)
      A[t] = *reinterpret_cast<const bf16x8*>(wt + ((size_t)col * kTaps + tap) * CIN + ci);
    else
      A[t] = bf16x8{};
  }
  float bco[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) bco[v] = bias[8 * (v >> 2) + 4 * hf + (v & 3)];
  stage_rows<3, CPC, kPx, kPitch>(&sm[0][0], inb, y, Wp, x0);
  __syncthreads();
  const int xl = wv * 32 + col;  // pixel within the segment
  bf16x8 Bv[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    int tap, ci;
    kmap<CIN>(t, hf, tap, ci);
    if (tap < kTaps) {
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      Bv[t] = sm[dy + 1][(xl + 1 + dx) * kPitch + ci / 8];
    } else {
      Bv[t] = bf16x8{};
    }
  }
  f32x16 acc = {};
#pragma unroll
  for (int t = 0; t < NS; ++t)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[t], Bv[t], acc, 0, 0, 0);
  const size_t pix = (size_t)b * plane + (size_t)(y + 1) * Wp + (x0 + xl + 1);
  store_out_row(out + pix * kC, acc, bco, hf);
}

// conv_fwd_lds_kernel over R output rows per block: the block stages R + 2
// input rows once (instead of 3 per output row) and each wave keeps its
// weight fragments in registers across the R rows (instead of reloading
// them per 32-pixel tile).  Each tile's MFMA sequence, bias and ReLU are
// those of conv_fwd_lds_kernel, so the output is bitwise the same.
// Grid (W / 128, H / R, B).
template <int CIN, int R>
__global__ __launch_bounds__(256) void conv_fwd_rows_kernel(
    const bf16* __restrict__ in, const bf16* __restrict__ wt,
    const float* __restrict__ bias, bf16* __restrict__ out, int H, int W) {
  constexpr int NS = nsteps<CIN>();
  constexpr int CPC = CIN / 8;
  constexpr int kPitch = CPC == 1 ? 1 : CPC + 1;
  constexpr int kPx = 130;
  __shared__ bf16x8 sm[R + 2][kPx * kPitch];
  int seg, grp, b;
  xcd_tile(seg, grp, b);
  const int y0 = grp * R;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int hf = lane >> 5, col = lane & 31;
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  const bf16* inb = in + (size_t)b * plane * CIN;
  const int x0 = seg * 128;
  // Weights and biases first (as conv_fwd_lds_kernel).
  bf16x8 A[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    int tap, ci;
    kmap<CIN>(t, hf, tap, ci);
    if (tap < kTaps)
      A[t] = *reinterpret_cast<const bf16x8*>(wt + ((size_t)col * kTaps + tap) * CIN + ci);
    else
      A[t] = bf16x8{};
  }
  float bco[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) bco[v] = bias[8 * (v >> 2) + 4 * hf + (v & 3)];
  stage_rows<R + 2, CPC, kPx, kPitch>(&sm[0][0], inb, y0, Wp, x0);
  __syncthreads();
  const int xl = wv * 32 + col;
#pragma unroll 1
  for (int r = 0; r < R; ++r) {
    bf16x8 Bv[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      int tap, ci;
      kmap<CIN>(t, hf, tap, ci);
      if (tap < kTaps) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        Bv[t] = sm[r + dy + 1][(xl + 1 + dx) * kPitch + ci / 8];
      } else {
        Bv[t] = bf16x8{};
      }
    }
    f32x16 acc = {};
#pragma unroll
    for (int t = 0; t < NS; ++t)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[t], Bv[t], acc, 0, 0, 0);
    const size_t pix = (size_t)b * plane + (size_t)(y0 + r + 1) * Wp + (x0 + xl + 1);
    store_out_row(out + pix * kC, acc, bco, hf);
  }
}

// ---- output layer (32 -> 1), FP32 on the vector units ------------------------

// out[b][y][x] = b6 + sum_{tap,ci} w6[tap][ci] * a5[b][y+1+dy][x+1+dx][ci].
__global__ __launch_bounds__(256) void out_layer_kernel(
    const bf16* __restrict__ a5, const float* __restrict__ w6,
    const float* __restrict__ bias, float* __restrict__ out, int B, int H,
    int W) {
  const float b6 = bias[5 * kC];
  // Rows in one contiguous range per XCD (xcd_span): the rows above and
  // below a block's pixels are then in its XCD's L2.
  const long idx = (long)xcd_span(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const long total = (long)B * H * W;
  if (idx >= total) return;
  const int x = (int)(idx % W);
  const int y = (int)((idx / W) % H);
  const int b = (int)(idx / ((long)W * H));
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  float acc = 0.0f;
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const bf16* p = a5 + ((size_t)b * plane + (size_t)(y + 1 + dy) * Wp + (x + 1 + dx)) * kC;
#pragma unroll
    for (int c8 = 0; c8 < kC; c8 += 8) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(p + c8);
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += w6[tap * kC + c8 + q] * (float)v[q];
    }
  }
  out[idx] = acc + b6;
}

// ---- loss: mean squared error, and its gradient -----------------------------

// dout = 2 (out - t) / n; per-block partial sums of (out - t)^2 in FP32.
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ out,
                                                  const float* __restrict__ t,
                                                  float* __restrict__ dout,
                                                  float* __restrict__ partial,
                                                  long n) {
  __shared__ float s[256];
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  float e2 = 0.0f;
  if (i < n) {
    const float d = out[i] - t[i];
    dout[i] = 2.0f * d / (float)n;
    e2 = d * d;
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// ---- output layer backward ---------------------------------------------------

// delta5[p][ci] = (a5[p][ci] > 0) * sum_tap dout[p - tap] * w6[tap][ci].
__global__ __launch_bounds__(256) void out_layer_bwd_data_kernel(
    const float* __restrict__ dout, const float* __restrict__ w6,
    const bf16* __restrict__ a5, bf16* __restrict__ d5, int B, int H, int W) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)B * H * W;
  if (idx >= total) return;
  const int x = (int)(idx % W);
  const int y = (int)((idx / W) % H);
  const int b = (int)(idx / ((long)W * H));
  float g[kC];
#pragma unroll
  for (int c = 0; c < kC; ++c) g[c] = 0.0f;
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int ys = y - dy, xs = x - dx;  // output pixel that reads p at tap
    if (ys < 0 || ys >= H || xs < 0 || xs >= W) continue;
    const float d = dout[((long)b * H + ys) * W + xs];
#pragma unroll
    for (int c = 0; c < kC; ++c) g[c] += d * w6[tap * kC + c];
  }
  const int Wp = W + 2;
  const size_t pix = (size_t)b * (H + 2) * Wp + (size_t)(y + 1) * Wp + (x + 1);
#pragma unroll
  for (int c8 = 0; c8 < kC; c8 += 8) {
    const bf16x8 m = *reinterpret_cast<const bf16x8*>(a5 + pix * kC + c8);
    bf16x8 o;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = (bf16)((float)m[q] > 0.0f ? g[c8 + q] : 0.0f);
    *reinterpret_cast<bf16x8*>(d5 + pix * kC + c8) = o;
  }
}

// Partial dW6[tap][ci] and db6 per block of rows: thread t < 288 handles
// (tap, ci) = (t / 32, t % 32); thread 288 accumulates db6.  Grid: nblk
// blocks, each handling rows [blk*rows_per, ...) of the B*H rows.
__global__ __launch_bounds__(320) void out_layer_wgrad_kernel(
    const float* __restrict__ dout, const bf16* __restrict__ a5,
    float* __restrict__ partial, int B, int H, int W, int rows_per) {
  const int t = threadIdx.x;
  if (t > kW6) return;
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  float acc = 0.0f;
  const int r0 = blockIdx.x * rows_per;
  const int r1 = min(r0 + rows_per, B * H);
  const int tap = t / kC, ci = t % kC;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  for (int row = r0; row < r1; ++row) {
    const int b = row / H, y = row % H;
    const float* d = dout + (long)row * W;
    if (t == kW6) {
      for (int x = 0; x < W; ++x) acc += d[x];
    } else {
      const bf16* a = a5 + ((size_t)b * plane + (size_t)(y + 1 + dy) * Wp + 1 + dx) * kC + ci;
      for (int x = 0; x < W; ++x) acc += d[x] * (float)a[(size_t)x * kC];
    }
  }
  partial[(size_t)blockIdx.x * (kW6 + 1) + t] = acc;
}

// ---- hidden-layer weight gradient on MFMA ------------------------------------
//
// dW[co][n] = sum over pixels p of delta[p][co] * In[p + tap(n)][ci(n)],
// n = tap*CIN + ci, and db[co] = sum_p delta[p][co].  One block per group of
// rows; the delta row and the three input rows it needs are staged in LDS,
// then each wave accumulates its share of the 32-wide n blocks over all
// pixels with MFMA steps of 16 pixels (A: lane co, 8 pixels; B: lane n,
// the same 8 pixels).  Per-block partials are summed by wgrad_reduce.
template <int CIN>
__global__ __launch_bounds__(256) void wgrad_mfma_kernel(
    const bf16* __restrict__ delta, const bf16* __restrict__ in,
    float* __restrict__ partial, int B, int H, int W, int rows_per) {
  constexpr int N = kTaps * CIN;         // 288 or 72
  constexpr int NB = (N + 31) / 32;      // 9 or 3 blocks of 32
  constexpr int kMaxW = kWgradMaxW;      // staged row width limit
  __shared__ bf16 sd[kMaxW * kC];        // delta row
  __shared__ bf16 sa[3 * (kMaxW + 2) * CIN];  // input rows y-1..y+1
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int hf = lane >> 5, col = lane & 31;
  const int Wp = W + 2;
  const size_t plane = (size_t)(H + 2) * Wp;
  f32x16 acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) acc[q] = f32x16{};
  float dbias = 0.0f;
  const int r0 = blockIdx.x * rows_per;
  const int r1 = min(r0 + rows_per, B * H);
  for (int row = r0; row < r1; ++row) {
    const int b = row / H, y = row % H;
    __syncthreads();
    // Stage delta[b][y+1][1..W][*] and in[b][y..y+2][0..W+1][*].
    for (int i = threadIdx.x; i < W * kC / 8; i += 256) {
      const int x = i / (kC / 8), c8 = (i % (kC / 8)) * 8;
      *reinterpret_cast<bf16x8*>(sd + x * kC + c8) = *reinterpret_cast<const bf16x8*>(
          delta + ((size_t)b * plane + (size_t)(y + 1) * Wp + x + 1) * kC + c8);
    }
    for (int i = threadIdx.x; i < 3 * Wp * CIN / 8; i += 256) {
      const int rr = i / (Wp * CIN / 8), rem = i % (Wp * CIN / 8);
      const int x = rem / (CIN / 8), c8 = (rem % (CIN / 8)) * 8;
      *reinterpret_cast<bf16x8*>(sa + (rr * Wp + x) * CIN + c8) =
          *reinterpret_cast<const bf16x8*>(in + ((size_t)b * plane + (size_t)(y + rr) * Wp + x) * CIN + c8);
    }
    __syncthreads();
    for (int x0 = 0; x0 < W; x0 += 16) {
      bf16x8 a;
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = sd[(x0 + 8 * hf + s) * kC + col];
      if (wv == 0) {
#pragma unroll
        for (int s = 0; s < 8; ++s) dbias += (float)a[s];
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int nb = wv + 4 * q;
        if (nb >= NB) break;
        const int n = nb * 32 + col;
        bf16x8 bv;
        if (n < N) {
          const int tap = n / CIN, ci = n % CIN;
          const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
          for (int s = 0; s < 8; ++s)
            bv[s] = sa[((1 + dy) * Wp + (x0 + 8 * hf + s + 1 + dx)) * CIN + ci];
        } else {
          bv = bf16x8{};
        }
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[q], 0, 0, 0);
      }
    }
  }
  // D[co][n]: lane l holds n = nb*32 + l%32 and co = 8*(v/4) + 4*hf + v%4.
  float* pb = partial + (size_t)blockIdx.x * (kC * N + kC);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int nb = wv + 4 * q;
    if (nb >= NB) break;
    const int n = nb * 32 + col;
    if (n >= N) continue;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int co = 8 * (v >> 2) + 4 * hf + (v & 3);
      pb[co * N + n] = acc[q][v];
    }
  }
  if (wv == 0) {
    // Lanes l and l+32 hold the two halves of channel l%32.
    const float other = __shfl_xor(dbias, 32);
    if (hf == 0) pb[kC * N + col] = dbias + other;
  }
}

// Sums nblk partial rows of `len` floats into `out` (fixed order).
__global__ __launch_bounds__(256) void partial_sum_kernel(
    const float* __restrict__ partial, int nblk, int len, float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * len + i];
  out[i] = s;
}

// Scatters a layer's summed wgrad ([co][tap][ci_pad] then db) into the
// master-layout gradient ([co][ci][ky][kx], then b).
__global__ __launch_bounds__(256) void scatter_grad_kernel(
    const float* __restrict__ sums, int cin_pad, int cin, int cout,
    float* __restrict__ gw, float* __restrict__ gb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nw = cout * kTaps * cin_pad;
  if (i < nw) {
    const int co = i / (kTaps * cin_pad), rem = i % (kTaps * cin_pad);
    const int tap = rem / cin_pad, ci = rem % cin_pad;
    if (ci < cin) gw[(co * cin + ci) * kTaps + tap] = sums[i];
  } else if (i < nw + cout) {
    gb[i - nw] = sums[i];
  }
}

// Output layer: partial sums are [tap][ci] then db6.
__global__ __launch_bounds__(320) void scatter_grad6_kernel(
    const float* __restrict__ sums, float* __restrict__ gw, float* __restrict__ gb) {
  const int t = threadIdx.x;
  if (t < kW6) {
    const int tap = t / kC, ci = t % kC;
    gw[ci * kTaps + tap] = sums[t];
  } else if (t == kW6) {
    gb[0] = sums[t];
  }
}

// ---- parameters ------------------------------------------------------------

// BF16 kernel layouts from the FP32 master: forward Wf[co][tap][ci_pad],
// backward-data Wb[ci][tap][co] = W[co][ci][flip(tap)], output-layer w6
// [tap][ci] (FP32) and the biases.
__global__ __launch_bounds__(256) void pack_params_kernel(
    const float* __restrict__ master, bf16* __restrict__ wf1,
    bf16* __restrict__ wfh, bf16* __restrict__ wbh, float* __restrict__ w6,
    float* __restrict__ bias) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  // Layer 1: [co][ci<3][ky][kx] -> wf1[co][tap][ci pad 8].
  if (i < kC * kTaps * kC0) {
    const int co = i / (kTaps * kC0), rem = i % (kTaps * kC0);
    const int tap = rem / kC0, ci = rem % kC0;
    wf1[i] = (bf16)(ci < kIn ? master[(co * kIn + ci) * kTaps + tap] : 0.0f);
  }
  // Hidden layers 2..5.
  if (i < 4 * kWh) {
    const int l = i / kWh, rem = i % kWh;
    const float* W = master + w_offset(l + 1);
    {
      const int co = rem / (kTaps * kC), r2 = rem % (kTaps * kC);
      const int tap = r2 / kC, ci = r2 % kC;
      wfh[i] = (bf16)W[(co * kC + ci) * kTaps + tap];
    }
    {
      const int ci = rem / (kTaps * kC), r2 = rem % (kTaps * kC);
      const int tap = r2 / kC, co = r2 % kC;
      wbh[i] = (bf16)W[(co * kC + ci) * kTaps + (kTaps - 1 - tap)];
    }
  }
  if (i < kW6) {
    const int tap = i / kC, ci = i % kC;
    w6[i] = (float)(bf16)master[w_offset(5) + ci * kTaps + tap];
  }
  if (i < 5 * kC + 1) bias[i] = i < 5 * kC ? master[b_offset(i / kC) + i % kC] : master[b_offset(5)];
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ m,
                                                   float* __restrict__ v, int n,
                                                   float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Input tensor from FP32 planes [b][y][x][3] (already scaled).
__global__ __launch_bounds__(256) void pack_input_kernel(const float* __restrict__ src,
                                                         bf16* __restrict__ dst, int B,
                                                         int H, int W) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)B * H * W) return;
  const int x = (int)(idx % W), y = (int)((idx / W) % H), b = (int)(idx / ((long)W * H));
  const size_t pix = (size_t)b * (H + 2) * (W + 2) + (size_t)(y + 1) * (W + 2) + (x + 1);
  bf16x8 o = bf16x8{};
  for (int c = 0; c < kIn; ++c) o[c] = (bf16)src[idx * kIn + c];
  *reinterpret_cast<bf16x8*>(dst + pix * kC0) = o;
}

}  // namespace nn
}  // namespace tea

// ---- host side -----------------------------------------------------------------

using namespace tea;
using namespace tea::nn;

struct tea_nn {
  int device = 0;
  int max_batch = 0, H = 0, W = 0;
  hipStream_t stream = nullptr;
  int fwd_rows = kFwdRows;  // rows per block of the forward kernel (0: one-row kernel)
  // parameters
  float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  long adam_t = 0;
  bf16 *wf1 = nullptr, *wfh = nullptr, *wbh = nullptr;
  float *w6 = nullptr, *bias = nullptr;
  // activations and gradients
  float* in_f32 = nullptr;   // [B][H][W][3]
  bf16* in0 = nullptr;       // [B][H+2][W+2][8]
  bf16* act[5] = {};         // a1..a5 [B][H+2][W+2][32]
  bf16* del[5] = {};         // delta1..delta5
  float *out = nullptr, *target = nullptr, *dout = nullptr;
  float *partial = nullptr, *sums = nullptr, *loss_part = nullptr;
  size_t partial_len = 0;
  int wg_blocks = 0;
  // device-resident training set
  float *ds_in = nullptr, *ds_tgt = nullptr;
  int ds_n = 0;
  int* ds_order = nullptr;
};

namespace {

size_t act_elems(const tea_nn* n, int c) {
  return (size_t)n->max_batch * (n->H + 2) * (n->W + 2) * c;
}

int nn_alloc(void** p, size_t bytes, hipStream_t st) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc failed in the network workspace");
    return TEA_HIP_ENOMEM;
  }
  TEA_HIP_TRY(hipMemsetAsync(*p, 0, bytes, st));
  return TEA_HIP_OK;
}

int pack_params(tea_nn* n) {
  hipLaunchKernelGGL(pack_params_kernel, dim3((4 * kWh + 255) / 256), dim3(256), 0,
                     n->stream, n->master, n->wf1, n->wfh, n->wbh, n->w6, n->bias);
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

dim3 rows_grid(const tea_nn* n, int batch) { return dim3(n->H, batch); }
unsigned flat_blocks(long count) { return (unsigned)((count + 255) / 256); }

// Forward pass over `batch` samples whose packed input is in n->in0, on
// stream `st` (the network's own, or a solver's); without the output layer
// unless with_output (its caller then runs it on the last activations); from
// hidden layer `first` on (0-based: the caller has written the activations
// of the layers before it).
int forward_on(tea_nn* n, int batch, hipStream_t st, bool with_output = true,
               int first = 0) {
  const int H = n->H, W = n->W;
  // Several rows per block only while the grid keeps at least 512 blocks:
  // on small grids the one-row kernel's parallelism matters more than the
  // saved weight and row loads.
  const int R = n->fwd_rows;
  if (W % 128 == 0 && (R == 2 || R == 4) && H % R == 0 &&
      (long)(W / 128) * (H / R) * batch >= 512) {
    const dim3 grid(W / 128, H / R, batch);
    if (first > 0)
      ;
    else if (R == 4)
      hipLaunchKernelGGL((conv_fwd_rows_kernel<kC0, 4>), grid, dim3(256), 0, st, n->in0,
                         n->wf1, n->bias, n->act[0], H, W);
    else
      hipLaunchKernelGGL((conv_fwd_rows_kernel<kC0, 2>), grid, dim3(256), 0, st, n->in0,
                         n->wf1, n->bias, n->act[0], H, W);
    for (int l = first > 1 ? first : 1; l < 5; ++l) {
      const bf16* w = n->wfh + (size_t)(l - 1) * kWh;
      if (R == 4)
        hipLaunchKernelGGL((conv_fwd_rows_kernel<kC, 4>), grid, dim3(256), 0, st,
                           n->act[l - 1], w, n->bias + l * kC, n->act[l], H, W);
      else
        hipLaunchKernelGGL((conv_fwd_rows_kernel<kC, 2>), grid, dim3(256), 0, st,
                           n->act[l - 1], w, n->bias + l * kC, n->act[l], H, W);
    }
  } else if (W % 128 == 0) {
    const dim3 grid(W / 128, H, batch);
    if (first == 0)
      hipLaunchKernelGGL((conv_fwd_lds_kernel<kC0>), grid, dim3(256), 0, st, n->in0, n->wf1,
                         n->bias, n->act[0], H, W);
    for (int l = first > 1 ? first : 1; l < 5; ++l)
      hipLaunchKernelGGL((conv_fwd_lds_kernel<kC>), grid, dim3(256), 0, st, n->act[l - 1],
                         n->wfh + (size_t)(l - 1) * kWh, n->bias + l * kC, n->act[l], H, W);
  } else {
    if (first == 0)
      hipLaunchKernelGGL((conv_mfma_kernel<kC0, 0>), rows_grid(n, batch), dim3(256), 0,
                         st, n->in0, n->wf1, n->bias, nullptr, n->act[0], H, W, batch);
    for (int l = first > 1 ? first : 1; l < 5; ++l)
      hipLaunchKernelGGL((conv_mfma_kernel<kC, 0>), rows_grid(n, batch), dim3(256), 0,
                         st, n->act[l - 1], n->wfh + (size_t)(l - 1) * kWh,
                         n->bias + l * kC, nullptr, n->act[l], H, W, batch);
  }
  if (with_output)
    hipLaunchKernelGGL(out_layer_kernel, dim3(flat_blocks((long)batch * H * W)), dim3(256),
                       0, st, n->act[4], n->w6, n->bias, n->out, batch, H, W);
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

int forward(tea_nn* n, int batch) { return forward_on(n, batch, n->stream); }

int wgrad_layer(tea_nn* n, int batch, int layer, const bf16* delta, const bf16* in) {
  const int rows = batch * n->H;
  const int rows_per = (rows + n->wg_blocks - 1) / n->wg_blocks;
  const int nblk = (rows + rows_per - 1) / rows_per;
  const int cin_pad = layer == 0 ? kC0 : kC;
  const int len = kC * kTaps * cin_pad + kC;
  if (layer == 0)
    hipLaunchKernelGGL((wgrad_mfma_kernel<kC0>), dim3(nblk), dim3(256), 0, n->stream,
                       delta, in, n->partial, batch, n->H, n->W, rows_per);
  else
    hipLaunchKernelGGL((wgrad_mfma_kernel<kC>), dim3(nblk), dim3(256), 0, n->stream,
                       delta, in, n->partial, batch, n->H, n->W, rows_per);
  hipLaunchKernelGGL(partial_sum_kernel, dim3((len + 255) / 256), dim3(256), 0, n->stream,
                     n->partial, nblk, len, n->sums);
  const int cin = layer == 0 ? kIn : kC;
  hipLaunchKernelGGL(scatter_grad_kernel, dim3((len + 255) / 256), dim3(256), 0,
                     n->stream, n->sums, cin_pad, cin, kC, n->grad + w_offset(layer),
                     n->grad + b_offset(layer));
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// Loss and gradients for `batch` samples (input packed, targets in place).
int loss_and_grad(tea_nn* n, int batch, float* loss) {
  const int H = n->H, W = n->W;
  TEA_RETURN_IF(forward(n, batch));
  const long cnt = (long)batch * H * W;
  const unsigned nb = flat_blocks(cnt);
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, n->stream, n->out, n->target,
                     n->dout, n->loss_part, cnt);
  // Output layer.
  hipLaunchKernelGGL(out_layer_bwd_data_kernel, dim3(nb), dim3(256), 0, n->stream,
                     n->dout, n->w6, n->act[4], n->del[4], batch, H, W);
  {
    const int rows = batch * H;
    const int rows_per = (rows + n->wg_blocks - 1) / n->wg_blocks;
    const int nblk = (rows + rows_per - 1) / rows_per;
    hipLaunchKernelGGL(out_layer_wgrad_kernel, dim3(nblk), dim3(320), 0, n->stream,
                       n->dout, n->act[4], n->partial, batch, H, W, rows_per);
    hipLaunchKernelGGL(partial_sum_kernel, dim3((kW6 + 1 + 255) / 256), dim3(256), 0,
                       n->stream, n->partial, nblk, kW6 + 1, n->sums);
    hipLaunchKernelGGL(scatter_grad6_kernel, dim3(1), dim3(320), 0, n->stream, n->sums,
                       n->grad + w_offset(5), n->grad + b_offset(5));
  }
  // Hidden layers 5..2, then layer 1.
  for (int l = 4; l >= 1; --l) {
    TEA_RETURN_IF(wgrad_layer(n, batch, l, n->del[l], n->act[l - 1]));
    hipLaunchKernelGGL((conv_mfma_kernel<kC, 1>), rows_grid(n, batch), dim3(256), 0,
                       n->stream, n->del[l], n->wbh + (size_t)(l - 1) * kWh, nullptr,
                       n->act[l - 1], n->del[l - 1], H, W, batch);
  }
  TEA_RETURN_IF(wgrad_layer(n, batch, 0, n->del[0], n->in0));
  TEA_HIP_TRY(hipGetLastError());
  std::vector<float> lp(nb);
  TEA_HIP_TRY(hipMemcpyAsync(lp.data(), n->loss_part, nb * sizeof(float),
                             hipMemcpyDeviceToHost, n->stream));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  double s = 0.0;
  for (float v : lp) s += v;
  if (loss) *loss = (float)(s / (double)cnt);
  return TEA_HIP_OK;
}

int adam_step(tea_nn* n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->adam_t += 1;
  // Bias corrections 1 - beta^t of the FP32 betas, formed in double: in FP32,
  // 1 - powf(0.999f, t) loses about 1e-5 of its value to cancellation.
  const float c1 = (float)(1.0 - pow((double)b1, (double)n->adam_t));
  const float c2 = (float)(1.0 - pow((double)b2, (double)n->adam_t));
  hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0,
                     n->stream, n->master, n->grad, n->adam_m, n->adam_v, kNumParams,
                     lr, b1, b2, eps, c1, c2);
  TEA_HIP_TRY(hipGetLastError());
  return pack_params(n);
}

int upload_batch(tea_nn* n, int batch, const float* input, const float* target) {
  const long px = (long)batch * n->H * n->W;
  TEA_HIP_TRY(hipMemcpyAsync(n->in_f32, input, px * kIn * sizeof(float),
                             hipMemcpyHostToDevice, n->stream));
  if (target)
    TEA_HIP_TRY(hipMemcpyAsync(n->target, target, px * sizeof(float),
                               hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(pack_input_kernel, dim3(flat_blocks(px)), dim3(256), 0, n->stream,
                     n->in_f32, n->in0, batch, n->H, n->W);
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// Gathers samples order[0..batch) of the resident training set.
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ ds_in,
                                                     const float* __restrict__ ds_tgt,
                                                     const int* __restrict__ order,
                                                     float* __restrict__ in,
                                                     float* __restrict__ tgt,
                                                     int batch, int px) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)batch * px) return;
  const int b = (int)(i / px), p = (int)(i % px);
  const long src = (long)order[b] * px + p;
  tgt[i] = ds_tgt[src];
  for (int c = 0; c < kIn; ++c) in[i * kIn + c] = ds_in[src * kIn + c];
}

#define NN_CHECK(n) TEA_CHECK_ARG((n) != nullptr, "null network handle")

}  // namespace

namespace tea {
namespace nn {
// Used by the flexible-CG preconditioner (fcg.hip).
int forward_on_stream(tea_nn* n, hipStream_t st) { return forward_on(n, 1, st); }
int forward_hidden_on_stream(tea_nn* n, hipStream_t st, int first) {
  return forward_on(n, 1, st, false, first);
}
const bf16* last_hidden(tea_nn* n) { return n->act[4]; }
bf16* hidden(tea_nn* n, int l) { return n->act[l]; }
bf16* input_tensor(tea_nn* n) { return n->in0; }
PackedWeights packed_weights(const tea_nn* n) {
  return PackedWeights{n->wf1, n->wfh, n->w6, n->bias};
}
const float* output(tea_nn* n) { return n->out; }
int sync(tea_nn* n) {
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}
}  // namespace nn
}  // namespace tea

extern "C" {

int tea_nn_num_params(void) { return kNumParams; }

int tea_nn_create(int device, int max_batch, int height, int width, tea_nn** out) {
  TEA_CHECK_ARG(out != nullptr, "null output handle");
  *out = nullptr;
  TEA_CHECK_ARG(max_batch >= 1 && height >= 1 && width >= 32 && width % 32 == 0,
                "need batch >= 1, height >= 1 and width a multiple of 32");
  TEA_HIP_TRY(hipSetDevice(device));
  tea_nn* n = new tea_nn();
  n->device = device;
  n->max_batch = max_batch;
  n->H = height;
  n->W = width;
  if (const char* v = getenv("TEA_NN_FWD_ROWS")) n->fwd_rows = atoi(v);
  auto fail = [&](int code) {
    tea_nn_destroy(n);
    return code;
  };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreateWithFlags failed");
    return fail(TEA_HIP_EHIP);
  }
  const size_t px = (size_t)max_batch * height * width;
  n->wg_blocks = 256;
  n->partial_len = (size_t)n->wg_blocks * (kC * kTaps * kC + kC);
  struct {
    void** p;
    size_t bytes;
  } bufs[] = {
      {(void**)&n->master, kNumParams * sizeof(float)},
      {(void**)&n->grad, kNumParams * sizeof(float)},
      {(void**)&n->adam_m, kNumParams * sizeof(float)},
      {(void**)&n->adam_v, kNumParams * sizeof(float)},
      {(void**)&n->wf1, (size_t)kC * kTaps * kC0 * sizeof(bf16)},
      {(void**)&n->wfh, (size_t)4 * kWh * sizeof(bf16)},
      {(void**)&n->wbh, (size_t)4 * kWh * sizeof(bf16)},
      {(void**)&n->w6, kW6 * sizeof(float)},
      {(void**)&n->bias, (5 * kC + 1) * sizeof(float)},
      {(void**)&n->in_f32, px * kIn * sizeof(float)},
      {(void**)&n->in0, act_elems(n, kC0) * sizeof(bf16)},
      {(void**)&n->out, px * sizeof(float)},
      {(void**)&n->target, px * sizeof(float)},
      {(void**)&n->dout, px * sizeof(float)},
      {(void**)&n->partial, n->partial_len * sizeof(float)},
      {(void**)&n->sums, (size_t)(kC * kTaps * kC + kC) * sizeof(float)},
      {(void**)&n->loss_part, (px / 256 + 2) * sizeof(float)},
  };
  for (auto& b : bufs) {
    const int st = nn_alloc(b.p, b.bytes, n->stream);
    if (st != TEA_HIP_OK) return fail(st);
  }
  for (int l = 0; l < 5; ++l) {
    int st = nn_alloc((void**)&n->act[l], act_elems(n, kC) * sizeof(bf16), n->stream);
    if (st == TEA_HIP_OK && width <= kWgradMaxW)
      st = nn_alloc((void**)&n->del[l], act_elems(n, kC) * sizeof(bf16), n->stream);
    if (st != TEA_HIP_OK) return fail(st);
  }
  if (hipStreamSynchronize(n->stream) != hipSuccess) {
    set_error("hipStreamSynchronize failed");
    return fail(TEA_HIP_EHIP);
  }
  *out = n;
  return TEA_HIP_OK;
}

int tea_nn_destroy(tea_nn* n) {
  if (!n) return TEA_HIP_OK;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  void* ptrs[] = {n->master, n->grad, n->adam_m, n->adam_v, n->wf1, n->wfh, n->wbh,
                  n->w6, n->bias, n->in_f32, n->in0, n->out, n->target, n->dout,
                  n->partial, n->sums, n->loss_part, n->ds_in, n->ds_tgt, n->ds_order};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int l = 0; l < 5; ++l) {
    if (n->act[l]) (void)hipFree(n->act[l]);
    if (n->del[l]) (void)hipFree(n->del[l]);
  }
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return TEA_HIP_OK;
}

int tea_nn_set_params(tea_nn* n, const float* host, size_t count) {
  NN_CHECK(n);
  TEA_CHECK_ARG(host && count == (size_t)kNumParams, "need exactly tea_nn_num_params() values");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_HIP_TRY(hipMemcpyAsync(n->master, host, count * sizeof(float),
                             hipMemcpyHostToDevice, n->stream));
  TEA_HIP_TRY(hipMemsetAsync(n->adam_m, 0, count * sizeof(float), n->stream));
  TEA_HIP_TRY(hipMemsetAsync(n->adam_v, 0, count * sizeof(float), n->stream));
  n->adam_t = 0;
  TEA_RETURN_IF(pack_params(n));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}

int tea_nn_get_params(tea_nn* n, float* host, size_t count) {
  NN_CHECK(n);
  TEA_CHECK_ARG(host && count == (size_t)kNumParams, "need exactly tea_nn_num_params() values");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_HIP_TRY(hipMemcpyAsync(host, n->master, count * sizeof(float),
                             hipMemcpyDeviceToHost, n->stream));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}

int tea_nn_forward(tea_nn* n, int batch, const float* input, float* output) {
  NN_CHECK(n);
  TEA_CHECK_ARG(batch >= 1 && batch <= n->max_batch, "batch out of range");
  TEA_CHECK_ARG(input && output, "null arrays");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_RETURN_IF(upload_batch(n, batch, input, nullptr));
  TEA_RETURN_IF(forward(n, batch));
  TEA_HIP_TRY(hipMemcpyAsync(output, n->out, (size_t)batch * n->H * n->W * sizeof(float),
                             hipMemcpyDeviceToHost, n->stream));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}

int tea_nn_loss_grad(tea_nn* n, int batch, const float* input, const float* target,
                     float* loss, float* grad) {
  NN_CHECK(n);
  TEA_CHECK_ARG(batch >= 1 && batch <= n->max_batch, "batch out of range");
  TEA_CHECK_ARG(input && target, "null arrays");
  TEA_CHECK_ARG(n->del[0] != nullptr, "training needs width <= 128");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_RETURN_IF(upload_batch(n, batch, input, target));
  TEA_RETURN_IF(loss_and_grad(n, batch, loss));
  if (grad) {
    TEA_HIP_TRY(hipMemcpyAsync(grad, n->grad, kNumParams * sizeof(float),
                               hipMemcpyDeviceToHost, n->stream));
    TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  }
  return TEA_HIP_OK;
}

int tea_nn_train_step(tea_nn* n, int batch, const float* input, const float* target,
                      float lr, float* loss) {
  TEA_RETURN_IF(tea_nn_loss_grad(n, batch, input, target, loss, nullptr));
  TEA_RETURN_IF(adam_step(n, lr));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}

int tea_nn_adam_step(tea_nn* n, const float* grad, float lr) {
  NN_CHECK(n);
  TEA_CHECK_ARG(grad != nullptr, "null gradient");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_HIP_TRY(hipMemcpyAsync(n->grad, grad, kNumParams * sizeof(float),
                             hipMemcpyHostToDevice, n->stream));
  TEA_RETURN_IF(adam_step(n, lr));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  return TEA_HIP_OK;
}

int tea_nn_get_adam_state(tea_nn* n, float* m, float* v, size_t count, long long* step) {
  NN_CHECK(n);
  TEA_CHECK_ARG(m && v && step && count == (size_t)kNumParams,
                "need two arrays of tea_nn_num_params() values and a step counter");
  TEA_HIP_TRY(hipSetDevice(n->device));
  TEA_HIP_TRY(hipMemcpyAsync(m, n->adam_m, count * sizeof(float), hipMemcpyDeviceToHost,
                             n->stream));
  TEA_HIP_TRY(hipMemcpyAsync(v, n->adam_v, count * sizeof(float), hipMemcpyDeviceToHost,
                             n->stream));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  *step = n->adam_t;
  return TEA_HIP_OK;
}

int tea_nn_load_dataset(tea_nn* n, int count, const float* inputs, const float* targets) {
  NN_CHECK(n);
  TEA_CHECK_ARG(count >= 1 && inputs && targets, "empty dataset");
  TEA_HIP_TRY(hipSetDevice(n->device));
  if (n->ds_in) (void)hipFree(n->ds_in);
  if (n->ds_tgt) (void)hipFree(n->ds_tgt);
  if (n->ds_order) (void)hipFree(n->ds_order);
  n->ds_in = n->ds_tgt = nullptr;
  n->ds_order = nullptr;
  const size_t px = (size_t)n->H * n->W;
  TEA_RETURN_IF(nn_alloc((void**)&n->ds_in, (size_t)count * px * kIn * sizeof(float), n->stream));
  TEA_RETURN_IF(nn_alloc((void**)&n->ds_tgt, (size_t)count * px * sizeof(float), n->stream));
  TEA_RETURN_IF(nn_alloc((void**)&n->ds_order, (size_t)count * sizeof(int), n->stream));
  TEA_HIP_TRY(hipMemcpyAsync(n->ds_in, inputs, (size_t)count * px * kIn * sizeof(float),
                             hipMemcpyHostToDevice, n->stream));
  TEA_HIP_TRY(hipMemcpyAsync(n->ds_tgt, targets, (size_t)count * px * sizeof(float),
                             hipMemcpyHostToDevice, n->stream));
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  n->ds_n = count;
  return TEA_HIP_OK;
}

// One pass over order[0..count) in batches of `batch` (a short last batch is
// dropped); Adam with learning rate lr; mean batch loss to *mean_loss.
int tea_nn_train_epoch(tea_nn* n, const int* order, int count, int batch, float lr,
                       float* mean_loss) {
  NN_CHECK(n);
  TEA_CHECK_ARG(n->ds_in != nullptr, "no dataset loaded");
  TEA_CHECK_ARG(n->del[0] != nullptr, "training needs width <= 128");
  TEA_CHECK_ARG(batch >= 1 && batch <= n->max_batch && count >= batch, "bad batch");
  TEA_HIP_TRY(hipSetDevice(n->device));
  for (int i = 0; i < count; ++i)
    TEA_CHECK_ARG(order[i] >= 0 && order[i] < n->ds_n, "sample index out of range");
  TEA_HIP_TRY(hipMemcpyAsync(n->ds_order, order, (size_t)count * sizeof(int),
                             hipMemcpyHostToDevice, n->stream));
  const int px = n->H * n->W;
  double total = 0.0;
  int steps = 0;
  for (int s = 0; s + batch <= count; s += batch) {
    hipLaunchKernelGGL(gather_kernel, dim3(flat_blocks((long)batch * px)), dim3(256), 0,
                       n->stream, n->ds_in, n->ds_tgt, n->ds_order + s, n->in_f32,
                       n->target, batch, px);
    hipLaunchKernelGGL(pack_input_kernel, dim3(flat_blocks((long)batch * px)), dim3(256),
                       0, n->stream, n->in_f32, n->in0, batch, n->H, n->W);
    float loss = 0.0f;
    TEA_RETURN_IF(loss_and_grad(n, batch, &loss));
    TEA_RETURN_IF(adam_step(n, lr));
    total += loss;
    ++steps;
  }
  TEA_HIP_TRY(hipStreamSynchronize(n->stream));
  if (mean_loss) *mean_loss = (float)(total / (steps ? steps : 1));
  return TEA_HIP_OK;
}

}  // extern "C"
