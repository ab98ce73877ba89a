// libteahip.so: flexible CG with the learned preconditioner, and
// Jacobi-preconditioned CG (no counterpart in the reference).
//
// Storage and precision follow the mixed-precision CG (mpcg.hip): p, r, w,
// z and the face coefficients in FP32; u, u0, dot products and scalars in
// FP64; the FP64 true-residual refresh is tea_hip_mpcg_refresh.
//
// Preconditioners, z = M(r):
//   Jacobi:  z = r / diag(A), diag = 1 + kx[i] + kx[i+1] + ky[i] + ky[i+x].
//   Learned: z = r / diag(A) + g P y,
//            y = D^-1/2 sigma N(D^-1/2 R r / sigma, f),
//            where R is the c x c block mean onto the coarse grid, T =
//            P^T A P / c^2 the Galerkin coarse operator, D its diagonal,
//            sigma the RMS of D^-1/2 R r over the 13 x 13 window around each
//            coarse cell (the receptive field: every region reaches the
//            network at unit scale), N the 6-layer network on the coarse
//            grid, P the piecewise-constant prolongation, f the couplings
//            of D^-1/2 T D^-1/2 (feature_kernel) and g = (y.Rr) / (y.T y)
//            the energy-optimal factor of the coarse direction (all
//            computed on the device).  N is trained to do what 6 sweeps of
//            TeaLeaf's Jacobi iteration on T do, so that the coarse term is a local
//            approximation of the two-level additive preconditioner
//            D^-1 + P (P^T A P)^-1 P^T.  With g,
//            z.r = r.D^-1 r + c^2 (y.Rr)^2 / (y.T y) > 0 for any y: the
//            network cannot make the flexible CG step degenerate.
// With c = 8 the network runs on 1/64 of the cells; its 13-cell receptive
// field spans 104 fine cells.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tea_hip_impl.h"

namespace tea {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// z = r / diag and z.w over the interior (span kernel body).
struct JacobiBody {
  const float* r;
  const float* kx;
  const float* ky;
  const float* w;
  float* z;
  int x;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    const double d = (1.0 + ((double)kx[i + 1] + (double)kx[i])) +
                     ((double)ky[i + x] + (double)ky[i]);
    const float zv = (float)((double)r[i] / d);
    z[i] = zv;
    out[0] = (double)zv * (double)w[i];
  }
};

// Local scale of the network input: sigma = RMS of D^-1/2 rbar over the
// (2 kWin + 1)^2 window around each coarse cell (zero beyond the edge), in
// two passes (rows into tmp, then columns), each summed in a fixed order,
// and at least kFloor times the RMS over the grid (so that regions of
// round-off do not reach the network magnified).
constexpr int kWin = 6;         // the network's receptive-field radius
constexpr float kFloor = 1e-3f;

// One coarse term of the learned preconditioner at a fine cell: g[0] times
// y at the coarse cell ((kk-h) >> shift, (jj-h) >> shift).
struct CoarseTerm {
  const float* y;
  const float* g;
  int shift, wc;
  __device__ int index(int kk, int jj, int h) const {
    return ((jj - h) >> shift) * wc + ((kk - h) >> shift);
  }
  __device__ double at(int kk, int jj, int h) const {
    return (double)g[0] * (double)y[index(kk, jj, h)];
  }
};

// z = r / diag + the NL coarse terms, and z.w (span kernel body).
template <int NL>
struct ProlongBody {
  const float* r;
  const float* kx;
  const float* ky;
  const float* w;
  CoarseTerm t0, t1;
  float* z;
  int x, h;
  __device__ void operator()(int i, int kk, int jj, double (&out)[1]) const {
    const double d = (1.0 + ((double)kx[i + 1] + (double)kx[i])) +
                     ((double)ky[i + x] + (double)ky[i]);
    double v = (double)r[i] / d;
    if (NL > 0) v = v + t0.at(kk, jj, h);
    if (NL > 1) v = v + t1.at(kk, jj, h);
    const float zv = (float)v;
    z[i] = zv;
    out[0] = (double)zv * (double)w[i];
  }
};

// The same, two cells per lane (x even): z = r / diag (+ the NL coarse
// terms) and z.w, the arithmetic of JacobiBody / ProlongBody cell by cell.
template <int NL>
__global__ __launch_bounds__(kBlock) void fcg_z2_kernel(
    Geo g, const float* __restrict__ r, const float* __restrict__ kx,
    const float* __restrict__ ky, const float* __restrict__ w, CoarseTerm t0,
    CoarseTerm t1, float* __restrict__ z, double* partials, int nspan,
    int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  const int lane = threadIdx.x & (kWave - 1);
  const int x = g.x;
  double psum[1][kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const int e = base + k * (2 * kBlock) + 2 * threadIdx.x;
    const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
    const int kk = e - row * x;
    const bool in0 = e < g.n && interior(g, kk, row);
    const bool in1 = e < g.n && interior(g, kk + 1, row);
    const bool any = in0 || in1;
    const int ic = e < g.n ? e : x + 2;
    const int ir = any ? e : x + 2;
    const float2 R = ld2(r, ir), KX = ld2(kx, ic), KY = ld2(ky, ir),
                 KYt = ld2(ky, ir + x), W = ld2(w, ir);
    float kxr = __shfl_down(KX.x, 1);
    if (lane == kWave - 1 && any) kxr = kx[e + 2];
    const double da = (1.0 + ((double)KX.y + (double)KX.x)) + ((double)KYt.x + (double)KY.x);
    const double db = (1.0 + ((double)kxr + (double)KX.y)) + ((double)KYt.y + (double)KY.y);
    float za = 0.0f, zb = 0.0f;
    if (in0) {
      double v = (double)R.x / da;
      if (NL > 0) v = v + t0.at(kk, row, g.h);
      if (NL > 1) v = v + t1.at(kk, row, g.h);
      za = (float)v;
    }
    if (in1) {
      double v = (double)R.y / db;
      if (NL > 0) v = v + t0.at(kk + 1, row, g.h);
      if (NL > 1) v = v + t1.at(kk + 1, row, g.h);
      zb = (float)v;
    }
    psum[0][k] = (in0 ? (double)za * (double)W.x : 0.0) + (in1 ? (double)zb * (double)W.y : 0.0);
    if (any) st2(z, e, in0, in1, za, zb);
  }
  span_reduce_pairs_n<1>(psum, partials, nspan, span, single_group != 0);
}

// fcg_z2_kernel's arithmetic with one pair per thread; the waves store their
// z.w sums for span_finish_kernel (see wave_subtree_pairs).  Lane 63 loads
// the right neighbour's kx with the pair's own loads.
template <int NL>
__global__ __launch_bounds__(kBlock) void fcg_z1_kernel(
    Geo g, const float* __restrict__ r, const float* __restrict__ kx,
    const float* __restrict__ ky, const float* __restrict__ w, CoarseTerm t0,
    CoarseTerm t1, float* __restrict__ z, double* sub, int nblk) {
  const int blk = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int x = g.x;
  const int e = blk * kPairBlock + 2 * threadIdx.x;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * x;
  const bool in0 = e < g.n && interior(g, kk, row);
  const bool in1 = e < g.n && interior(g, kk + 1, row);
  const bool any = in0 || in1;
  const int ic = e < g.n ? e : x + 2;
  const int ir = any ? e : x + 2;
  const float2 R = ld2(r, ir), KX = ld2(kx, ic), KY = ld2(ky, ir),
               KYt = ld2(ky, ir + x), W = ld2(w, ir);
  float kxn = 0.0f;
  if (lane == kWave - 1) kxn = kx[ir + 2];
  // The coarse values of both cells, issued with the pair's own loads (index
  // 0 for cells outside the interior, whose value is not used).
  float ya0 = 0.0f, yb0 = 0.0f, ya1 = 0.0f, yb1 = 0.0f, g0 = 0.0f, g1 = 0.0f;
  if (NL > 0) {
    g0 = t0.g[0];
    ya0 = t0.y[in0 ? t0.index(kk, row, g.h) : 0];
    yb0 = t0.y[in1 ? t0.index(kk + 1, row, g.h) : 0];
  }
  if (NL > 1) {
    g1 = t1.g[0];
    ya1 = t1.y[in0 ? t1.index(kk, row, g.h) : 0];
    yb1 = t1.y[in1 ? t1.index(kk + 1, row, g.h) : 0];
  }
  float kxr = __shfl_down(KX.x, 1);
  if (lane == kWave - 1) kxr = kxn;
  const double da = (1.0 + ((double)KX.y + (double)KX.x)) + ((double)KYt.x + (double)KY.x);
  const double db = (1.0 + ((double)kxr + (double)KX.y)) + ((double)KYt.y + (double)KY.y);
  float za = 0.0f, zb = 0.0f;
  if (in0) {
    double v = (double)R.x / da;
    if (NL > 0) v = v + (double)g0 * (double)ya0;
    if (NL > 1) v = v + (double)g1 * (double)ya1;
    za = (float)v;
  }
  if (in1) {
    double v = (double)R.y / db;
    if (NL > 0) v = v + (double)g0 * (double)yb0;
    if (NL > 1) v = v + (double)g1 * (double)yb1;
    zb = (float)v;
  }
  double v[1] = {(in0 ? (double)za * (double)W.x : 0.0) +
                 (in1 ? (double)zb * (double)W.y : 0.0)};
  if (any) st2(z, e, in0, in1, za, zb);
  wave_subtree_pairs<1>(v, sub, nblk, blk);
}

// The coarse values of a quad's four cells for one level.  With a
// coarsening factor of 4 or more (shift >= 2) the quad's interior cells lie
// in at most two coarse cells, those of its first and last interior cells
// (klo, khi), so two loads serve all four.
__device__ inline void quad_coarse(const CoarseTerm& t, int kk, int klo, int khi, int row,
                                   int h, bool any, float (&y)[4]) {
  const float ylo = t.y[any ? t.index(klo, row, h) : 0];
  const float yhi = t.y[any ? t.index(khi, row, h) : 0];
  const int xlo = (klo - h) >> t.shift;
#pragma unroll
  for (int c = 0; c < 4; ++c) y[c] = ((kk + c - h) >> t.shift) == xlo ? ylo : yhi;
}

// fcg_z1_kernel's arithmetic four cells per lane (x a multiple of 4; every
// level's coarsening factor at least 4).  Lane 63 loads the right
// neighbour's kx with the quad's own loads.
template <int NL>
__global__ __launch_bounds__(kQuadThreads) void fcg_z4_kernel(
    Geo g, const float* __restrict__ r, const float* __restrict__ kx,
    const float* __restrict__ ky, const float* __restrict__ w, CoarseTerm t0,
    CoarseTerm t1, float* __restrict__ z, double* sub, int nblk) {
  const int blk = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int x = g.x;
  const int e = blk * kQuadBlock + 4 * threadIdx.x;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * x;
  bool in[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) in[c] = e < g.n && interior(g, kk + c, row);
  const bool any = in[0] || in[1] || in[2] || in[3];
  const int ic = e < g.n ? e : x + 4;
  const int ir = any ? e : x + 4;
  const float4 R = ld4(r, ir), KX = ld4(kx, ic), KY = ld4(ky, ir), KYt = ld4(ky, ir + x),
               W = ld4(w, ir);
  float kxn = 0.0f;
  if (lane == kWave - 1) kxn = kx[ir + 4];
  const int klo = kk < g.h ? g.h : kk;
  const int khi = kk + 3 > x - g.h - 1 ? x - g.h - 1 : kk + 3;
  float y0[4] = {}, y1[4] = {};
  float g0 = 0.0f, g1 = 0.0f;
  if (NL > 0) {
    g0 = t0.g[0];
    quad_coarse(t0, kk, klo, khi, row, g.h, any, y0);
  }
  if (NL > 1) {
    g1 = t1.g[0];
    quad_coarse(t1, kk, klo, khi, row, g.h, any, y1);
  }
  float kxr = from_next_lane(KX.x);
  if (lane == kWave - 1) kxr = kxn;
  float zv[4];
  double prod[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float kx0 = q_at(KX, c), kx1 = c == 3 ? kxr : q_at(KX, c + 1);
    const double d = (1.0 + ((double)kx1 + (double)kx0)) + ((double)q_at(KYt, c) + (double)q_at(KY, c));
    zv[c] = 0.0f;
    if (in[c]) {
      double v = (double)q_at(R, c) / d;
      if (NL > 0) v = v + (double)g0 * (double)y0[c];
      if (NL > 1) v = v + (double)g1 * (double)y1[c];
      zv[c] = (float)v;
    }
    prod[c] = in[c] ? (double)zv[c] * (double)q_at(W, c) : 0.0;
  }
  double v[1] = {(prod[0] + prod[1]) + (prod[2] + prod[3])};
  if (any) st4(z, e, in, zv);
  wave_subtree_quads<1>(v, sub, nblk, blk, threadIdx.x / kWave);
}

// The coarse operator T = P^T A P / c^2: for each coarse cell, the face
// coefficients of its left and bottom faces (kc[0][..], kc[1][..]): the sum
// of the c fine face coefficients along the face over c^2, zero on the
// domain edge (the reflective boundary).
__global__ __launch_bounds__(256) void coarse_coef_kernel(Geo g, int c, int hc, int wc,
                                                          const float* __restrict__ kx,
                                                          const float* __restrict__ ky,
                                                          float* __restrict__ kc) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= hc * wc) return;
  const int X = idx % wc, Y = idx / wc;
  const int corner = (g.h + Y * c) * g.x + g.h + X * c;
  double sx = 0.0, sy = 0.0;
  for (int b = 0; b < c; ++b) {
    sx += (double)kx[corner + b * g.x];
    sy += (double)ky[corner + b];
  }
  const double cc = (double)(c * c);
  kc[idx] = X == 0 ? 0.0f : (float)(sx / cc);
  kc[hc * wc + idx] = Y == 0 ? 0.0f : (float)(sy / cc);
}

// Diagonal of T at coarse cell (X, Y): 1 + its four face coefficients.
__device__ inline double coarse_diag(const float* kc, int hc, int wc, int X, int Y) {
  const int idx = Y * wc + X;
  const float* kxc = kc;
  const float* kyc = kc + hc * wc;
  const double kr = X + 1 < wc ? (double)kxc[idx + 1] : 0.0;
  const double kt = Y + 1 < hc ? (double)kyc[idx + wc] : 0.0;
  return (1.0 + ((double)kxc[idx] + kr)) + ((double)kyc[idx] + kt);
}

// Network features: the couplings of each coarse cell to its left and
// bottom neighbours in D^-1/2 T D^-1/2, k / sqrt(D_i D_j) (zero on the
// domain edge), normalised by (m, s); also dis = D^-1/2 per coarse cell.
__global__ __launch_bounds__(256) void feature_kernel(int hc, int wc,
                                                      const float* __restrict__ kc,
                                                      float m0, float s0, float m1,
                                                      float s1,
                                                      float* __restrict__ feat,
                                                      float* __restrict__ dis) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= hc * wc) return;
  const int X = idx % wc, Y = idx / wc;
  const double di = 1.0 / sqrt(coarse_diag(kc, hc, wc, X, Y));
  double left = 0.0, bottom = 0.0;
  if (X > 0) left = (double)kc[idx] * di / sqrt(coarse_diag(kc, hc, wc, X - 1, Y));
  if (Y > 0) bottom = (double)kc[hc * wc + idx] * di / sqrt(coarse_diag(kc, hc, wc, X, Y - 1));
  feat[idx] = (float)((left - m0) / s0);
  feat[hc * wc + idx] = (float)((bottom - m1) / s1);
  dis[idx] = (float)di;
}

// ---- the coarse pipeline with Jacobi sweeps as the coarse solve ---------
// Level 0 coarsens the interior by 8, level 1 by 32.  Every sum below is
// taken in double in a fixed order, restated by the CPU oracle
// (orc_fcg_restrict8 / _restrict4 / _sweeps / _gamma), so the whole
// preconditioner is reproducible bit for bit.

// Block means of r over 8 x 8 fine blocks (level 0) and, with rbar1, of
// the level-0 means over 4 x 4 blocks (level 1).  Order: each half row of a
// block (4 cells) as ((r0 + r1) + (r2 + r3)), the 8 half rows of a half
// block pairwise ((q0 + q1) + (q2 + q3)) + ((q4 + q5) + (q6 + q7)), then the
// two halves; level 1 likewise over its 4 x 4 level-0 means.
// Grid (ceil(wc / 32), ceil(hc / 4)): wave w takes level-0 row 4 by + w,
// lane l column 32 bx + l / 2 and half l % 2 of the block's rows (two
// 8-byte loads per row when x and h are even).  With dis0 (dis1), also each
// block's sum of (D^-1/2 rbar)^2 over its cells of the level into rms0
// (rms1): the network's input scale.
template <bool kPairs>
__global__ __launch_bounds__(256) void restrict8_kernel(
    Geo g, int hc, int wc, const float* __restrict__ r, float* __restrict__ rbar0,
    float* __restrict__ rbar1, const float* __restrict__ dis0,
    const float* __restrict__ dis1, double* __restrict__ rms0, double* __restrict__ rms1) {
  __shared__ float sm[4][32];
  __shared__ double sq[4];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
  const int X = blockIdx.x * 32 + (lane >> 1), Y = blockIdx.y * 4 + wv;
  const int half = lane & 1;
  const bool ok = X < wc && Y < hc;
  // D^-1/2 of the cell for the RMS partial, loaded with the rest (a pointer
  // that is always valid stands in when there is none).
  const float* d0p = dis0 ? dis0 : r;
  const float d0 = d0p[dis0 && ok ? Y * wc + X : 0];
  double t = 0.0;
  if (ok) {
    const int base = (g.h + 8 * Y) * g.x + g.h + 8 * X + 4 * half;
    double q[8];
    // kPairs: x and h even, two 8-byte loads per row.  A template argument,
    // not a test: the two forms' last row was merged into one load after the
    // branch, waited for on its own.
    if (kPairs) {
      float2 a[8], c[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        a[b] = ld2(r, base + b * g.x);
        c[b] = ld2(r, base + b * g.x + 2);
      }
#pragma unroll
      for (int b = 0; b < 8; ++b)
        q[b] = ((double)a[b].x + (double)a[b].y) + ((double)c[b].x + (double)c[b].y);
    } else {
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const float* p = r + base + b * g.x;
        q[b] = ((double)p[0] + (double)p[1]) + ((double)p[2] + (double)p[3]);
      }
    }
    t = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
  }
  const double o = __shfl_xor(t, 1);
  const float m = ok ? (float)((half ? o + t : t + o) / 64.0) : 0.0f;
  if (ok && half == 0) rbar0[Y * wc + X] = m;
  if (rbar1 == nullptr && rms0 == nullptr) return;  // uniform over the grid
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  if (half == 0) sm[wv][lane >> 1] = m;
  double s2 = 0.0;
  if (rms0 && ok && half == 0) {
    const double v = (double)m * (double)d0;
    s2 = v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off);
  if (lane == 0) sq[wv] = s2;
  __syncthreads();
  if (threadIdx.x == 0 && rms0) rms0[bid] = (sq[0] + sq[1]) + (sq[2] + sq[3]);
  if (rbar1 && threadIdx.x < 8) {
    const int wc1 = wc >> 2, hc1 = hc >> 2;
    const int X1 = blockIdx.x * 8 + threadIdx.x, Y1 = blockIdx.y;
    double v2 = 0.0;
    if (X1 < wc1 && Y1 < hc1) {
      double rs[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float* p = &sm[j][4 * threadIdx.x];
        rs[j] = ((double)p[0] + (double)p[1]) + ((double)p[2] + (double)p[3]);
      }
      const float m1 = (float)(((rs[0] + rs[1]) + (rs[2] + rs[3])) / 16.0);
      rbar1[Y1 * wc1 + X1] = m1;
      if (rms1) {
        const double v = (double)m1 * (double)dis1[Y1 * wc1 + X1];
        v2 = v * v;
      }
    }
    if (rms1) {
      for (int off = 4; off > 0; off >>= 1) v2 += __shfl_xor(v2, off);
      if (threadIdx.x == 0) rms1[bid] = v2;
    }
  }
}

constexpr int kSweepTile = 16;
constexpr int kMaxSweeps = 8;
constexpr int kSweepR = kSweepTile + 2 * (kMaxSweeps + 1);  // staged side, at most

// `sweeps` Jacobi sweeps on T y = rbar from y = 0 (the iteration of
// TeaLeaf/c_kernels/sycl/jacobi.cpp:85-91 on the coarse operator: left,
// bottom, right and top face coefficients kl = kc[i], kb = kc[n + i],
// kr = kc[i + 1], kt = kc[n + i + wc], zero on the domain edge), in FP64,
// over one 16 x 16 tile of the level.  The block stages rbar and the face
// coefficients of the tile and the H = sweeps + 1 cells around it in LDS
// (zero beyond the grid, where y then stays zero) and takes sweep s over
// the cells within H - s of the tile, so that the last sweep leaves y
// exact on the tile and the ring around it.  Then, per tile cell: y
// rounded to FP32 into y, T y of the rounded y in the operand order of
// TeaLeaf's SMVP (shared.h:59-63), and the tile's partial sums of y.rbar
// and y.T y (the block's LDS tree) at part[tile] and part[ntiles + tile],
// tile = blockIdx.y * gridDim.x + blockIdx.x; with xsum, the sums of T y
// over the tile's 4 x 4 blocks (the level-1 cells; the cross term).
__global__ __launch_bounds__(256) void sweeps_kernel(int hc, int wc, int sweeps,
                                                     const float* __restrict__ kc,
                                                     const float* __restrict__ rbar,
                                                     float* __restrict__ y,
                                                     double* __restrict__ part, int ntiles,
                                                     double* __restrict__ xsum) {
  __shared__ float s_rb[kSweepR * kSweepR];
  __shared__ float s_kl[kSweepR * (kSweepR + 1)];  // left faces, one more column
  __shared__ float s_kb[(kSweepR + 1) * kSweepR];  // bottom faces, one more row
  __shared__ double s_y[2][kSweepR * kSweepR];
  __shared__ double s_a[256], s_b[256], s_t[256];
  const int H = sweeps + 1, R = kSweepTile + 2 * H;
  const int X0 = blockIdx.x * kSweepTile - H, Y0 = blockIdx.y * kSweepTile - H;
  const long n = (long)hc * wc;
  for (int q = threadIdx.x; q < (R + 1) * (R + 1); q += 256) {
    const int lx = q % (R + 1), ly = q / (R + 1);
    const int X = X0 + lx, Y = Y0 + ly;
    const bool in = X >= 0 && X < wc && Y >= 0 && Y < hc;
    const long i = (long)Y * wc + X;
    float kl = 0.0f, kb = 0.0f, rb = 0.0f;
    if (in) {
      kl = kc[i];
      kb = kc[n + i];
      rb = rbar[i];
    }
    if (ly < R) s_kl[ly * (R + 1) + lx] = kl;
    if (lx < R) s_kb[ly * R + lx] = kb;
    if (lx < R && ly < R) {
      s_rb[ly * R + lx] = rb;
      s_y[0][ly * R + lx] = 0.0;
    }
  }
  __syncthreads();
  int cur = 0;
  for (int s = 1; s <= sweeps; ++s) {
    const int lo = s, w = R - 2 * s;
    const double* yp = s_y[cur];
    double* yn = s_y[cur ^ 1];
    for (int q = threadIdx.x; q < w * w; q += 256) {
      const int lx = lo + q % w, ly = lo + q / w;
      const int k = ly * R + lx;
      const double kl = s_kl[ly * (R + 1) + lx], kr = s_kl[ly * (R + 1) + lx + 1];
      const double kb = s_kb[ly * R + lx], kt = s_kb[(ly + 1) * R + lx];
      yn[k] = ((double)s_rb[k] + (kr * yp[k + 1] + kl * yp[k - 1]) +
               (kt * yp[k + R] + kb * yp[k - R])) /
              ((1.0 + (kl + kr)) + (kb + kt));
    }
    __syncthreads();
    cur ^= 1;
  }
  const double* yv = s_y[cur];
  const int tx = threadIdx.x & (kSweepTile - 1), ty = threadIdx.x / kSweepTile;
  const int lx = H + tx, ly = H + ty, k = ly * R + lx;
  const int X = blockIdx.x * kSweepTile + tx, Y = blockIdx.y * kSweepTile + ty;
  auto yr = [&](int kk) { return (double)(float)yv[kk]; };
  const double kl = s_kl[ly * (R + 1) + lx], kr = s_kl[ly * (R + 1) + lx + 1];
  const double kb = s_kb[ly * R + lx], kt = s_kb[(ly + 1) * R + lx];
  const double yc = yr(k);
  const double tyv = (((1.0 + (kr + kl)) + (kt + kb)) * yc - (kr * yr(k + 1) + kl * yr(k - 1))) -
                     (kt * yr(k + R) + kb * yr(k - R));
  double a = 0.0, b = 0.0;
  if (X < wc && Y < hc) {
    y[(long)Y * wc + X] = (float)yv[k];
    a = yc * (double)s_rb[k];
    b = yc * tyv;
  }
  s_a[threadIdx.x] = a;
  s_b[threadIdx.x] = b;
  s_t[threadIdx.x] = tyv;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s_a[threadIdx.x] += s_a[threadIdx.x + st];
      s_b[threadIdx.x] += s_b[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int tile = blockIdx.y * gridDim.x + blockIdx.x;
    part[tile] = s_a[0];
    part[ntiles + tile] = s_b[0];
  }
  if (xsum && threadIdx.x < 16) {
    const int bx = threadIdx.x & 3, by = threadIdx.x >> 2;
    const int X1 = blockIdx.x * 4 + bx, Y1 = blockIdx.y * 4 + by;
    if (X1 < (wc >> 2) && Y1 < (hc >> 2)) {
      double rs[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double* p = &s_t[(4 * by + j) * kSweepTile + 4 * bx];
        rs[j] = (p[0] + p[1]) + (p[2] + p[3]);
      }
      xsum[(long)Y1 * (wc >> 2) + X1] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
    }
  }
}

// ---- the learned coarse solve of one level in one launch -----------------
// net_level_kernel computes, for one 16 x 16 tile of a coarse level, what
// box_conv_kernel (with the first two layers), the other three hidden
// layers of the network (nn.hip: conv_*_kernel), out_tile_kernel and the
// coarse dots compute for it, keeping every intermediate in LDS: the staged
// rbar and D^-1/2 (the tile and 13 cells around it), the local scale sigma and the packed BF16
// input (7 cells around it: the network's receptive field plus the ring of
// T y), and the activations of each layer on its shrinking region (28^2,
// 26^2, ..., 20^2 pixels), then N and y = N sigma D^-1/2 on the tile and its
// ring.  Each output pixel takes the same MFMA sequence (18 K-steps of 16
// over (tap, channel), weights as the A operand, the pixel's patch as B) as
// in the per-layer kernels, and the output layer is out_tile_kernel's
// (fused_out), so N is bitwise theirs; pixels outside the grid are set to zero after every layer (the
// zero ring of the per-layer kernels).  The tile's partials of y.rbar and
// y.T y, the level-1 block sums of T y and y itself are those of
// sweeps_kernel.
namespace fused {
constexpr int kT = 16;                  // output tile side
constexpr int kRN = 7;                  // network radius 6 + the ring of T y
constexpr int kS0 = kT + 2 * kRN;       // 30: network input region
constexpr int kSV = kS0 + 2 * kWin;     // 42: staged rbar, D^-1/2
constexpr int kSN = kT + 2;             // 18: network output region
constexpr int kCh = 32;                 // hidden channels
constexpr int kPx1 = (kS0 - 2) * (kS0 - 2);  // 784: layer-1 output pixels
constexpr int kPx2 = (kS0 - 4) * (kS0 - 4);  // 676
// LDS: activations A (layer 1, 3, 5 outputs) and B (input, layer 2, 4
// outputs), 32 BF16 channels per pixel in four 16-byte chunks at a pitch of
// five (act_slot), then the per-tile arrays kept to the end.
constexpr int kBytesA = kPx1 * 80;
constexpr int kBytesB = kPx2 * 80;
constexpr int kLdsBytes = kBytesA + kBytesB + 4 * kSN * kSN * 4;
static_assert(kS0 * kS0 * 16 <= kBytesB, "packed input must fit in B");
static_assert((kSV * kSV * 2 + kSV * kS0 + kS0 * kS0) * 4 <= kBytesA,
              "local-scale stage must fit in A");
static_assert(3 * 256 * 8 <= kBytesB, "reductions must fit in B");
}  // namespace fused

typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef float f32x16v __attribute__((ext_vector_type(16)));

// Chunk c of pixel p is 16-byte slot 5 p + c: a pixel pitch of five chunks
// (80 bytes; 5 is prime to 8 and 16), so the same chunk of any 8 or 16
// consecutive pixels falls in distinct 16-byte bank groups, wherever the
// run starts (a tap offset moves it), and a 16-byte LDS read by
// consecutive lanes is conflict-free.
__device__ inline int act_slot(int p, int c) { return 5 * p + c; }

// A wave's operands of one layer: its weight fragments (the A operand of
// each of the NS K-steps: lane l holds output channel l % 32 and the step's
// 8 K-values of half l / 32) and the biases of the output channels it holds.
template <int kCin>
struct LayerFrag {
  static constexpr int NS = kCin == fused::kCh ? 18 : 5;
  bf16x8 A[NS];
  float bco[16];
  __device__ inline void load(const bf16* __restrict__ wt, const float* __restrict__ bias) {
    const int lane = threadIdx.x & (kWave - 1);
    const int hf = lane >> 5, col = lane & 31;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int tap = kCin == fused::kCh ? t >> 1 : 2 * t + hf;
      const int ci = kCin == fused::kCh ? ((t & 1) << 4) + 8 * hf : 0;
      A[t] = tap < 9 ? *reinterpret_cast<const bf16x8*>(wt + ((size_t)col * 9 + tap) * kCin + ci)
                     : bf16x8{};
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) bco[v] = bias[8 * (v >> 2) + 4 * hf + (v & 3)];
  }
};

// One 32 -> 32 (or 8 -> 32 for the first layer) convolution over a square
// region: input side si (chunks per pixel kCin / 8), output side si - 2,
// ReLU(D + bias), BF16, zero at pixels outside the grid.  Output pixel q of
// the region lies at grid (gx0 + q % so, gy0 + q / so).  Each wave takes NT
// 32-pixel tiles at a time (tiles base, base + nwaves, ...), so NT
// independent MFMA chains hide the LDS reads of one another.
template <int kCin, int NT>
__device__ inline void fused_conv(const bf16x8* __restrict__ in, bf16x8* __restrict__ out,
                                  int si, const LayerFrag<kCin>& w, int gx0, int gy0, int hc,
                                  int wc, int nwaves) {
  constexpr int NS = LayerFrag<kCin>::NS;
  const int so = si - 2, npx = so * so;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int hf = lane >> 5, col = lane & 31;
  const int ntiles = (npx + 31) / 32;
  for (int base = wv; base < ntiles; base += NT * nwaves) {
    int r[NT], c[NT], q[NT];
    f32x16v acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      q[j] = min((base + j * nwaves) * 32 + col, npx - 1);
      r[j] = q[j] / so;
      c[j] = q[j] - r[j] * so;
      acc[j] = f32x16v{};
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int tap = kCin == fused::kCh ? t >> 1 : 2 * t + hf;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        bf16x8 b = bf16x8{};
        if (tap < 9) {
          const int pin = (r[j] + tap / 3) * si + (c[j] + tap % 3);
          b = kCin == fused::kCh ? in[act_slot(pin, ((t & 1) << 1) + hf)] : in[pin];
        }
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.A[t], b, acc[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int tile = base + j * nwaves;
      if (tile >= ntiles || tile * 32 + col >= npx) continue;
      const int gx = gx0 + c[j], gy = gy0 + r[j];
      const bool ing = gx >= 0 && gx < wc && gy >= 0 && gy < hc;
      bf16x4v* o = reinterpret_cast<bf16x4v*>(out);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        bf16x4v v;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float vv = acc[j][4 * g + k] + w.bco[4 * g + k];
          v[k] = (bf16)(ing && vv > 0.0f ? vv : 0.0f);
        }
        o[2 * act_slot(q[j], g) + hf] = v;
      }
    }
  }
}

typedef float f32x4v __attribute__((ext_vector_type(4)));

// The output layer (32 -> 1) on the matrix cores: for the side-so region of
// pixels whose 3 x 3 patches lie in `in` (side so + 2, act_slot layout),
// N - b6 = the sum over (tap, channel) of w6 a, as 9 K-steps (one per tap)
// of v_mfma_f32_16x16x32_bf16 with w6 in row 0 of the A operand (rows 1-15
// zero) and the patches of 16 pixels as the columns of B: lane l of a wave
// supplies chunk l / 16 of tap t's input pixel for pixel l % 16, and lanes
// 0-15 end with the 16 sums in accumulator 0.  w6 holds BF16 values, so
// its BF16 operand is exact; the sum is in the MFMA's order.  (On the
// vector units, 864 FP32 operations per pixel without contraction, this
// layer was the slowest launch of the chain.)  f(q, sum) takes pixel q's
// sum.
struct OutFrag {
  bf16x8 A[9];
  // Every lane loads (a load under the lane test was waited for inside it);
  // lanes other than 0, 16, 32 and 48 then zero theirs.
  __device__ inline void load(const float* __restrict__ w6) {
    const int lane = threadIdx.x & (kWave - 1);
    const bool row0 = (lane & 15) == 0;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const float4 lo = *reinterpret_cast<const float4*>(w6 + t * fused::kCh + 8 * (lane >> 4));
      const float4 hi =
          *reinterpret_cast<const float4*>(w6 + t * fused::kCh + 8 * (lane >> 4) + 4);
      const float f[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
      bf16x8 a;
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = row0 ? (bf16)f[k] : (bf16)0.0f;
      A[t] = a;
    }
  }
};

template <class F>
__device__ inline void fused_out(const bf16x8* __restrict__ in, int so, const OutFrag& w,
                                 int nwaves, F f) {
  const int npx = so * so, si = so + 2;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int px = lane & 15, ch = lane >> 4;
  const int ntiles = (npx + 15) / 16;
  for (int tile = wv; tile < ntiles; tile += nwaves) {
    const int q = min(tile * 16 + px, npx - 1);
    const int r = q / so, c = q - r * so;
    f32x4v acc = {};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int pin = (r + t / 3) * si + (c + t % 3);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.A[t], in[act_slot(pin, ch)], acc, 0, 0,
                                                     0);
    }
    if (ch == 0 && tile * 16 + px < npx) f(q, acc[0]);
  }
}

template <int kThreads, int NT>
__global__ __launch_bounds__(kThreads) void net_level_kernel(
    int hc, int wc, const float* __restrict__ rbar, const float* __restrict__ dis,
    const float* __restrict__ feat, const float* __restrict__ kc,
    const double* __restrict__ rms, int nrms, nn::PackedWeights wts,
    float* __restrict__ y, double* __restrict__ part, int ntiles,
    double* __restrict__ xsum) {
  using namespace fused;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  bf16x8* actA = reinterpret_cast<bf16x8*>(lds);
  bf16x8* actB = reinterpret_cast<bf16x8*>(lds + kBytesA);
  float* s_sig = reinterpret_cast<float*>(lds + kBytesA + kBytesB);  // 18 x 18
  float* s_ds = s_sig + kSN * kSN;
  float* s_n = s_ds + kSN * kSN;
  float* s_rb = s_n + kSN * kSN;  // the tile's rbar (16 x 16 of the 18 x 18)
  // The local-scale stage, in A (dead until layer 1 writes it).
  float* v_rb = reinterpret_cast<float*>(lds);
  float* v_ds = v_rb + kSV * kSV;
  float* v_tmp = v_ds + kSV * kSV;   // rows kSV x columns kS0
  float* v_sig = v_tmp + kSV * kS0;  // kS0 x kS0
  double* red = reinterpret_cast<double*>(actB);  // at the end, B is free
  constexpr int kWaves_ = kThreads / kWave;
  const int tid = threadIdx.x;
  const int X0 = blockIdx.x * kT, Y0 = blockIdx.y * kT;
  const long n = (long)hc * wc;
  auto in_grid = [&](int X, int Y) { return X >= 0 && X < wc && Y >= 0 && Y < hc; };

  // The RMS of D^-1/2 rbar over the grid from the restriction's partials,
  // in box_conv_kernel's order (thread t of 256 adds t, t + 256, ...).
  __shared__ double s_scale[256];
  __shared__ float s_scale0;
  if (tid < 256) {
    // Thread t adds partials t, t + 256, ... in turn, loaded 8 at a time.
    double v = 0.0;
    for (int i0 = tid; i0 < nrms; i0 += 256 * 8) {
      double t8[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t8[j] = i0 + 256 * j < nrms ? rms[i0 + 256 * j] : 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i0 + 256 * j < nrms) v += t8[j];
    }
    s_scale[tid] = v;
  }
  // Stage rbar and D^-1/2 on the tile and 13 cells around it.
  const int vx0 = X0 - kRN - kWin, vy0 = Y0 - kRN - kWin;
  for (int q = tid; q < kSV * kSV; q += kThreads) {
    const int X = vx0 + q % kSV, Y = vy0 + q / kSV;
    float rb = 0.0f, d = 0.0f;
    if (in_grid(X, Y)) {
      rb = rbar[(long)Y * wc + X];
      d = dis[(long)Y * wc + X];
    }
    v_rb[q] = rb;
    v_ds[q] = d;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) s_scale[tid] += s_scale[tid + st];
    __syncthreads();
  }
  if (tid == 0) s_scale0 = (float)sqrt(s_scale[0] / (double)n);
  // Row pass of the window sums (box_conv_kernel's): every staged row, the
  // input region's columns.
  for (int q = tid; q < kSV * kS0; q += kThreads) {
    const int rr = q / kS0, cc = q - rr * kS0 + kWin;
    const int X = vx0 + cc;
    double s = 0.0;
    for (int d = -kWin; d <= kWin; ++d) {
      if (X + d < 0 || X + d >= wc) continue;
      const double v = (double)v_rb[rr * kSV + cc + d] * (double)v_ds[rr * kSV + cc + d];
      s += v * v;
    }
    v_tmp[q] = (float)s;
  }
  __syncthreads();
  // Column pass, sigma and the packed input (box_conv_kernel's), zero outside
  // the grid.
  const float floor_s = kFloor * s_scale0;
  for (int q = tid; q < kS0 * kS0; q += kThreads) {
    const int r = q / kS0, c = q - r * kS0;
    const int X = X0 - kRN + c, Y = Y0 - kRN + r;
    bf16x8 o = bf16x8{};
    float sg = 0.0f;
    if (in_grid(X, Y)) {
      double s = 0.0;
      for (int d = -kWin; d <= kWin; ++d) {
        if (Y + d < 0 || Y + d >= hc) continue;
        s += (double)v_tmp[(r + kWin + d) * kS0 + c];
      }
      sg = fmaxf((float)sqrt(s / (double)((2 * kWin + 1) * (2 * kWin + 1))), floor_s);
      const int vi = (r + kWin) * kSV + c + kWin;
      o[0] = sg > 0.0f ? (bf16)((v_rb[vi] * v_ds[vi]) / sg) : (bf16)0.0f;
      o[1] = (bf16)feat[(long)Y * wc + X];
      o[2] = (bf16)feat[n + (long)Y * wc + X];
    }
    v_sig[q] = sg;
    actB[q] = o;
  }
  __syncthreads();
  // Keep what the end needs: sigma and D^-1/2 on the output region, rbar on
  // the tile.
  for (int q = tid; q < kSN * kSN; q += kThreads) {
    const int r = q / kSN, c = q - r * kSN;
    s_sig[q] = v_sig[(r + kRN - 1) * kS0 + c + kRN - 1];
    const int vi = (r + kRN - 1 + kWin) * kSV + c + kRN - 1 + kWin;
    s_ds[q] = v_ds[vi];
    s_rb[q] = v_rb[vi];
  }
  // Layers 1..5: input (B, side 30) -> A (28) -> B (26) -> A (24) -> B (22)
  // -> A (20).  Each layer's weights are loaded before the barrier that
  // waits for the layer before it, so the loads overlap the wait.
  {
    LayerFrag<8> w1;
    w1.load(wts.wf1, wts.bias);
    __syncthreads();
    fused_conv<8, NT>(actB, actA, kS0, w1, X0 - kRN + 1, Y0 - kRN + 1, hc, wc, kWaves_);
  }
#pragma unroll 1
  for (int l = 1; l < 5; ++l) {
    const int si = kS0 - 2 * l;
    const int off = kRN - l - 1;  // output region origin relative to the tile
    const bf16x8* src = (l & 1) ? actA : actB;
    bf16x8* dst = (l & 1) ? actB : actA;
    LayerFrag<kCh> w;
    w.load(wts.wfh + (size_t)(l - 1) * kCh * 9 * kCh, wts.bias + l * kCh);
    __syncthreads();
    fused_conv<kCh, NT>(src, dst, si, w, X0 - off, Y0 - off, hc, wc, kWaves_);
  }
  // Output layer on the 18 x 18 region from A (side 20) on the matrix cores
  // (fused_out, as out_tile_kernel); then y = N sigma D^-1/2 (zero outside
  // the grid).
  OutFrag wo;
  wo.load(wts.w6);
  __syncthreads();
  const float b6 = wts.bias[5 * kCh];
  fused_out(actA, kSN, wo, kWaves_, [&](int q, float sum) {
    const int r = q / kSN, c = q - r * kSN;
    const int X = X0 - 1 + c, Y = Y0 - 1 + r;
    s_n[q] = in_grid(X, Y) ? (sum + b6) * (s_sig[q] * s_ds[q]) : 0.0f;
  });
  __syncthreads();
  // The tile: y, T y of y (SMVP order, as sweeps_kernel), the partials and
  // the level-1 block sums.
  double* s_a = red;
  double* s_b = red + 256;
  double* s_t = red + 512;
  if (tid < 256) {
    const int tx = tid & (kT - 1), ty = tid / kT;
    const int X = X0 + tx, Y = Y0 + ty;
    const int k = (ty + 1) * kSN + tx + 1;
    const long i = (long)Y * wc + X;
    const bool in = X < wc && Y < hc;
    double kl = 0.0, kr = 0.0, kb = 0.0, kt = 0.0;
    if (in) {
      kl = kc[i];
      kb = kc[n + i];
      kr = X + 1 < wc ? (double)kc[i + 1] : 0.0;
      kt = Y + 1 < hc ? (double)kc[n + i + wc] : 0.0;
    }
    const double yc = (double)s_n[k];
    const double tyv = (((1.0 + (kr + kl)) + (kt + kb)) * yc -
                        (kr * (double)s_n[k + 1] + kl * (double)s_n[k - 1])) -
                       (kt * (double)s_n[k + kSN] + kb * (double)s_n[k - kSN]);
    double a = 0.0, b = 0.0;
    if (in) {
      y[i] = s_n[k];
      a = yc * (double)s_rb[k];
      b = yc * tyv;
    }
    s_a[tid] = a;
    s_b[tid] = b;
    s_t[tid] = tyv;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      s_a[tid] += s_a[tid + st];
      s_b[tid] += s_b[tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int tile = blockIdx.y * gridDim.x + blockIdx.x;
    part[tile] = s_a[0];
    part[ntiles + tile] = s_b[0];
  }
  if (xsum && tid < 16) {
    const int bx = tid & 3, by = tid >> 2;
    const int X1 = blockIdx.x * 4 + bx, Y1 = blockIdx.y * 4 + by;
    if (X1 < (wc >> 2) && Y1 < (hc >> 2)) {
      double rs[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double* p = &s_t[(4 * by + j) * kT + 4 * bx];
        rs[j] = (p[0] + p[1]) + (p[2] + p[3]);
      }
      xsum[(long)Y1 * (wc >> 2) + X1] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
    }
  }
}

// The start of the per-layer pipeline, for one 16 x 16 tile per 256-thread
// block: sigma and the network's packed input on the tile and 2 cells
// around it, the first layer on the tile and its ring, and the second layer
// on the tile, with the input and the first layer's activations in LDS.
// - The RMS floor from the restriction's per-block partials (thread t adds
//   partials t, t + 256, ... in turn, then the LDS tree).
// - rbar and D^-1/2 staged on the tile and 8 cells around it; the row pass
//   of the window sums (FP64, rounded to FP32) on the staged rows over the
//   input region's columns; the column pass (FP64 over the 13 row sums);
//   sigma = max(RMS of the window, kFloor times the grid's RMS); the input:
//   channel 0 = D^-1/2 rbar / sigma (0 where sigma is 0), channels 1, 2 =
//   the features, zero outside the grid.  Cells beyond the grid's edge add
//   nothing (adding +0 to a sum of squares, which is >= +0, is skipping
//   them).
// - Layer 1 (8 -> 32 channels) on the 18 x 18 region (fused_conv, zero
//   outside the grid) and layer 2 (32 -> 32) on the tile, each pixel in the
//   network's own layer kernels' MFMA sequence, so the second layer's
//   activations are bitwise theirs; stored at its padded positions.
// Each block redoes the RMS, the row sums of the rows around its tile and
// layer 1 on the tile's ring: cheaper than the separate launches (the
// window passes, layer 1 and layer 2 as four launches took 4.1 + 4.9 + 8.9
// + 12.1 us at 512^2, as one local-scale-and-layer-1 launch and layer 2
// 15.3 + 12.1 us).
constexpr int kBoxT = 16;
constexpr int kBoxA = kBoxT + 2;          // 18: layer-1 region
constexpr int kBoxC = kBoxT + 4;          // 20: sigma and input region
constexpr int kBoxS = kBoxC + 2 * kWin;   // 32: staged region
constexpr int kBoxStage = (2 * kBoxS * kBoxS + kBoxS * kBoxC) * 4;  // bytes
constexpr int kBoxA1 = kBoxA * kBoxA * 5 * 16;                       // bytes
__global__ __launch_bounds__(256) void box_conv_kernel(
    int hc, int wc, const float* __restrict__ rbar, const float* __restrict__ dis,
    const double* __restrict__ part, int npart, const float* __restrict__ feat,
    float* __restrict__ sig, nn::PackedWeights wts, bf16* __restrict__ a2) {
  constexpr int kN = kBoxS * kBoxS;
  constexpr int kIt = (kN + 255) / 256;
  constexpr int kNq = (kBoxC * kBoxC + 255) / 256;
  // The local-scale stage is dead once the input is formed: layer 1's
  // activations (act_slot layout) take its place.
  __shared__ __attribute__((aligned(16)))
  unsigned char s_u[kBoxStage > kBoxA1 ? kBoxStage : kBoxA1];
  float* s_rb = reinterpret_cast<float*>(s_u);
  float* s_ds = s_rb + kN;
  float* s_tmp = s_ds + kN;
  bf16x8* s_a1 = reinterpret_cast<bf16x8*>(s_u);
  __shared__ bf16x8 s_in[kBoxC * kBoxC];
  __shared__ double s_sc[256];
  __shared__ float s_floor;
  const int tid = threadIdx.x;
  const int X0 = blockIdx.x * kBoxT, Y0 = blockIdx.y * kBoxT;
  const long n = (long)hc * wc;
  auto in_grid = [&](int X, int Y) { return X >= 0 && X < wc && Y >= 0 && Y < hc; };
  // Every global load first: the first layer's weights, the RMS partials (8
  // per thread and round), the staged region, the features of the thread's
  // input pixels.
  LayerFrag<8> w1;
  w1.load(wts.wf1, wts.bias);
  double v = 0.0;
  for (int i0 = tid; i0 < npart; i0 += 256 * 8) {
    double t8[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t8[j] = part[i0 + 256 * j < npart ? i0 + 256 * j : 0];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i0 + 256 * j < npart) v += t8[j];
  }
  // Staged cell q: (X0 - 2 - kWin + q % kBoxS, Y0 - 2 - kWin + q / kBoxS).
  float rb[kIt], ds[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const int q = tid + 256 * k;
    const int X = X0 - 2 - kWin + q % kBoxS, Y = Y0 - 2 - kWin + q / kBoxS;
    const bool ok = q < kN && in_grid(X, Y);
    const long j = ok ? (long)Y * wc + X : 0;
    rb[k] = rbar[j];
    ds[k] = dis[j];
    if (!ok) rb[k] = ds[k] = 0.0f;
  }
  float f0[kNq], f1[kNq];
#pragma unroll
  for (int k = 0; k < kNq; ++k) {
    const int q = tid + 256 * k;
    const int X = X0 - 2 + q % kBoxC, Y = Y0 - 2 + q / kBoxC;
    const long j = q < kBoxC * kBoxC && in_grid(X, Y) ? (long)Y * wc + X : 0;
    f0[k] = feat[j];
    f1[k] = feat[n + j];
  }
  s_sc[tid] = v;
#pragma unroll
  for (int k = 0; k < kIt; ++k)
    if (tid + 256 * k < kN) {
      s_rb[tid + 256 * k] = rb[k];
      s_ds[tid + 256 * k] = ds[k];
    }
  __syncthreads();
  // Row pass: staged row r, region column c (staged column c + kWin).
  for (int q = tid; q < kBoxS * kBoxC; q += 256) {
    const int r = q / kBoxC, c = q - r * kBoxC + kWin;
    const int Xc = X0 - 2 - kWin + c;
    double s = 0.0;
#pragma unroll
    for (int d = -kWin; d <= kWin; ++d) {
      const double w = (double)s_rb[r * kBoxS + c + d] * (double)s_ds[r * kBoxS + c + d];
      s += Xc + d < 0 || Xc + d >= wc ? 0.0 : w * w;
    }
    s_tmp[q] = (float)s;
  }
  // The LDS tree of the RMS sum (element t adds element t + st): steps 128
  // and 64 through LDS, 32 .. 1 by shuffles in wave 0.
  if (tid < 128) s_sc[tid] += s_sc[tid + 128];
  __syncthreads();
  if (tid < 64) {
    double x = s_sc[tid] + s_sc[tid + 64];
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) x = x + __shfl_down(x, st);
    if (tid == 0) s_floor = kFloor * (float)sqrt(x / (double)n);
  }
  __syncthreads();
  // Column pass, sigma and the input on the tile and 2 cells around it.
  const float floor_s = s_floor;
#pragma unroll
  for (int k = 0; k < kNq; ++k) {
    const int q = tid + 256 * k;
    if (q >= kBoxC * kBoxC) break;
    const int r = q / kBoxC, c = q - r * kBoxC;
    const int X = X0 - 2 + c, Y = Y0 - 2 + r;
    bf16x8 o = bf16x8{};
    if (in_grid(X, Y)) {
      double s = 0.0;
#pragma unroll
      for (int d = -kWin; d <= kWin; ++d)
        s += Y + d < 0 || Y + d >= hc ? 0.0 : (double)s_tmp[(r + kWin + d) * kBoxC + c];
      const float sg =
          fmaxf((float)sqrt(s / (double)((2 * kWin + 1) * (2 * kWin + 1))), floor_s);
      const long i = (long)Y * wc + X;
      if (r >= 2 && r < 2 + kBoxT && c >= 2 && c < 2 + kBoxT) sig[i] = sg;
      const int vi = (r + kWin) * kBoxS + c + kWin;
      o[0] = sg > 0.0f ? (bf16)((s_rb[vi] * s_ds[vi]) / sg) : (bf16)0.0f;
      o[1] = (bf16)f0[k];
      o[2] = (bf16)f1[k];
    }
    s_in[q] = o;
  }
  // Layer 2's weights, loaded before the barrier that waits for the input.
  LayerFrag<fused::kCh> w2;
  w2.load(wts.wfh, wts.bias + fused::kCh);
  __syncthreads();
  // Layer 1 on the 18 x 18 region (the stage's LDS is dead now).
  fused_conv<8, 1>(s_in, s_a1, kBoxC, w1, X0 - 1, Y0 - 1, hc, wc, kWaves);
  __syncthreads();
  // Layer 2 on the tile: wave w takes 32-pixel tiles w and w + 4 (tile rows
  // 2j, 2j + 1).
  const int lane = tid & (kWave - 1), wv = tid / kWave;
  const int hf = lane >> 5, col = lane & 31;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int j = wv + 4 * k;
    const int q = 32 * j + col;
    const int r = q / kBoxT, c = q - r * kBoxT;
    f32x16v acc = {};
#pragma unroll
    for (int t = 0; t < LayerFrag<fused::kCh>::NS; ++t) {
      const int tap = t >> 1;
      const int pin = (r + tap / 3) * kBoxA + (c + tap % 3);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
          w2.A[t], s_a1[act_slot(pin, ((t & 1) << 1) + hf)], acc, 0, 0, 0);
    }
    const int X = X0 + c, Y = Y0 + r;
    bf16* px = a2 + ((size_t)(Y + 1) * (wc + 2) + (X + 1)) * fused::kCh;
    nn::store_out_row(px, acc, w2.bco, hf, X < wc && Y < hc);
  }
}

// The end of net_level_kernel for the per-layer pipeline (box_conv_kernel
// and the network's last three hidden layers): the output layer on the
// tile and its ring, from the last hidden layer's activations staged in LDS
// (the tile and 2 cells around it), by fused_out as in net_level_kernel, so
// N is bitwise its; then y = N sigma D^-1/2 on the tile and its ring, and
// the same y, partials and level-1 block sums as net_level_kernel.  One
// 256-thread block per 16 x 16 tile.  (A separate output-layer launch read
// each activation about three times from HBM and wrote N for this kernel
// to read back: 14 + 6.7 us at 512^2.)
__global__ __launch_bounds__(256) void out_tile_kernel(
    int hc, int wc, const bf16* __restrict__ a5, const float* __restrict__ w6,
    const float* __restrict__ bias, const float* __restrict__ sig,
    const float* __restrict__ dis, const float* __restrict__ rbar,
    const float* __restrict__ kc, float* __restrict__ y, double* __restrict__ part,
    int ntiles, double* __restrict__ xsum) {
  using namespace fused;
  constexpr int kSA = kSN + 2;                // 20: staged activation region
  constexpr int kChunks = kSA * kSA * 4;      // 16-byte chunks staged
  constexpr int kIt = (kChunks + 255) / 256;  // per thread
  constexpr int kNq = (kSN * kSN + 255) / 256;
  __shared__ bf16x8 s_act[kSA * kSA * 5];
  __shared__ float s_n[kSN * kSN];
  __shared__ float s_sd[kSN * kSN];
  __shared__ double s_a[256], s_b[256], s_t[256];
  const int tid = threadIdx.x;
  const int X0 = blockIdx.x * kT, Y0 = blockIdx.y * kT;
  const long n = (long)hc * wc;
  // Every global load of the kernel is issued before the first LDS store:
  // the staged activations, sigma D^-1/2 on the tile and its ring, and the
  // coarse coefficients and rbar of the thread's tile cell.
  // Region pixel q is coarse cell (X0 - 2 + q % kSA, Y0 - 2 + q / kSA), at
  // padded position (+1, +1) in a5 (zero ring; cells beyond it are only
  // read for ring pixels outside the grid, whose N is not used).
  OutFrag wo;
  wo.load(w6);
  const float b6 = bias[5 * kCh];
  const int Wp = wc + 2;
  bf16x8 v[kIt];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const int j = tid + 256 * k;
    const int q = j >> 2;
    const int Xp = X0 - 1 + q % kSA, Yp = Y0 - 1 + q / kSA;
    const bool ok = j < kChunks && Xp >= 0 && Xp <= wc + 1 && Yp >= 0 && Yp <= hc + 1;
    const size_t at = ok ? ((size_t)Yp * Wp + Xp) * kCh + 8 * (j & 3) : 0;
    v[k] = *reinterpret_cast<const bf16x8*>(a5 + at);
    if (!ok) v[k] = bf16x8{};
  }
  float sd[kNq];
#pragma unroll
  for (int k = 0; k < kNq; ++k) {
    const int q = tid + 256 * k;
    const int X = X0 - 1 + q % kSN, Y = Y0 - 1 + q / kSN;
    const bool ok = q < kSN * kSN && X >= 0 && X < wc && Y >= 0 && Y < hc;
    const long j = ok ? (long)Y * wc + X : 0;
    sd[k] = sig[j] * dis[j];  // >= 0; -1 marks a pixel outside the grid
    if (!ok) sd[k] = -1.0f;
  }
  const int tx = tid & (kT - 1), ty = tid / kT;
  const int X = X0 + tx, Y = Y0 + ty;
  const bool in = X < wc && Y < hc;
  const long i = in ? (long)Y * wc + X : 0;
  const float fkl = kc[i], fkb = kc[n + i];
  const float fkr = kc[X + 1 < wc ? i + 1 : i], fkt = kc[Y + 1 < hc ? n + i + wc : n + i];
  const float frb = rbar[i];
#pragma unroll
  for (int k = 0; k < kIt; ++k) {
    const int j = tid + 256 * k;
    if (j < kChunks) s_act[act_slot(j >> 2, j & 3)] = v[k];
  }
#pragma unroll
  for (int k = 0; k < kNq; ++k)
    if (tid + 256 * k < kSN * kSN) s_sd[tid + 256 * k] = sd[k];
  __syncthreads();
  fused_out(s_act, kSN, wo, 4, [&](int q, float sum) {
    s_n[q] = s_sd[q] < 0.0f ? 0.0f : (sum + b6) * s_sd[q];
  });
  __syncthreads();
  const int k = (ty + 1) * kSN + tx + 1;
  double kl = 0.0, kr = 0.0, kb = 0.0, kt = 0.0;
  if (in) {
    kl = fkl;
    kb = fkb;
    kr = X + 1 < wc ? (double)fkr : 0.0;
    kt = Y + 1 < hc ? (double)fkt : 0.0;
  }
  const double yc = (double)s_n[k];
  const double tyv = (((1.0 + (kr + kl)) + (kt + kb)) * yc -
                      (kr * (double)s_n[k + 1] + kl * (double)s_n[k - 1])) -
                     (kt * (double)s_n[k + kSN] + kb * (double)s_n[k - kSN]);
  double a = 0.0, b = 0.0;
  if (in) {
    y[i] = s_n[k];
    a = yc * (double)frb;
    b = yc * tyv;
  }
  s_a[tid] = a;
  s_b[tid] = b;
  s_t[tid] = tyv;
  __syncthreads();
  if (xsum && tid >= 64 && tid < 80) {
    const int bx = tid & 3, by = (tid >> 2) & 3;
    const int X1 = blockIdx.x * 4 + bx, Y1 = blockIdx.y * 4 + by;
    if (X1 < (wc >> 2) && Y1 < (hc >> 2)) {
      double rs[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double* p = &s_t[(4 * by + j) * kT + 4 * bx];
        rs[j] = (p[0] + p[1]) + (p[2] + p[3]);
      }
      xsum[(long)Y1 * (wc >> 2) + X1] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
    }
  }
  // The tile sums: the LDS tree's steps 128 and 64, then steps 32..1 by
  // shuffles in wave 0 (element t still adds element t + st at each step).
  if (tid < 128) {
    s_a[tid] += s_a[tid + 128];
    s_b[tid] += s_b[tid + 128];
  }
  __syncthreads();
  if (tid < 64) {
    double va = s_a[tid] + s_a[tid + 64], vb = s_b[tid] + s_b[tid + 64];
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
      va = va + __shfl_down(va, st);
      vb = vb + __shfl_down(vb, st);
    }
    if (tid == 0) {
      const int tile = blockIdx.y * gridDim.x + blockIdx.x;
      part[tile] = va;
      part[ntiles + tile] = vb;
    }
  }
}

// The factors g of the coarse directions, into g[0] and g[1].  One level:
// g = (y.rbar) / (y.T y), 0 unless y.T y > 0 and the ratio is finite.  Two
// levels: the joint minimiser of the energy norm of e - g0 P0 y0 - g1 P1 y1
// (the system of gamma2_kernel), whose cross entry c0^2 y0.T0 (Q y1) is the
// sum over level-1 cells of xsum * y1 (T0 is symmetric).  Each sum in the
// order of one kGammaThreads-thread block: thread t adds elements t,
// t + kGammaThreads, ... in turn, then the LDS tree.
constexpr int kGammaThreads = 1024;
__global__ __launch_bounds__(kGammaThreads) void joint_gamma_kernel(
    int levels, const double* __restrict__ p0, int n0, const double* __restrict__ p1,
    int n1, const double* __restrict__ xsum, const float* __restrict__ y1, int nx,
    double c0sq, double c1sq, float* __restrict__ g) {
  constexpr int kT = kGammaThreads;
  __shared__ double sh[5][kT];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // Each thread's elements are added in turn (the order above).  All five
  // sums' loads of a round are issued before any is added: a round per sum
  // (loads waited for sum by sum) took 21 us at 4096^2.  A round takes kB
  // elements of each tile sum and kBx of the cross sum, whose arrays are 16
  // times longer (level-1 cells against level-0 tiles): one round at
  // 4096^2 (four rounds of four took 10 us).
  // Every load is unconditional, from a clamped index (p0's first element
  // stands in for a level that is not there): a load under a test on the
  // level count was branched around and waited for inside the branch, one
  // memory latency per element (10 us at 4096^2).
  constexpr int kB = 4, kBx = 16;
  const int n2 = levels == 2 ? n1 : 0, nxx = levels == 2 ? nx : 0;
  const int nt = max(n0, n2);
  const double* q1 = n2 > 0 ? p1 : p0;
  const int q1b = n2 > 0 ? n1 : 0;
  const double* xs = nxx > 0 ? xsum : p0;
  const float* ys = nxx > 0 ? y1 : reinterpret_cast<const float*>(p0);
  for (int k0 = threadIdx.x, kx0 = threadIdx.x; k0 < nt || kx0 < nxx;
       k0 += kT * kB, kx0 += kT * kBx) {
    double t[4][kB], tx[kBx];
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      const int k = k0 + kT * j;
      const int a = k < n0 ? k : 0, b = k < n2 ? k : 0;
      t[0][j] = p0[a];
      t[1][j] = p0[n0 + a];
      t[2][j] = q1[b];
      t[3][j] = q1[q1b + b];
    }
#pragma unroll
    for (int j = 0; j < kBx; ++j) {
      const int k = kx0 + kT * j;
      const int a = k < nxx ? k : 0;
      tx[j] = xs[a] * (double)ys[a];
    }
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      const int k = k0 + kT * j;
      if (k < n0) {
        v[0] += t[0][j];
        v[1] += t[1][j];
      }
      if (k < n2) {
        v[2] += t[2][j];
        v[3] += t[3][j];
      }
    }
#pragma unroll
    for (int j = 0; j < kBx; ++j)
      if (kx0 + kT * j < nxx) v[4] += tx[j];
  }
  for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  // The LDS tree (element t adds element t + st, st = kT / 2 .. 1): the
  // steps down to 64 through LDS, the last six by shuffles in wave 0.
  for (int st = kT / 2; st >= kWave; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x >= kWave) return;
  double tot[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    double x = sh[q][threadIdx.x];
#pragma unroll
    for (int st = kWave / 2; st > 0; st >>= 1) x = x + __shfl_down(x, st);
    tot[q] = x;
  }
  if (threadIdx.x != 0) return;
  const double a0 = tot[0], b0 = tot[1];
  if (levels == 1) {
    const double q = a0 / b0;
    g[0] = (b0 > 0.0 && isfinite(q)) ? (float)q : 0.0f;
    g[1] = 0.0f;
    return;
  }
  const double a1 = c0sq * a0, b11 = c0sq * b0;
  const double a2 = c1sq * tot[2], b22 = c1sq * tot[3];
  const double b12 = c0sq * tot[4];
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

int ensure_fcg_buffers(tea_hip_chunk* c) {
  if (c->fcg_z) return TEA_HIP_OK;
  if (hipMalloc(&c->fcg_z, (size_t)c->geo.n * sizeof(float)) != hipSuccess) {
    c->fcg_z = nullptr;
    set_error("hipMalloc of the FCG z buffer failed");
    return TEA_HIP_ENOMEM;
  }
  TEA_HIP_TRY(hipMemsetAsync(c->fcg_z, 0, (size_t)c->geo.n * sizeof(float), c->stream));
  return TEA_HIP_OK;
}

}  // namespace
}  // namespace tea

using namespace tea;

// net_level_kernel's launch shapes: threads per block and 32-pixel tiles
// in flight per wave (TEA_FCG_NET_SHAPE=threads x tiles, e.g. 512x2).
struct NetShape {
  int threads, tiles;
};
static const NetShape kNetShapes[] = {{512, 1}, {512, 2}, {512, 4}, {1024, 1}, {1024, 2}};
static NetShape net_shape() {
  static const NetShape v = [] {
    NetShape s{512, 2};
    const char* e = getenv("TEA_FCG_NET_SHAPE");
    int t = 0, n = 0;
    if (e && sscanf(e, "%dx%d", &t, &n) == 2)
      for (const NetShape& k : kNetShapes)
        if (k.threads == t && k.tiles == n) s = k;
    return s;
  }();
  return v;
}
typedef void (*NetLevelFn)(int, int, const float*, const float*, const float*, const float*,
                           const double*, int, nn::PackedWeights, float*, double*, int,
                           double*);
static NetLevelFn net_level_fn(NetShape s) {
  if (s.threads == 1024) return s.tiles == 2 ? net_level_kernel<1024, 2> : net_level_kernel<1024, 1>;
  return s.tiles == 4 ? net_level_kernel<512, 4>
                      : s.tiles == 2 ? net_level_kernel<512, 2> : net_level_kernel<512, 1>;
}

// Whether the learned coarse solve runs as the per-layer pipeline (the
// default: box_conv_kernel, the network's last three hidden layers,
// out_tile_kernel) or as one fused launch per level (TEA_FCG_NET=fused),
// which gives the same y bit for bit.  The fused kernel recomputes each
// tile's halo at every layer (2.4 times the MFMA work of the per-layer
// pipeline with 16 x 16 tiles) and measured slower at 4096^2.
static bool net_per_layer() {
  static const bool v = [] {
    const char* e = getenv("TEA_FCG_NET");
    return e == nullptr || strcmp(e, "fused") != 0;
  }();
  return v;
}

// The learned coarse solve of level l on stream st: y, the tile partials
// and (level 0 of two) the level-1 block sums of T y.
static int launch_net_level(tea_hip_chunk* c, int l, hipStream_t st) {
  FcgLevel& L = c->fcg_lv[l];
  const dim3 tiles((unsigned)(L.wc + fused::kT - 1) / fused::kT,
                   (unsigned)(L.hc + fused::kT - 1) / fused::kT);
  double* xs = (l == 0 && c->fcg_levels == 2) ? c->fcg_xsum : nullptr;
  Timed t(c, K_FCG_NET, st);
  if (!net_per_layer()) {
    const NetShape shape = net_shape();
    hipLaunchKernelGGL(net_level_fn(shape), tiles, dim3(shape.threads), fused::kLdsBytes, st,
                       L.hc, L.wc, L.rbar, L.dis, L.feat, L.kc, L.rms, L.nrms,
                       nn::packed_weights(L.net), L.y, L.tpart, L.ntiles, xs);
  } else {
    const dim3 boxes((unsigned)(L.wc + kBoxT - 1) / kBoxT,
                     (unsigned)(L.hc + kBoxT - 1) / kBoxT);
    hipLaunchKernelGGL(box_conv_kernel, boxes, dim3(256), 0, st, L.hc, L.wc, L.rbar, L.dis,
                       L.rms, L.nrms, L.feat, L.sig, nn::packed_weights(L.net),
                       nn::hidden(L.net, 1));
    TEA_RETURN_IF(nn::forward_hidden_on_stream(L.net, st, 2));
    const nn::PackedWeights wts = nn::packed_weights(L.net);
    hipLaunchKernelGGL(out_tile_kernel, tiles, dim3(256), 0, st, L.hc, L.wc,
                       nn::last_hidden(L.net), wts.w6, wts.bias, L.sig, L.dis, L.rbar, L.kc,
                       L.y, L.tpart, L.ntiles, xs);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// The non-learned coarse solve of level l on stream st (same outputs).
static int launch_sweeps_level(tea_hip_chunk* c, int l, hipStream_t st) {
  FcgLevel& L = c->fcg_lv[l];
  Timed t(c, K_FCG_SWEEPS, st);
  const dim3 grid((unsigned)(L.wc + kSweepTile - 1) / kSweepTile,
                  (unsigned)(L.hc + kSweepTile - 1) / kSweepTile);
  hipLaunchKernelGGL(sweeps_kernel, grid, dim3(256), 0, st, L.hc, L.wc, c->fcg_sweeps,
                     L.kc, L.rbar, L.y, L.tpart, L.ntiles,
                     (l == 0 && c->fcg_levels == 2) ? c->fcg_xsum : nullptr);
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// The coarse part of the two-level preconditioner: the restriction of both
// levels in one pass (with the RMS partials the network's input scale
// needs), each level's coarse solve (level 1 on the side stream, beside
// level 0 on the chunk's), and the factors into c->fcg_g.
static int launch_coarse(tea_hip_chunk* c) {
  const int nl = c->fcg_levels;
  const bool net = c->fcg_sweeps == 0;
  FcgLevel& L0 = c->fcg_lv[0];
  FcgLevel& L1 = c->fcg_lv[1];
  {
    Timed t(c, K_FCG_RESTRICT);
    const dim3 grid((unsigned)(L0.wc + 31) / 32, (unsigned)(L0.hc + 3) / 4);
    const bool pairs = ((c->geo.x | c->geo.h) & 1) == 0;
    hipLaunchKernelGGL(pairs ? restrict8_kernel<true> : restrict8_kernel<false>, grid,
                       dim3(256), 0, c->stream, c->geo, L0.hc, L0.wc, c->mp_r, L0.rbar,
                       nl == 2 ? L1.rbar : nullptr, net ? L0.dis : nullptr,
                       net && nl == 2 ? L1.dis : nullptr, net ? L0.rms : nullptr,
                       net && nl == 2 ? L1.rms : nullptr);
    TEA_HIP_TRY(hipGetLastError());
  }
  auto level = [&](int l, hipStream_t st) {
    return net ? launch_net_level(c, l, st) : launch_sweeps_level(c, l, st);
  };
  // The network's two level chains overlap on two streams; the sweeps'
  // single kernels are shorter than the fork and join between the streams
  // (about 17 us at 4096^2), so they run one after the other.
  if (nl == 2 && !net) {
    TEA_RETURN_IF(level(0, c->stream));
    TEA_RETURN_IF(level(1, c->stream));
  } else if (nl == 2) {
    TEA_HIP_TRY(hipEventRecord(c->fcg_fork, c->stream));
    TEA_HIP_TRY(hipStreamWaitEvent(c->fcg_side, c->fcg_fork, 0));
    TEA_RETURN_IF(level(1, c->fcg_side));
    TEA_RETURN_IF(level(0, c->stream));
    TEA_HIP_TRY(hipEventRecord(c->fcg_join, c->fcg_side));
    TEA_HIP_TRY(hipStreamWaitEvent(c->stream, c->fcg_join, 0));
  } else {
    TEA_RETURN_IF(level(0, c->stream));
  }
  Timed t(c, net ? K_FCG_NET : K_FCG_SWEEPS);
  hipLaunchKernelGGL(joint_gamma_kernel, dim3(1), dim3(kGammaThreads), 0, c->stream, nl,
                     L0.tpart,
                     L0.ntiles, L1.tpart, L1.ntiles, c->fcg_xsum, L1.y,
                     nl == 2 ? L1.hc * L1.wc : 0, (double)L0.c * L0.c,
                     (double)L1.c * L1.c, c->fcg_g);
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

static int log2_of(int c) {
  int shift = 0;
  while ((1 << shift) < c) ++shift;
  return shift;
}

// z = M(r) and the partials of z.w into array `slot` of c->partials (not
// waited for); *len is the number of partials left to finish.
// Whether z = M(r) takes the four-cells-per-lane kernel.
static bool precondition_quads(const tea_hip_chunk* c, int use_net) {
  bool quads = use_quad_blocks(c, 2);
  const int nl = use_net ? c->fcg_levels : 0;
  for (int l = 0; l < nl; ++l) quads = quads && c->fcg_lv[l].c >= 4;
  return quads;
}

// With sub_q >= 0 (four-cells-per-lane kernel only) the z.w subtree sums go
// to array sub_q of c->sub and are left for the caller to finish.
static int launch_precondition(tea_hip_chunk* c, int use_net, int slot, int* len,
                               int sub_q = -1) {
  const Geo g = c->geo;
  const int sg = c->ngroup == 1 ? 1 : 0;
  const int nl = use_net ? c->fcg_levels : 0;
  double* partials = c->partials + (size_t)slot * c->nspan;
  *len = c->nspan;
  if (nl > 0) TEA_RETURN_IF(launch_coarse(c));
  const CoarseTerm none{nullptr, nullptr, 0, 0};
  auto term = [&](int l) {
    const FcgLevel& L = c->fcg_lv[l];
    return CoarseTerm{L.y, c->fcg_g + l, log2_of(L.c), L.wc};
  };
  const CoarseTerm t0 = nl > 0 ? term(0) : none;
  const CoarseTerm t1 = nl > 1 ? term(1) : none;
  const int kid = nl == 0 ? K_FCG_JACOBI : K_FCG_PROLONG;
  if (precondition_quads(c, use_net)) {
    double* sub = c->sub + (size_t)(sub_q > 0 ? sub_q : 0) * kWaves * c->nblk;
    {
      Timed t(c, kid);
      if (nl == 0)
        hipLaunchKernelGGL((fcg_z4_kernel<0>), quad_grid(c), dim3(kQuadThreads), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, sub,
                           c->nblk);
      else if (nl == 1)
        hipLaunchKernelGGL((fcg_z4_kernel<1>), quad_grid(c), dim3(kQuadThreads), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, sub,
                           c->nblk);
      else
        hipLaunchKernelGGL((fcg_z4_kernel<2>), quad_grid(c), dim3(kQuadThreads), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, sub,
                           c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    if (sub_q >= 0) return TEA_HIP_OK;
    return launch_span_finish(c, 1, len, slot);
  }
  if (use_pair_blocks(c, 2)) {
    {
      Timed t(c, kid);
      if (nl == 0)
        hipLaunchKernelGGL((fcg_z1_kernel<0>), pair_grid(c), dim3(kBlock), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, c->sub,
                           c->nblk);
      else if (nl == 1)
        hipLaunchKernelGGL((fcg_z1_kernel<1>), pair_grid(c), dim3(kBlock), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, c->sub,
                           c->nblk);
      else
        hipLaunchKernelGGL((fcg_z1_kernel<2>), pair_grid(c), dim3(kBlock), 0, c->stream, g,
                           c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, c->sub,
                           c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    return launch_span_finish(c, 1, len, slot);
  }
  Timed t(c, kid);
  if (c->pairs) {
    if (nl == 0)
      hipLaunchKernelGGL((fcg_z2_kernel<0>), span_grid(c), dim3(kBlock), 0, c->stream, g,
                         c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, partials,
                         c->nspan, sg);
    else if (nl == 1)
      hipLaunchKernelGGL((fcg_z2_kernel<1>), span_grid(c), dim3(kBlock), 0, c->stream, g,
                         c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, partials,
                         c->nspan, sg);
    else
      hipLaunchKernelGGL((fcg_z2_kernel<2>), span_grid(c), dim3(kBlock), 0, c->stream, g,
                         c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, partials,
                         c->nspan, sg);
  } else if (nl == 0) {
    JacobiBody body{c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, c->fcg_z, g.x};
    hipLaunchKernelGGL((span_kernel<1, JacobiBody>), span_grid(c), dim3(kBlock), 0,
                       c->stream, g, body, partials, c->nspan, sg);
  } else if (nl == 1) {
    ProlongBody<1> body{c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, g.x, g.h};
    hipLaunchKernelGGL((span_kernel<1, ProlongBody<1>>), span_grid(c), dim3(kBlock), 0,
                       c->stream, g, body, partials, c->nspan, sg);
  } else {
    ProlongBody<2> body{c->mp_r, c->mp_kx, c->mp_ky, c->mp_w, t0, t1, c->fcg_z, g.x, g.h};
    hipLaunchKernelGGL((span_kernel<1, ProlongBody<2>>), span_grid(c), dim3(kBlock), 0,
                       c->stream, g, body, partials, c->nspan, sg);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

static void free_level(FcgLevel& L) {
  for (float* b : {L.feat, L.rbar, L.kc, L.dis, L.sig, L.y})
    if (b) (void)hipFree(b);
  if (L.rms) (void)hipFree(L.rms);
  if (L.tpart) (void)hipFree(L.tpart);
  if (L.net) tea_nn_destroy(L.net);
  L = FcgLevel();
}

// net_level_kernel's dynamic LDS exceeds the default limit of a launch:
// raise it once per process.
static int ensure_net_lds(int device) {
  static bool done[64] = {};
  if (device >= 0 && device < 64 && done[device]) return TEA_HIP_OK;
  for (const NetShape& k : kNetShapes)
    TEA_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(net_level_fn(k)),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    fused::kLdsBytes));
  if (device >= 0 && device < 64) done[device] = true;
  return TEA_HIP_OK;
}

// Frees every coarse level and the buffers shared between them.
static void release_levels(tea_hip_chunk* c) {
  for (FcgLevel& L : c->fcg_lv) free_level(L);
  if (c->fcg_xsum) (void)hipFree(c->fcg_xsum);
  c->fcg_xsum = nullptr;
  if (c->fcg_g) (void)hipFree(c->fcg_g);
  c->fcg_g = nullptr;
  c->fcg_levels = 0;
  c->fcg_sweeps = 0;
}

// The side stream and the fork / join events of a two-level coarse part.
static int ensure_side_stream(tea_hip_chunk* c) {
  if (!c->fcg_side) TEA_HIP_TRY(hipStreamCreateWithFlags(&c->fcg_side, hipStreamNonBlocking));
  if (!c->fcg_fork) TEA_HIP_TRY(hipEventCreateWithFlags(&c->fcg_fork, hipEventDisableTiming));
  if (!c->fcg_join) TEA_HIP_TRY(hipEventCreateWithFlags(&c->fcg_join, hipEventDisableTiming));
  return TEA_HIP_OK;
}

extern "C" {

int tea_hip_fcg_attach_sweeps(tea_hip_chunk* c, int sweeps, int coarsen, int levels) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(sweeps >= 1 && sweeps <= kMaxSweeps, "sweeps must be 1 to 8");
  TEA_CHECK_ARG(coarsen == 8, "the coarse sweeps need a coarsening factor of 8");
  TEA_CHECK_ARG(levels >= 1 && levels <= kFcgMaxLevels, "levels must be 1 or 2");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(ensure_fcg_buffers(c));
  release_levels(c);
  const int nx = c->geo.x - 2 * c->geo.h, ny = c->geo.y - 2 * c->geo.h;
  TEA_CHECK_ARG(nx % coarsen == 0 && ny % coarsen == 0,
                "the interior must be a multiple of the coarsening factor");
  auto fail = [] {
    set_error("hipMalloc of the preconditioner's coarse buffers failed");
    return TEA_HIP_ENOMEM;
  };
  // Level l coarsens by 8 * 4^l while the interior divides.
  for (int l = 0; l < levels; ++l) {
    const int cl = coarsen << (2 * l);
    if (nx % cl != 0 || ny % cl != 0) break;
    FcgLevel& L = c->fcg_lv[l];
    L.c = cl;
    L.hc = ny / cl;
    L.wc = nx / cl;
    L.ntiles = ((L.wc + kSweepTile - 1) / kSweepTile) * ((L.hc + kSweepTile - 1) / kSweepTile);
    const size_t cn = (size_t)L.hc * L.wc;
    if (hipMalloc(&L.kc, 2 * cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.rbar, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.y, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.tpart, 2 * (size_t)L.ntiles * sizeof(double)) != hipSuccess)
      return fail();
    c->fcg_levels = l + 1;
  }
  if (hipMalloc(&c->fcg_g, 2 * sizeof(float)) != hipSuccess) return fail();
  TEA_HIP_TRY(hipMemsetAsync(c->fcg_g, 0, 2 * sizeof(float), c->stream));
  if (c->fcg_levels == 2) {
    const size_t n1 = (size_t)c->fcg_lv[1].hc * c->fcg_lv[1].wc;
    if (hipMalloc(&c->fcg_xsum, n1 * sizeof(double)) != hipSuccess) return fail();
    TEA_RETURN_IF(ensure_side_stream(c));
  }
  c->fcg_sweeps = sweeps;
  return TEA_HIP_OK;
}

int tea_hip_fcg_attach(tea_hip_chunk* c, const float* params, size_t count,
                       const float* feature_norm, int coarsen, int levels) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(ensure_fcg_buffers(c));
  release_levels(c);
  if (!params) return TEA_HIP_OK;  // Jacobi only
  TEA_CHECK_ARG(feature_norm != nullptr, "null feature normalisation");
  TEA_CHECK_ARG(coarsen >= 1 && (coarsen & (coarsen - 1)) == 0,
                "coarsening factor must be a power of two");
  TEA_CHECK_ARG(levels >= 1 && levels <= kFcgMaxLevels, "levels must be 1 or 2");
  for (int q = 0; q < 4; ++q) c->fcg_norm[q] = feature_norm[q];
  TEA_CHECK_ARG(c->fcg_norm[1] > 0.0f && c->fcg_norm[3] > 0.0f,
                "feature scales must be positive");
  const int nx = c->geo.x - 2 * c->geo.h, ny = c->geo.y - 2 * c->geo.h;
  TEA_CHECK_ARG(nx % coarsen == 0 && ny % coarsen == 0,
                "the interior must be a multiple of the coarsening factor");
  TEA_CHECK_ARG((nx / coarsen) % 32 == 0, "coarse width must be a multiple of 32");
  // Level l coarsens by coarsen * 4^l; a level is added while the grid
  // divides and its width stays a multiple of 32.
  auto fail = [] {
    set_error("hipMalloc of the preconditioner's coarse buffers failed");
    return TEA_HIP_ENOMEM;
  };
  TEA_CHECK_ARG(coarsen == 8, "the network's restriction needs a coarsening factor of 8");
  for (int l = 0; l < levels; ++l) {
    const int cl = coarsen << (2 * l);
    if (nx % cl != 0 || ny % cl != 0 || (nx / cl) % 32 != 0) break;
    FcgLevel& L = c->fcg_lv[l];
    L.c = cl;
    L.hc = ny / cl;
    L.wc = nx / cl;
    const size_t cn = (size_t)L.hc * L.wc;
    TEA_RETURN_IF(tea_nn_create(c->device, 1, L.hc, L.wc, &L.net));
    TEA_RETURN_IF(tea_nn_set_params(L.net, params, count));
    // One RMS partial per block of the restriction (restrict8_kernel's grid
    // over level 0), for either level.
    const FcgLevel& L0 = c->fcg_lv[0];
    L.nrms = ((L0.wc + 31) / 32) * ((L0.hc + 3) / 4);
    L.ntiles = ((L.wc + fused::kT - 1) / fused::kT) * ((L.hc + fused::kT - 1) / fused::kT);
    if (hipMalloc(&L.feat, 2 * cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.kc, 2 * cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.dis, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.sig, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.rbar, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.y, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.rms, (size_t)L.nrms * sizeof(double)) != hipSuccess ||
        hipMalloc(&L.tpart, 2 * (size_t)L.ntiles * sizeof(double)) != hipSuccess)
      return fail();
    c->fcg_levels = l + 1;
  }
  if (hipMalloc(&c->fcg_g, 2 * sizeof(float)) != hipSuccess) return fail();
  TEA_HIP_TRY(hipMemsetAsync(c->fcg_g, 0, 2 * sizeof(float), c->stream));
  if (c->fcg_levels == 2) {
    const size_t n1 = (size_t)c->fcg_lv[1].hc * c->fcg_lv[1].wc;
    if (hipMalloc(&c->fcg_xsum, n1 * sizeof(double)) != hipSuccess) return fail();
    TEA_RETURN_IF(ensure_side_stream(c));
  }
  TEA_RETURN_IF(ensure_net_lds(c->device));
  return TEA_HIP_OK;
}

int tea_hip_fcg_levels(tea_hip_chunk* c, int* levels) {
  TEA_CHECK_ARG(c != nullptr && levels != nullptr, "null argument");
  *levels = c->fcg_levels;
  return TEA_HIP_OK;
}

// Coarse operators and features of every level from the FP32 coefficients
// (after tea_hip_mpcg_init).
static int fcg_features(tea_hip_chunk* c) {
  for (int l = 0; l < c->fcg_levels; ++l) {
    FcgLevel& L = c->fcg_lv[l];
    const dim3 grid(((unsigned)(L.hc * L.wc) + 255) / 256);
    hipLaunchKernelGGL(coarse_coef_kernel, grid, dim3(256), 0, c->stream, c->geo, L.c,
                       L.hc, L.wc, c->mp_kx, c->mp_ky, L.kc);
    if (c->fcg_sweeps > 0) continue;  // the sweeps need no network features
    hipLaunchKernelGGL(feature_kernel, grid, dim3(256), 0, c->stream, L.hc, L.wc, L.kc,
                       c->fcg_norm[0], c->fcg_norm[1], c->fcg_norm[2], c->fcg_norm[3],
                       L.feat, L.dis);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

int tea_hip_fcg_init(tea_hip_chunk* c, int coefficient, double rx, double ry,
                     double* rr, double* bb) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->fcg_z != nullptr, "tea_hip_fcg_attach has not been called");
  TEA_RETURN_IF(tea_hip_mpcg_init(c, coefficient, rx, ry, rr, bb));
  return fcg_features(c);
}

int tea_hip_fcg_precondition(tea_hip_chunk* c, int use_net, double* zw) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->fcg_z != nullptr && c->mp_r != nullptr, "FCG not initialised");
  TEA_CHECK_ARG(zw != nullptr, "null zw");
  TEA_CHECK_ARG(!use_net || c->fcg_levels > 0, "no coarse correction attached");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD_KEEP_U(c);
  int len;
  TEA_RETURN_IF(launch_precondition(c, use_net, 0, &len));
  return finish_reduction_n(c, 1, len, zw);
}

int tea_hip_fcg_calc_r_precondition(tea_hip_chunk* c, double alpha, int use_net,
                                    double* rr, double* zw) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->fcg_z != nullptr && c->mp_r != nullptr, "FCG not initialised");
  TEA_CHECK_ARG(rr != nullptr && zw != nullptr, "null outputs");
  TEA_CHECK_ARG(!use_net || c->fcg_levels > 0, "no coarse correction attached");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD_KEEP_U(c);
  double out[2];
  if (use_quad_blocks(c, 2) && precondition_quads(c, use_net)) {
    // Both kernels leave their subtree sums in c->sub (r.r in array 0, z.w
    // in array 1; nothing between them uses it) for one finisher launch.
    int len;
    TEA_RETURN_IF(launch_mp_r(c, alpha, &len, false, false));
    TEA_RETURN_IF(launch_precondition(c, use_net, 0, &len, 1));
    TEA_RETURN_IF(launch_span_finish(c, 2, &len, 0, true));
    TEA_RETURN_IF(finish_reduction_n(c, 2, len, out));
  } else {
    int lens[2] = {c->nspan, c->nspan};
    TEA_RETURN_IF(launch_mp_r(c, alpha, &lens[0], false));
    TEA_RETURN_IF(launch_precondition(c, use_net, 1, &lens[1]));
    TEA_RETURN_IF(finish_reduction_lens(c, 2, lens, out));
  }
  *rr = out[0];
  *zw = out[1];
  return TEA_HIP_OK;
}

int tea_hip_fcg_calc_w(tea_hip_chunk* c, double alpha, int apply_u, double beta,
                       double* pw, double* pr) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->fcg_z != nullptr && c->mp_r != nullptr, "FCG not initialised");
  TEA_CHECK_ARG(pw && pr, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD_KEEP_U(c);
  // p = z + beta p_old (the mixed-precision kernel with z in place of r),
  // w = A p, and the dots p.w and p.r.
  MpArgs a{c->mp_p[c->mp_cur], c->mp_p[mp_next(c->mp_cur)], c->fcg_z,
           c->field[TEA_FIELD_U], c->mp_w, c->mp_kx, c->mp_ky, c->mp_r,
           alpha, beta, apply_u ? 1 : 0, 1};
  int len = c->nspan;
  if (use_pair_blocks(c, 1)) {
    if (c->quads) {
      TEA_RETURN_IF(launch_mp_w4(c, a, K_FCG_W));
    } else {
      if (c->hold) TEA_RETURN_IF(apply_mp_hold(c));
      Timed t(c, K_FCG_W);
      hipLaunchKernelGGL((mp_w1_kernel<true>), pair_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, a, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    TEA_RETURN_IF(launch_span_finish(c, 2, &len, 0, true));
  } else {
    if (c->hold) TEA_RETURN_IF(apply_mp_hold(c));
    Timed t(c, K_FCG_W);
    if (c->pairs)
      hipLaunchKernelGGL((mp_w2_kernel<true, 4>), span_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, a, c->partials, c->nspan,
                         c->ngroup == 1 ? 1 : 0);
    else
      hipLaunchKernelGGL((mp_w_kernel<4, true>), span_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, a, c->partials, c->nspan,
                         c->ngroup == 1 ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  c->mp_cur = mp_next(c->mp_cur);
  double out[2];
  TEA_RETURN_IF(finish_reduction_n(c, 2, len, out));
  *pw = out[0];
  *pr = out[1];
  return TEA_HIP_OK;
}

int tea_hip_fcg_read_z(tea_hip_chunk* c, float* host, size_t count) {
  TEA_CHECK_ARG(c != nullptr && c->fcg_z != nullptr, "FCG not initialised");
  TEA_CHECK_ARG(count <= (size_t)c->geo.n, "count exceeds the field length");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_HIP_TRY(hipMemcpyAsync(host, c->fcg_z, count * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
  return sync_chunk(c);
}

}  // extern "C"
