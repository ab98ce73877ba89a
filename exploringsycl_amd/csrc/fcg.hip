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

// Coarse block means of r, and per-block partial sums of (D^-1/2 rbar)^2.
__global__ __launch_bounds__(256) void restrict_kernel(Geo g, int c, int hc, int wc,
                                                       const float* __restrict__ r,
                                                       const float* __restrict__ dis,
                                                       float* __restrict__ rbar,
                                                       double* __restrict__ part) {
  __shared__ double sh[256];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  double v2 = 0.0;
  if (idx < hc * wc) {
  const int X = idx % wc, Y = idx / wc;
  double sum = 0.0;
  if (((g.x | g.h | c) & 1) == 0) {
    // Every block row starts on an even index: 8-byte loads.
    for (int b = 0; b < c; ++b) {
      const int row = (g.h + Y * c + b) * g.x + g.h + X * c;
      for (int a = 0; a < c; a += 2) {
        const float2 v = ld2(r, row + a);
        sum += (double)v.x + (double)v.y;
      }
    }
  } else {
    for (int b = 0; b < c; ++b) {
      const int row = (g.h + Y * c + b) * g.x + g.h + X * c;
      for (int a = 0; a < c; ++a) sum += (double)r[row + a];
    }
  }
  const float m = (float)(sum / (double)(c * c));
  rbar[idx] = m;
  const double ms = (double)m * (double)dis[idx];
  v2 = ms * ms;
  }
  sh[threadIdx.x] = v2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Local scale of the network input: sigma = RMS of D^-1/2 rbar over the
// (2 kWin + 1)^2 window around each coarse cell (zero beyond the edge), in
// two passes (rows into tmp, then columns), each summed in a fixed order,
// and at least kFloor times the RMS over the grid (so that regions of
// round-off do not reach the network magnified).
constexpr int kWin = 6;         // the network's receptive-field radius
constexpr float kFloor = 1e-3f;

// Row pass of the window sums; the grid's extra last block computes the
// RMS scale[0] (of D^-1/2 rbar over the coarse grid, from the restriction's
// per-block partials) at the same time, since the row pass does not need it.
__global__ __launch_bounds__(256) void box_rows_kernel(int hc, int wc,
                                                       const float* __restrict__ rbar,
                                                       const float* __restrict__ dis,
                                                       float* __restrict__ tmp,
                                                       const double* __restrict__ part,
                                                       int npart,
                                                       float* __restrict__ scale) {
  if (blockIdx.x == gridDim.x - 1) {
    __shared__ double sh[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < npart; i += 256) v += part[i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) scale[0] = (float)sqrt(sh[0] / (double)((long)hc * wc));
    return;
  }
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= hc * wc) return;
  const int X = idx % wc;
  double s = 0.0;
  for (int d = -kWin; d <= kWin; ++d) {
    if (X + d < 0 || X + d >= wc) continue;
    const double v = (double)rbar[idx + d] * (double)dis[idx + d];
    s += v * v;
  }
  tmp[idx] = (float)s;
}

// Column pass: sigma (floored), then the cell's network input: channel 0 =
// D^-1/2 rbar / sigma (0 where sigma is 0), channels 1, 2 = features.
__global__ __launch_bounds__(256) void box_cols_kernel(int hc, int wc,
                                                       const float* __restrict__ tmp,
                                                       const float* __restrict__ scale,
                                                       float* __restrict__ sig,
                                                       const float* __restrict__ rbar,
                                                       const float* __restrict__ dis,
                                                       const float* __restrict__ feat,
                                                       bf16* __restrict__ in0) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= hc * wc) return;
  const int X = idx % wc, Y = idx / wc;
  double s = 0.0;
  for (int d = -kWin; d <= kWin; ++d) {
    if (Y + d < 0 || Y + d >= hc) continue;
    s += (double)tmp[idx + d * wc];
  }
  const float sg = fmaxf((float)sqrt(s / (double)((2 * kWin + 1) * (2 * kWin + 1))),
                         kFloor * scale[0]);
  sig[idx] = sg;
  bf16x8 o = bf16x8{};
  o[0] = sg > 0.0f ? (bf16)((rbar[idx] * dis[idx]) / sg) : (bf16)0.0f;
  o[1] = (bf16)feat[idx];
  o[2] = (bf16)feat[hc * wc + idx];
  *reinterpret_cast<bf16x8*>(in0 + ((size_t)(Y + 1) * (wc + 2) + (X + 1)) * 8) = o;
}

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

// Block means of a coarser level from a finer one (factor f) and the
// per-block partial sums of (D^-1/2 rbar)^2 of the coarser level.
__global__ __launch_bounds__(256) void restrict_coarse_kernel(
    int wf, int f, int hc, int wc, const float* __restrict__ fine,
    const float* __restrict__ dis, float* __restrict__ rbar, double* __restrict__ part) {
  __shared__ double sh[256];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  double v2 = 0.0;
  if (idx < hc * wc) {
    const int X = idx % wc, Y = idx / wc;
    double sum = 0.0;
    for (int b = 0; b < f; ++b)
      for (int a = 0; a < f; ++a) sum += (double)fine[(Y * f + b) * wf + X * f + a];
    const float m = (float)(sum / (double)(f * f));
    rbar[idx] = m;
    const double ms = (double)m * (double)dis[idx];
    v2 = ms * ms;
  }
  sh[threadIdx.x] = v2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
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

// The network's output back in T's variables, y = D^-1/2 sigma N (into y,
// not in place: neighbours read N), and the per-block partial sums over
// coarse cells of y.rbar and y.(T y), with T the coarse operator (identity
// plus the 5-point operator of the face coefficients kc, zero beyond the
// domain edge).  Neighbours' y are formed by the same expression.
__global__ __launch_bounds__(256) void coarse_dots_kernel(int hc, int wc,
                                                          const float* __restrict__ net,
                                                          const float* __restrict__ dis,
                                                          const float* __restrict__ sig,
                                                          float* __restrict__ y,
                                                          const float* __restrict__ rbar,
                                                          const float* __restrict__ kc,
                                                          double* __restrict__ part) {
  __shared__ double s0[256], s1[256];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (idx < hc * wc) {
    auto ys = [&](int j) { return net[j] * (sig[j] * dis[j]); };
    const int X = idx % wc, Y = idx / wc;
    const float* kxc = kc;
    const float* kyc = kc + hc * wc;
    const float yf = ys(idx);
    y[idx] = yf;
    const double yc = yf;
    const double kl = kxc[idx], kb = kyc[idx];
    const double kr = X + 1 < wc ? (double)kxc[idx + 1] : 0.0;
    const double kt = Y + 1 < hc ? (double)kyc[idx + wc] : 0.0;
    double ty = (1.0 + kl + kr + kb + kt) * yc;
    if (X > 0) ty -= kl * (double)ys(idx - 1);
    if (X + 1 < wc) ty -= kr * (double)ys(idx + 1);
    if (Y > 0) ty -= kb * (double)ys(idx - wc);
    if (Y + 1 < hc) ty -= kt * (double)ys(idx + wc);
    a = yc * (double)rbar[idx];
    b = yc * ty;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s0[threadIdx.x] += s0[threadIdx.x + st];
      s1[threadIdx.x] += s1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0[0];
    part[gridDim.x + blockIdx.x] = s1[0];
  }
}

// The A-optimal factor of the coarse correction P y: (y.rbar) / (y.T y)
// (the minimiser of the energy norm of e - g P y, since (P y)^T r =
// c^2 y.rbar and (P y)^T A (P y) = c^2 y.T y); 0 if y.T y is not positive.
// Written to scale[1].
__global__ __launch_bounds__(256) void gamma_kernel(const double* __restrict__ part,
                                                    int nblk, float* __restrict__ scale) {
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    a += part[i];
    b += part[nblk + i];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s0[threadIdx.x] += s0[threadIdx.x + st];
      s1[threadIdx.x] += s1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double g = s0[0] / s1[0];
    scale[1] = (s1[0] > 0.0 && isfinite(g)) ? (float)g : 0.0f;
  }
}

// Partial sums over level-1 cells of y1.(T1 Q y2): Q y2 is the level-2
// correction on the level-1 grid (constant over blocks of f x f), T1 the
// level-1 coarse operator.  With P2 = P1 Q, (P1 y1)^T A (P2 y2) =
// c1^2 y1.(T1 Q y2): the cross term of the two coarse directions.
__global__ __launch_bounds__(256) void cross_dots_kernel(int hc, int wc, int f, int wc2,
                                                         const float* __restrict__ y1,
                                                         const float* __restrict__ y2,
                                                         const float* __restrict__ kc,
                                                         double* __restrict__ part) {
  __shared__ double sh[256];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (idx < hc * wc) {
    const int X = idx % wc, Y = idx / wc;
    const float* kxc = kc;
    const float* kyc = kc + hc * wc;
    auto q = [&](int xx, int yy) { return (double)y2[(yy / f) * wc2 + xx / f]; };
    const double qc = q(X, Y);
    const double kl = kxc[idx], kb = kyc[idx];
    const double kr = X + 1 < wc ? (double)kxc[idx + 1] : 0.0;
    const double kt = Y + 1 < hc ? (double)kyc[idx + wc] : 0.0;
    double t = (1.0 + kl + kr + kb + kt) * qc;
    if (X > 0) t -= kl * q(X - 1, Y);
    if (X + 1 < wc) t -= kr * q(X + 1, Y);
    if (Y > 0) t -= kb * q(X, Y - 1);
    if (Y + 1 < hc) t -= kt * q(X, Y + 1);
    v = (double)y1[idx] * t;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Joint factors of the two coarse directions: minimise the energy norm of
// e - g1 P1 y1 - g2 P2 y2, i.e. solve [B11 B12; B12 B22] g = a with
// a_l = c_l^2 y_l.rbar_l, B_ll = c_l^2 y_l.T_l y_l, B12 = c1^2 y1.T1 Q y2
// (the Gram matrix of the two directions in the A inner product, SPD when
// they are independent).  Falls back to the separate factors a_l / B_ll
// when it is not safely positive definite.  Writes scale1[1], scale2[1].
__global__ __launch_bounds__(256) void gamma2_kernel(const double* __restrict__ p1, int n1,
                                                     const double* __restrict__ p2, int n2,
                                                     const double* __restrict__ px, int nx,
                                                     double c1sq, double c2sq,
                                                     float* __restrict__ scale1,
                                                     float* __restrict__ scale2) {
  __shared__ double sh[5][256];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < n1; i += 256) {
    v[0] += p1[i];
    v[1] += p1[n1 + i];
  }
  for (int i = threadIdx.x; i < n2; i += 256) {
    v[2] += p2[i];
    v[3] += p2[n2 + i];
  }
  for (int i = threadIdx.x; i < nx; i += 256) v[4] += px[i];
  for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a1 = c1sq * sh[0][0], b11 = c1sq * sh[1][0];
    const double a2 = c2sq * sh[2][0], b22 = c2sq * sh[3][0];
    const double b12 = c1sq * sh[4][0];
    const double det = b11 * b22 - b12 * b12;
    double g1, g2;
    if (b11 > 0.0 && b22 > 0.0 && det > 1e-12 * b11 * b22) {
      g1 = (a1 * b22 - a2 * b12) / det;
      g2 = (a2 * b11 - a1 * b12) / det;
    } else {
      g1 = b11 > 0.0 ? a1 / b11 : 0.0;
      g2 = 0.0;
    }
    scale1[1] = isfinite(g1) ? (float)g1 : 0.0f;
    scale2[1] = isfinite(g2) ? (float)g2 : 0.0f;
  }
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
  double t = 0.0;
  if (ok) {
    const int base = (g.h + 8 * Y) * g.x + g.h + 8 * X + 4 * half;
    double q[8];
    if (((g.x | g.h) & 1) == 0) {
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const float2 a = ld2(r, base + b * g.x), c = ld2(r, base + b * g.x + 2);
        q[b] = ((double)a.x + (double)a.y) + ((double)c.x + (double)c.y);
      }
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
    const double v = (double)m * (double)dis0[Y * wc + X];
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

// The factors g of the coarse directions, into g[0] and g[1].  One level:
// g = (y.rbar) / (y.T y), 0 unless y.T y > 0 and the ratio is finite.  Two
// levels: the joint minimiser of the energy norm of e - g0 P0 y0 - g1 P1 y1
// (the system of gamma2_kernel), whose cross entry c0^2 y0.T0 (Q y1) is the
// sum over level-1 cells of xsum * y1 (T0 is symmetric).  Each sum in the
// order of one 256-thread block: thread t adds elements t, t + 256, ... in
// turn, then the LDS tree.
__global__ __launch_bounds__(256) void joint_gamma_kernel(
    int levels, const double* __restrict__ p0, int n0, const double* __restrict__ p1,
    int n1, const double* __restrict__ xsum, const float* __restrict__ y1, int nx,
    double c0sq, double c1sq, float* __restrict__ g) {
  __shared__ double sh[5][256];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < n0; k += 256) {
    v[0] += p0[k];
    v[1] += p0[n0 + k];
  }
  if (levels == 2) {
    for (int k = threadIdx.x; k < n1; k += 256) {
      v[2] += p1[k];
      v[3] += p1[n1 + k];
    }
    for (int k = threadIdx.x; k < nx; k += 256) v[4] += xsum[k] * (double)y1[k];
  }
  for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double a0 = sh[0][0], b0 = sh[1][0];
  if (levels == 1) {
    const double q = a0 / b0;
    g[0] = (b0 > 0.0 && isfinite(q)) ? (float)q : 0.0f;
    g[1] = 0.0f;
    return;
  }
  const double a1 = c0sq * a0, b11 = c0sq * b0;
  const double a2 = c1sq * sh[2][0], b22 = c1sq * sh[3][0];
  const double b12 = c0sq * sh[4][0];
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

// Coarse pipeline of one level: the coarse residual (from r for level 0,
// from the level above otherwise), its local scale, the network, and the
// factor g of the coarse direction.
// The restriction of level l (block means of r, or of the finer level's
// means), on stream st.
static int run_restrict(tea_hip_chunk* c, int l, hipStream_t st) {
  FcgLevel& L = c->fcg_lv[l];
  const int nblk = (L.hc * L.wc + 255) / 256;
  Timed t(c, K_FCG_RESTRICT, st);
  if (l == 0) {
    hipLaunchKernelGGL(restrict_kernel, dim3(nblk), dim3(256), 0, st, c->geo, L.c, L.hc,
                       L.wc, c->mp_r, L.dis, L.rbar, L.part);
  } else {
    const FcgLevel& F = c->fcg_lv[l - 1];
    hipLaunchKernelGGL(restrict_coarse_kernel, dim3(nblk), dim3(256), 0, st, F.wc,
                       L.c / F.c, L.hc, L.wc, F.rbar, L.dis, L.rbar, L.part);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// The rest of level l after its restriction, on stream st: the scales and
// the network input, the network, the unscaled output (into L.tmp) with its
// dots, and -- unless the joint factors of two levels follow -- the level's
// own factor.
static int run_level(tea_hip_chunk* c, int l, hipStream_t st, bool own_factor) {
  FcgLevel& L = c->fcg_lv[l];
  const int cn = L.hc * L.wc;
  const int nblk = (cn + 255) / 256;
  {
    Timed t(c, K_FCG_RESTRICT, st);
    hipLaunchKernelGGL(box_rows_kernel, dim3(nblk + 1), dim3(256), 0, st, L.hc, L.wc,
                       L.rbar, L.dis, L.tmp, L.part, nblk, L.scale);
    hipLaunchKernelGGL(box_cols_kernel, dim3(nblk), dim3(256), 0, st, L.hc, L.wc, L.tmp,
                       L.scale, L.sig, L.rbar, L.dis, L.feat, nn::input_tensor(L.net));
  }
  {
    Timed t(c, K_FCG_NET, st);
    TEA_RETURN_IF(nn::forward_on_stream(L.net, st));
    hipLaunchKernelGGL(coarse_dots_kernel, dim3(nblk), dim3(256), 0, st, L.hc, L.wc,
                       nn::output(L.net), L.dis, L.sig, L.tmp, L.rbar, L.kc, L.part);
    if (own_factor)
      hipLaunchKernelGGL(gamma_kernel, dim3(1), dim3(256), 0, st, L.part, nblk, L.scale);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

static CoarseTerm coarse_term(const tea_hip_chunk* c, int l) {
  const FcgLevel& L = c->fcg_lv[l];
  int shift = 0;
  while ((1 << shift) < L.c) ++shift;
  return CoarseTerm{L.tmp, L.scale + 1, shift, L.wc};
}

// The coarse part of the two-level preconditioner with Jacobi sweeps as
// the coarse solve: the restriction of both levels in one pass, the sweeps
// of level 1 on the side stream beside level 0's, and the factors into
// c->fcg_g.
static int launch_sweeps_coarse(tea_hip_chunk* c) {
  const int nl = c->fcg_levels;
  FcgLevel& L0 = c->fcg_lv[0];
  FcgLevel& L1 = c->fcg_lv[1];
  {
    Timed t(c, K_FCG_RESTRICT);
    const dim3 grid((unsigned)(L0.wc + 31) / 32, (unsigned)(L0.hc + 3) / 4);
    hipLaunchKernelGGL(restrict8_kernel, grid, dim3(256), 0, c->stream, c->geo, L0.hc, L0.wc,
                       c->mp_r, L0.rbar, nl == 2 ? L1.rbar : nullptr, nullptr, nullptr,
                       nullptr, nullptr);
    TEA_HIP_TRY(hipGetLastError());
  }
  auto sweeps = [&](int l, hipStream_t st) -> int {
    FcgLevel& L = c->fcg_lv[l];
    Timed t(c, K_FCG_SWEEPS, st);
    const dim3 grid((unsigned)(L.wc + kSweepTile - 1) / kSweepTile,
                    (unsigned)(L.hc + kSweepTile - 1) / kSweepTile);
    hipLaunchKernelGGL(sweeps_kernel, grid, dim3(256), 0, st, L.hc, L.wc, c->fcg_sweeps,
                       L.kc, L.rbar, L.y, L.tpart, L.ntiles,
                       (l == 0 && nl == 2) ? c->fcg_xsum : nullptr);
    TEA_HIP_TRY(hipGetLastError());
    return TEA_HIP_OK;
  };
  if (nl == 2) {
    TEA_HIP_TRY(hipEventRecord(c->fcg_fork, c->stream));
    TEA_HIP_TRY(hipStreamWaitEvent(c->fcg_side, c->fcg_fork, 0));
    TEA_RETURN_IF(sweeps(1, c->fcg_side));
    TEA_RETURN_IF(sweeps(0, c->stream));
    TEA_HIP_TRY(hipEventRecord(c->fcg_join, c->fcg_side));
    TEA_HIP_TRY(hipStreamWaitEvent(c->stream, c->fcg_join, 0));
  } else {
    TEA_RETURN_IF(sweeps(0, c->stream));
  }
  Timed t(c, K_FCG_SWEEPS);
  hipLaunchKernelGGL(joint_gamma_kernel, dim3(1), dim3(256), 0, c->stream, nl, L0.tpart,
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
static int launch_precondition(tea_hip_chunk* c, int use_net, int slot, int* len) {
  const Geo g = c->geo;
  const int sg = c->ngroup == 1 ? 1 : 0;
  const int nl = use_net ? c->fcg_levels : 0;
  double* partials = c->partials + (size_t)slot * c->nspan;
  *len = c->nspan;
  const bool sweep = c->fcg_sweeps > 0;
  if (nl > 0 && sweep) {
    TEA_RETURN_IF(launch_sweeps_coarse(c));
  } else if (nl == 1) {
    TEA_RETURN_IF(run_restrict(c, 0, c->stream));
    TEA_RETURN_IF(run_level(c, 0, c->stream, true));
  } else if (nl == 2) {
    // Level 2 (from level 1's block means) on the side stream while level 1
    // runs on the chunk's stream; both are small grids, so their launch
    // latencies overlap.
    TEA_RETURN_IF(run_restrict(c, 0, c->stream));
    TEA_HIP_TRY(hipEventRecord(c->fcg_fork, c->stream));
    TEA_HIP_TRY(hipStreamWaitEvent(c->fcg_side, c->fcg_fork, 0));
    TEA_RETURN_IF(run_restrict(c, 1, c->fcg_side));
    TEA_RETURN_IF(run_level(c, 1, c->fcg_side, false));
    TEA_RETURN_IF(run_level(c, 0, c->stream, false));
    TEA_HIP_TRY(hipEventRecord(c->fcg_join, c->fcg_side));
    TEA_HIP_TRY(hipStreamWaitEvent(c->stream, c->fcg_join, 0));
  }
  if (nl == 2 && !sweep) {
    FcgLevel& L1 = c->fcg_lv[0];
    FcgLevel& L2 = c->fcg_lv[1];
    const int n1 = (L1.hc * L1.wc + 255) / 256, n2 = (L2.hc * L2.wc + 255) / 256;
    Timed t(c, K_FCG_NET);
    hipLaunchKernelGGL(cross_dots_kernel, dim3(n1), dim3(256), 0, c->stream, L1.hc, L1.wc,
                       L2.c / L1.c, L2.wc, L1.tmp, L2.tmp, L1.kc,
                       c->fcg_cross);
    hipLaunchKernelGGL(gamma2_kernel, dim3(1), dim3(256), 0, c->stream, L1.part, n1,
                       L2.part, n2, c->fcg_cross, n1, (double)L1.c * L1.c,
                       (double)L2.c * L2.c, L1.scale, L2.scale);
    TEA_HIP_TRY(hipGetLastError());
  }
  const CoarseTerm none{nullptr, nullptr, 0, 0};
  auto term = [&](int l) {
    const FcgLevel& L = c->fcg_lv[l];
    return sweep ? CoarseTerm{L.y, c->fcg_g + l, log2_of(L.c), L.wc} : coarse_term(c, l);
  };
  const CoarseTerm t0 = nl > 0 ? term(0) : none;
  const CoarseTerm t1 = nl > 1 ? term(1) : none;
  const int kid = nl == 0 ? K_FCG_JACOBI : K_FCG_PROLONG;
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
  for (float* b : {L.feat, L.rbar, L.kc, L.dis, L.sig, L.tmp, L.scale, L.y})
    if (b) (void)hipFree(b);
  if (L.part) (void)hipFree(L.part);
  if (L.tpart) (void)hipFree(L.tpart);
  if (L.net) tea_nn_destroy(L.net);
  L = FcgLevel();
}

// Frees every coarse level and the buffers shared between them.
static void release_levels(tea_hip_chunk* c) {
  for (FcgLevel& L : c->fcg_lv) free_level(L);
  for (double** b : {&c->fcg_cross, &c->fcg_xsum})
    if (*b) {
      (void)hipFree(*b);
      *b = nullptr;
    }
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
  for (int l = 0; l < levels; ++l) {
    const int cl = coarsen << (2 * l);
    if (nx % cl != 0 || ny % cl != 0 || (nx / cl) % 32 != 0) break;
    FcgLevel& L = c->fcg_lv[l];
    L.c = cl;
    L.hc = ny / cl;
    L.wc = nx / cl;
    const size_t cn = (size_t)L.hc * L.wc;
    const size_t nblk = (cn + 255) / 256;
    TEA_RETURN_IF(tea_nn_create(c->device, 1, L.hc, L.wc, &L.net));
    TEA_RETURN_IF(tea_nn_set_params(L.net, params, count));
    if (hipMalloc(&L.feat, 2 * cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.kc, 2 * cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.dis, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.sig, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.tmp, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.rbar, cn * sizeof(float)) != hipSuccess ||
        hipMalloc(&L.part, 2 * nblk * sizeof(double)) != hipSuccess ||
        hipMalloc(&L.scale, 2 * sizeof(float)) != hipSuccess) {
      set_error("hipMalloc of the preconditioner's coarse buffers failed");
      return TEA_HIP_ENOMEM;
    }
    c->fcg_levels = l + 1;
  }
  if (c->fcg_levels == 2) {
    const size_t n1 = ((size_t)c->fcg_lv[0].hc * c->fcg_lv[0].wc + 255) / 256;
    if (hipMalloc(&c->fcg_cross, n1 * sizeof(double)) != hipSuccess) {
      c->fcg_cross = nullptr;
      set_error("hipMalloc of the preconditioner's cross partials failed");
      return TEA_HIP_ENOMEM;
    }
    TEA_RETURN_IF(ensure_side_stream(c));
  }
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
  TEA_FLUSH_HELD(c);
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
  TEA_FLUSH_HELD(c);
  int lens[2] = {c->nspan, c->nspan};
  TEA_RETURN_IF(launch_mp_r(c, alpha, &lens[0], false));
  TEA_RETURN_IF(launch_precondition(c, use_net, 1, &lens[1]));
  double out[2];
  TEA_RETURN_IF(finish_reduction_lens(c, 2, lens, out));
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
  TEA_FLUSH_HELD(c);
  // p = z + beta p_old (the mixed-precision kernel with z in place of r),
  // w = A p, and the dots p.w and p.r.
  MpArgs a{c->mp_p[c->mp_cur], c->mp_p[c->mp_cur ^ 1], c->fcg_z,
           c->field[TEA_FIELD_U], c->mp_w, c->mp_kx, c->mp_ky, c->mp_r,
           alpha, beta, apply_u ? 1 : 0, 1};
  int len = c->nspan;
  if (use_pair_blocks(c, 1)) {
    {
      Timed t(c, K_FCG_W);
      hipLaunchKernelGGL((mp_w1_kernel<true>), pair_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, a, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    TEA_RETURN_IF(launch_span_finish(c, 2, &len, 0, true));
  } else {
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
  c->mp_cur ^= 1;
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
