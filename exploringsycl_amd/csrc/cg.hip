// libteahip.so: the CG solve (TeaLeaf/c_kernels/sycl/cg.cpp) and the shared
// solver kernels (TeaLeaf/c_kernels/sycl/solver_methods.cpp).
//
// Every FP64 expression below is evaluated in the reference's operand order
// and the library is compiled with -ffp-contract=off, so each product and sum
// rounds on its own; together with the reference's reduction tree
// (span_reduce on the device, the last passes on the host) the results are
// bitwise those of the
// FP64 CPU restatement.

#include <hip/hip_runtime.h>

#include "tea_hip_impl.h"

namespace tea {
namespace {

// The 5-point operator of TeaLeaf/shared.h:59-63 at cell i, given the
// operand at the cell (c) and at its left/right/bottom/top neighbours.
__device__ inline double smvp(double c, double l, double r, double b, double t,
                              const double* __restrict__ kx,
                              const double* __restrict__ ky, int i, int x) {
  const double kx0 = kx[i], kx1 = kx[i + 1];
  const double ky0 = ky[i], ky1 = ky[i + x];
  const double diag = (1.0 + (kx1 + kx0)) + (ky1 + ky0);
  return (diag * c - (kx1 * r + kx0 * l)) - (ky1 * t + ky0 * b);
}

// ---- cg_init (cg.cpp:7-134) ----------------------------------------------

struct InitUBody {
  double* __restrict__ p;
  double* __restrict__ r;
  double* __restrict__ u;
  double* __restrict__ w;
  const double* __restrict__ density;
  const double* __restrict__ energy;
  int coefficient, x, y;
  __device__ void operator()(int i, int kk, int jj) const {
    p[i] = 0.0;
    r[i] = 0.0;
    u[i] = energy[i] * density[i];
    if (jj > 0 && jj < y - 1 && kk > 0 && kk < x - 1)
      w[i] = coefficient == TEA_CONDUCTIVITY ? density[i] : 1.0 / density[i];
  }
};

struct InitKBody {
  const double* __restrict__ w;
  double* __restrict__ kx;
  double* __restrict__ ky;
  double rx, ry;
  int x, y, h;
  __device__ void operator()(int i, int kk, int jj) const {
    if (jj >= h && jj < y - 1 && kk >= h && kk < x - 1) {
      kx[i] = rx * (w[i - 1] + w[i]) / (2.0 * w[i - 1] * w[i]);
      ky[i] = ry * (w[i - x] + w[i]) / (2.0 * w[i - x] * w[i]);
    }
  }
};

struct InitOthersBody {
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  double* __restrict__ p;
  double* __restrict__ r;
  const double* __restrict__ u;
  double* __restrict__ w;
  int x;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    const double s = smvp(u[i], u[i - 1], u[i + 1], u[i - x], u[i + x], kx, ky,
                          i, x);
    w[i] = s;
    const double rr = u[i] - s;
    r[i] = rr;
    p[i] = rr;
    out[0] = rr * rr;
  }
};

// ---- one launch per reference kernel -------------------------------------

struct CalcWBody {  // cg.cpp:137-195
  const double* __restrict__ p;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  double* __restrict__ w;
  int x;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    const double pc = p[i];
    const double s = smvp(pc, p[i - 1], p[i + 1], p[i - x], p[i + x], kx, ky,
                          i, x);
    w[i] = s;
    out[0] = s * pc;
  }
};

struct CalcURBody {  // cg.cpp:198-254
  double* __restrict__ u;
  double* __restrict__ r;
  const double* __restrict__ p;
  const double* __restrict__ w;
  double alpha;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    u[i] = u[i] + alpha * p[i];
    const double rn = r[i] - alpha * w[i];
    r[i] = rn;
    out[0] = rn * rn;
  }
};

struct CalcPBody {  // cg.cpp:257-281
  double* __restrict__ p;
  const double* __restrict__ r;
  double beta;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) p[i] = beta * p[i] + r[i];
  }
};

// ---- fused iteration ------------------------------------------------------
//
// Kernel 1 (fused_w), in one pass over the grid:
//   u += alpha_prev * p_old                  (deferred from cg_calc_ur)
//   p_new = beta * p_old + r                 (deferred cg_calc_p)
//   p_new, u on the depth-1 halo ring         (deferred run_local_halos)
//   w = A p_new,  pw = sum p_new . w          (cg_calc_w)
// reads p_old, r, u, kx, ky and writes p_new, u, w: 8 words per cell.
// p_new at a neighbour is recomputed from p_old and r there; a neighbour on
// the halo ring is the reflection of the cell itself, so its p_new is pc.
// Kernel 2 (fused_r): r -= alpha w, rrn = r.r (3 words per cell).
// Together 11 words = 88 bytes per cell per iteration against the
// reference's 13 (calc_w 4 + calc_ur 6 + calc_p 3).

struct FusedWArgs {
  const double* p_old;
  double* p_new;
  const double* r;
  double* u;
  double* w;
  const double* kx;
  const double* ky;
  double alpha, beta;
  int apply_u, ring;
  // Pair-per-thread kernel only: the new u goes to u_out (u itself for the
  // iteration the host asked for, a spare buffer when it runs ahead), and
  // with dev_beta non-null beta is read from the device (the value the run
  // ahead formed from the device's own partials).
  double* u_out;
  const double* dev_beta;
};

// Cells of one thread are visited in batches of BATCH: all loads of a batch
// are issued before any of its arithmetic and stores, so each wave keeps
// BATCH cells' worth of reads in flight (loads cannot move past the stores
// of the previous batch, as the compiler cannot prove the arrays distinct).
// BATCH trades reads in flight per wave against waves per SIMD.
template <bool APPLY_P, int kBatch>
__global__ __launch_bounds__(kBlock) void fused_w_kernel(Geo g, FusedWArgs a,
                                                         double* partials,
                                                         int nspan,
                                                         int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  const int x = g.x;
  const int xl = g.h, xr = g.x - g.h - 1, yb = g.h, yt = g.y - g.h - 1;
  const double* __restrict__ P = a.p_old;
  const double* __restrict__ R = a.r;
  const double* __restrict__ KX = a.kx;
  const double* __restrict__ KY = a.ky;
  double prod[1][kItems];

#pragma unroll
  for (int b0 = 0; b0 < kItems; b0 += kBatch) {
    int ci[kBatch], ck[kBatch], cj[kBatch];
    bool in[kBatch];
    double po[kBatch], pl[kBatch], pr[kBatch], pb[kBatch], pt[kBatch];
    double rc[kBatch], rl[kBatch], rr[kBatch], rb[kBatch], rt[kBatch];
    double uc[kBatch], kx0[kBatch], kx1[kBatch], ky0[kBatch], ky1[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int e = base + (b0 + j) * kBlock + threadIdx.x;
      const int jj = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
      ci[j] = e;
      cj[j] = jj;
      ck[j] = e - jj * x;
      in[j] = e < g.n && interior(g, ck[j], jj);
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int i = in[j] ? ci[j] : x + 1;  // any interior-safe index
      po[j] = P[i];
      pl[j] = P[i - 1];
      pr[j] = P[i + 1];
      pb[j] = P[i - x];
      pt[j] = P[i + x];
      if (APPLY_P) {
        rc[j] = R[i];
        rl[j] = R[i - 1];
        rr[j] = R[i + 1];
        rb[j] = R[i - x];
        rt[j] = R[i + x];
      }
      kx0[j] = KX[i];
      kx1[j] = KX[i + 1];
      ky0[j] = KY[i];
      ky1[j] = KY[i + x];
      uc[j] = a.apply_u ? a.u[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int i = ci[j], kk = ck[j], jj = cj[j];
      double pc, l, rgt, bot, top;
      if (APPLY_P) {
        // p_new = beta p_old + r, here and at the four neighbours; a
        // neighbour on the reflective halo ring is this cell's mirror image.
        pc = a.beta * po[j] + rc[j];
        l = kk == xl ? pc : a.beta * pl[j] + rl[j];
        rgt = kk == xr ? pc : a.beta * pr[j] + rr[j];
        bot = jj == yb ? pc : a.beta * pb[j] + rb[j];
        top = jj == yt ? pc : a.beta * pt[j] + rt[j];
      } else {
        pc = po[j];
        l = pl[j];
        rgt = pr[j];
        bot = pb[j];
        top = pt[j];
      }
      const double diag = (1.0 + (kx1[j] + kx0[j])) + (ky1[j] + ky0[j]);
      const double s = (diag * pc - (kx1[j] * rgt + kx0[j] * l)) -
                       (ky1[j] * top + ky0[j] * bot);
      prod[0][b0 + j] = in[j] ? s * pc : 0.0;
      if (!in[j]) continue;
      a.w[i] = s;
      const bool L = kk == xl, Rr = kk == xr, B = jj == yb, T = jj == yt;
      const bool edge = L | Rr | B | T;
      if (a.apply_u) {
        const double un = uc[j] + a.alpha * po[j];
        a.u[i] = un;
        if (a.ring && edge) {
          if (L) a.u[i - 1] = un;
          if (Rr) a.u[i + 1] = un;
          if (B) a.u[i - x] = un;
          if (T) a.u[i + x] = un;
          if (L && B) a.u[i - 1 - x] = un;
          if (Rr && B) a.u[i + 1 - x] = un;
          if (L && T) a.u[i - 1 + x] = un;
          if (Rr && T) a.u[i + 1 + x] = un;
        }
      }
      if (APPLY_P) {
        double* __restrict__ Q = a.p_new;
        Q[i] = pc;
        if (edge) {
          // With the halo refresh folded in, the ring gets the new values;
          // without it, p keeps its old halo.
          if (L) Q[i - 1] = a.ring ? pc : pl[j];
          if (Rr) Q[i + 1] = a.ring ? pc : pr[j];
          if (B) Q[i - x] = a.ring ? pc : pb[j];
          if (T) Q[i + x] = a.ring ? pc : pt[j];
          if (L && B) Q[i - 1 - x] = a.ring ? pc : P[i - 1 - x];
          if (Rr && B) Q[i + 1 - x] = a.ring ? pc : P[i + 1 - x];
          if (L && T) Q[i - 1 + x] = a.ring ? pc : P[i - 1 + x];
          if (Rr && T) Q[i + 1 + x] = a.ring ? pc : P[i + 1 + x];
        }
      }
    }
  }
  span_reduce<1>(prod, partials, nspan, span, single_group != 0);
}

// ---- two cells per lane --------------------------------------------------
//
// When x is even every cell pair (2m, 2m+1) lies in one row and starts on a
// 16-byte boundary, so p, r, u, kx, ky and the rows above and below are read
// with one 16-byte load per pair, and the left/right neighbours come from
// the adjacent lanes (lane l-1 holds 2m-1, lane l+1 holds 2m+2); only lanes
// 0 and 63 load theirs.  A wave covers 128 consecutive cells = two groups
// of the reduction tree (lanes 0-31 and 32-63); a 256-thread block covers a
// 4096-cell span in 8 slots of 512 cells.  The first tree level is the
// in-lane pair sum, then the butterfly runs over each 32-lane half.

__device__ inline void span_reduce_pairs(double (&v)[kSlots], double* partials,
                                         int nspan, int span,
                                         bool single_group) {
  double (&vv)[1][kSlots] = *reinterpret_cast<double (*)[1][kSlots]>(&v);
  span_reduce_pairs_n<1>(vv, partials, nspan, span, single_group);
}

// One cell pair of the fused iteration: its indices, its loads and its
// arithmetic, shared by the span-per-block and pair-per-thread kernels so
// both evaluate every cell identically.
struct FusedPair {
  int e0, kk, jj;
  bool in0, in1, any;
  double2 P0, Pt, Pb, R0, Rt, Rb, KX, KY, KYt, U;
  // Wave-edge neighbours: lane 0 holds p and r at e0 - 1, lane 63 p, r and
  // kx at e0 + 2 (the other lanes take theirs from adjacent lanes).
  double PN, RN, KXN;
};

__device__ inline void fused_pair_index(const Geo& g, int e, FusedPair& s) {
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  s.e0 = e;
  s.jj = row;
  s.kk = e - row * g.x;
  s.in0 = e < g.n && interior(g, s.kk, row);
  s.in1 = e < g.n && interior(g, s.kk + 1, row);
  s.any = s.in0 || s.in1;
}

// The pair's own p, r and kx are also the left/right neighbours the adjacent
// lanes take by shuffle, so they are read whenever the pair is in range; the
// rows above and below only for interior pairs.  x + 2 is an even index whose
// +-x rows are in range.
template <bool APPLY_P>
__device__ inline void fused_pair_load(const Geo& g, const FusedWArgs& a,
                                       FusedPair& s) {
  const int x = g.x;
  const double* __restrict__ P = a.p_old;
  const double* __restrict__ R = a.r;
  const int ic = s.e0 < g.n ? s.e0 : x + 2;
  const int ir = s.any ? s.e0 : x + 2;
  s.P0 = ld2(P, ic);
  s.Pt = ld2(P, ir + x);
  s.Pb = ld2(P, ir - x);
  if (APPLY_P) {
    s.R0 = ld2(R, ic);
    s.Rt = ld2(R, ir + x);
    s.Rb = ld2(R, ir - x);
  }
  s.KX = ld2(a.kx, ic);
  s.KY = ld2(a.ky, ir);
  s.KYt = ld2(a.ky, ir + x);
  if (a.apply_u) s.U = ld2_nt(a.u, ir);
  // Issued with the pair's own loads so the wave waits once; ir - 1 and
  // ir + 2 are in range (ir is an interior pair's index or x + 2).
  // Lanes 0 and 63 share each load instruction, at their own addresses
  // (lane 0's kx value is not used).
  const int lane = threadIdx.x & (kWave - 1);
  if (lane == 0 || lane == kWave - 1) {
    const int ni = lane == 0 ? ir - 1 : ir + 2;
    s.PN = P[ni];
    if (APPLY_P) s.RN = R[ni];
    s.KXN = a.kx[ni];
  }
}

// Computes and stores the pair (w, u, p_new and their halo ring) and returns
// the in-lane pair sum of p_new . w (the first level of the reduction tree).
template <bool APPLY_P>
__device__ inline double fused_pair_compute(const Geo& g, const FusedWArgs& a,
                                            const FusedPair& s, int lane) {
  const int x = g.x;
  const int xl = g.h, xr = g.x - g.h - 1, yb = g.h, yt = g.y - g.h - 1;
  const double* __restrict__ P = a.p_old;
  const double beta = a.dev_beta ? *a.dev_beta : a.beta;
  const int i = s.e0;
  // Neighbours across the pair boundary, from the adjacent lanes.
  double pl = from_prev_lane(s.P0.y), pr = from_next_lane(s.P0.x);
  double rl = 0.0, rr = 0.0;
  if (APPLY_P) {
    rl = from_prev_lane(s.R0.y);
    rr = from_next_lane(s.R0.x);
  }
  double kxr = from_next_lane(s.KX.x);
  if (lane == 0) {
    pl = s.PN;
    if (APPLY_P) rl = s.RN;
  }
  if (lane == kWave - 1) {
    pr = s.PN;
    if (APPLY_P) rr = s.RN;
    kxr = s.KXN;
  }
  const int ka = s.kk, kb = s.kk + 1, row = s.jj;
  double ca, cb, la, ra, lb, rb2, ba, ta, bb, tb;
  if (APPLY_P) {
    ca = beta * s.P0.x + s.R0.x;
    cb = beta * s.P0.y + s.R0.y;
    la = ka == xl ? ca : beta * pl + rl;
    ra = ka == xr ? ca : cb;
    lb = kb == xl ? cb : ca;
    rb2 = kb == xr ? cb : beta * pr + rr;
    ba = row == yb ? ca : beta * s.Pb.x + s.Rb.x;
    ta = row == yt ? ca : beta * s.Pt.x + s.Rt.x;
    bb = row == yb ? cb : beta * s.Pb.y + s.Rb.y;
    tb = row == yt ? cb : beta * s.Pt.y + s.Rt.y;
  } else {
    ca = s.P0.x;
    cb = s.P0.y;
    la = pl;
    ra = cb;
    lb = ca;
    rb2 = pr;
    ba = s.Pb.x;
    ta = s.Pt.x;
    bb = s.Pb.y;
    tb = s.Pt.y;
  }
  const double kxa0 = s.KX.x, kxa1 = s.KX.y, kxb0 = s.KX.y, kxb1 = kxr;
  const double diag_a = (1.0 + (kxa1 + kxa0)) + (s.KYt.x + s.KY.x);
  const double diag_b = (1.0 + (kxb1 + kxb0)) + (s.KYt.y + s.KY.y);
  const double sa = (diag_a * ca - (kxa1 * ra + kxa0 * la)) -
                    (s.KYt.x * ta + s.KY.x * ba);
  const double sb = (diag_b * cb - (kxb1 * rb2 + kxb0 * lb)) -
                    (s.KYt.y * tb + s.KY.y * bb);
  const double psum = (s.in0 ? sa * ca : 0.0) + (s.in1 ? sb * cb : 0.0);
  if (!s.any) return psum;
  st2(a.w, i, s.in0, s.in1, sa, sb);
  double ua = 0.0, ub = 0.0;
  if (a.apply_u) {
    ua = s.U.x + a.alpha * s.P0.x;
    ub = s.U.y + a.alpha * s.P0.y;
    st2_nt(a.u_out, i, s.in0, s.in1, ua, ub);
  }
  if (APPLY_P) st2_nt(a.p_new, i, s.in0, s.in1, ca, cb);
  // Halo ring of boundary-adjacent cells (rare; per cell).
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const bool inc = c == 0 ? s.in0 : s.in1;
    const int ic = i + c, kc = ka + c;
    const bool L = kc == xl, Rg = kc == xr, B = row == yb, T = row == yt;
    if (!inc || !(L | Rg | B | T)) continue;
    const double pv = c == 0 ? ca : cb;
    const double uv = c == 0 ? ua : ub;
    int ring[8];
    int nr = 0;
    if (L) ring[nr++] = ic - 1;
    if (Rg) ring[nr++] = ic + 1;
    if (B) ring[nr++] = ic - x;
    if (T) ring[nr++] = ic + x;
    if (L && B) ring[nr++] = ic - 1 - x;
    if (Rg && B) ring[nr++] = ic + 1 - x;
    if (L && T) ring[nr++] = ic - 1 + x;
    if (Rg && T) ring[nr++] = ic + 1 + x;
    for (int q = 0; q < nr; ++q) {
      if (a.apply_u && (a.ring || a.u_out != a.u))
        a.u_out[ring[q]] = a.ring ? uv : a.u[ring[q]];
      if (APPLY_P) a.p_new[ring[q]] = a.ring ? pv : P[ring[q]];
    }
  }
  return psum;
}

// One span per block, 8 pairs per thread in batches of SB.
template <bool APPLY_P, int SB>
__global__ __launch_bounds__(kBlock) void fused_w2_kernel(Geo g, FusedWArgs a,
                                                          double* partials,
                                                          int nspan,
                                                          int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  const int lane = threadIdx.x & (kWave - 1);
  double psum[kSlots];
#pragma unroll
  for (int s0 = 0; s0 < kSlots; s0 += SB) {
    FusedPair s[SB];
#pragma unroll
    for (int j = 0; j < SB; ++j)
      fused_pair_index(g, base + (s0 + j) * (2 * kBlock) + 2 * threadIdx.x, s[j]);
#pragma unroll
    for (int j = 0; j < SB; ++j) fused_pair_load<APPLY_P>(g, a, s[j]);
#pragma unroll
    for (int j = 0; j < SB; ++j)
      psum[s0 + j] = fused_pair_compute<APPLY_P>(g, a, s[j], lane);
  }
  span_reduce_pairs(psum, partials, nspan, span, single_group != 0);
}

// One pair per thread: a 256-thread block covers 512 cells (8 groups of the
// reduction tree) and stores their 8-leaf subtree sum in sub[block];
// span_finish_kernel then combines the 8 blocks of each span.  Small blocks
// keep 4x more waves per SIMD in flight than the span-per-block kernel.
template <bool APPLY_P, int KS>
__global__ __launch_bounds__(kBlock * KS) void fused_w1_kernel(Geo g, FusedWArgs a,
                                                               double* sub, int nblk, int S) {
  int tid;
  const int blk = stacked_block<KS, kBlock>(S, &tid);
  if (blk >= nblk) return;
  const int lane = threadIdx.x & (kWave - 1);
  FusedPair s;
  fused_pair_index(g, blk * kPairBlock + 2 * tid, s);
  fused_pair_load<APPLY_P>(g, a, s);
  double v[1] = {fused_pair_compute<APPLY_P>(g, a, s, lane)};
  wave_subtree_pairs<1>(v, sub, nblk, blk, tid / kWave);
}

// Launches fused_w1_kernel with stack_factor(c) blocks one row apart per
// workgroup (stacked_block, tea_hip_impl.h): their rows above and below come
// mostly from the CU's L1.  At 4096^2 two blocks per workgroup take the FP64
// iteration from 269.2 to 267.6 us and the mixed-precision w-step from 83.2
// to 79.0 us (four: 270.6 us).
template <bool APPLY_P>
void launch_fused_w1(tea_hip_chunk* c, const FusedWArgs& a) {
  const int S = stack_rows(c), ks = stack_factor(c);
  const dim3 grid = stacked_grid(c, ks);
  if (ks == 1)
    hipLaunchKernelGGL((fused_w1_kernel<APPLY_P, 1>), grid, dim3(kBlock), 0, c->stream,
                       c->geo, a, c->sub, c->nblk, S);
  else if (ks == 2)
    hipLaunchKernelGGL((fused_w1_kernel<APPLY_P, 2>), grid, dim3(2 * kBlock), 0, c->stream,
                       c->geo, a, c->sub, c->nblk, S);
  else
    hipLaunchKernelGGL((fused_w1_kernel<APPLY_P, 4>), grid, dim3(4 * kBlock), 0, c->stream,
                       c->geo, a, c->sub, c->nblk, S);
}

__global__ __launch_bounds__(kBlock) void fused_r2_kernel(
    Geo g, double* __restrict__ r, const double* __restrict__ w, double alpha,
    double* partials, int nspan, int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  double2 rv[kSlots], wv[kSlots];
  bool in0[kSlots], in1[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const int e = base + k * (2 * kBlock) + 2 * threadIdx.x;
    const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
    const int kk = e - row * g.x;
    in0[k] = e < g.n && interior(g, kk, row);
    in1[k] = e < g.n && interior(g, kk + 1, row);
    const int i = (in0[k] || in1[k]) ? e : 0;
    rv[k] = ld2_nt(r, i);
    wv[k] = ld2_nt(w, i);
  }
  double psum[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const double ra = rv[k].x - alpha * wv[k].x;
    const double rb = rv[k].y - alpha * wv[k].y;
    psum[k] = (in0[k] ? ra * ra : 0.0) + (in1[k] ? rb * rb : 0.0);
    st2_nt(r, base + k * (2 * kBlock) + 2 * threadIdx.x, in0[k], in1[k], ra, rb);
  }
  span_reduce_pairs(psum, partials, nspan, span, single_group != 0);
}

// fused_r2's arithmetic with one pair per thread; the block stores its
// 8-leaf subtree of r.r (see fused_w1_kernel).  The new r goes to r_out (r
// itself, or a spare buffer when the iteration runs ahead), and with
// dev_alpha non-null alpha is read from the device.
__global__ __launch_bounds__(kBlock) void fused_r1_kernel(
    Geo g, const double* r, double* r_out, const double* __restrict__ w, double alpha,
    const double* __restrict__ dev_alpha, double* sub, int nblk) {
  if (dev_alpha) alpha = *dev_alpha;
  const int blk = blockIdx.x;
  const int e = blk * kPairBlock + 2 * threadIdx.x;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * g.x;
  const bool in0 = e < g.n && interior(g, kk, row);
  const bool in1 = e < g.n && interior(g, kk + 1, row);
  const int i = (in0 || in1) ? e : 0;
  const double2 rv = ld2_nt(r, i), wv = ld2_nt(w, i);
  const double ra = rv.x - alpha * wv.x;
  const double rb = rv.y - alpha * wv.y;
  double v[1] = {(in0 ? ra * ra : 0.0) + (in1 ? rb * rb : 0.0)};
  st2_nt(r_out, e, in0, in1, ra, rb);
  wave_subtree_pairs<1>(v, sub, nblk, blk);
}

// r -= alpha w and r.r: all 16 cells' loads first, then the updates.
__global__ __launch_bounds__(kBlock) void fused_r_kernel(
    Geo g, double* __restrict__ r, const double* __restrict__ w, double alpha,
    double* partials, int nspan, int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  double rv[kItems], wv[kItems];
  bool in[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int e = base + k * kBlock + threadIdx.x;
    const int jj = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
    in[k] = e < g.n && interior(g, e - jj * g.x, jj);
    const int i = in[k] ? e : 0;
    rv[k] = r[i];
    wv[k] = w[i];
  }
  double prod[1][kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const double rn = rv[k] - alpha * wv[k];
    prod[0][k] = in[k] ? rn * rn : 0.0;
    if (in[k]) r[base + k * kBlock + threadIdx.x] = rn;
  }
  span_reduce<1>(prod, partials, nspan, span, single_group != 0);
}

struct ApplyPendingBody {
  double* __restrict__ u;
  double* __restrict__ p;
  const double* __restrict__ r;
  double alpha, beta;
  int apply_u, apply_p;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const double po = p[i];
    if (apply_u) u[i] = u[i] + alpha * po;
    if (apply_p) p[i] = beta * po + r[i];
  }
};

// ---- shared solver kernels (solver_methods.cpp) ---------------------------

struct CopyUBody {
  const double* __restrict__ u;
  double* __restrict__ u0;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) u0[i] = u[i];
  }
};

struct ResidualBody {
  const double* __restrict__ u;
  const double* __restrict__ u0;
  double* __restrict__ r;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const int x = g.x;
    r[i] = u0[i] - smvp(u[i], u[i - 1], u[i + 1], u[i - x], u[i + x], kx, ky,
                        i, x);
  }
};

struct NormBody {
  const double* __restrict__ b;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    const double v = b[i];
    out[0] = v * v;
  }
};

// The 2-norm over cell pairs, one pair per lane (even rows, more than one
// group): 8 bytes per cell, the sum tree's first level in-lane.
__global__ __launch_bounds__(kBlock) void norm_pair_kernel(
    Geo g, const double* __restrict__ b, double* sub, int nblk) {
  const PairAt q = pair_at(g, blockIdx.x * kPairBlock + 2 * threadIdx.x);
  const double2 v = ld2_nt(b, q.ic);
  double s[1] = {(q.in0 ? v.x * v.x : 0.0) + (q.in1 ? v.y * v.y : 0.0)};
  wave_subtree_pairs<1>(s, sub, nblk, blockIdx.x);
}

struct FinaliseBody {
  const double* __restrict__ u;
  const double* __restrict__ density;
  double* __restrict__ energy;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) energy[i] = u[i] / density[i];
  }
};

template <class Body>
int launch_each(tea_hip_chunk* c, int kid, const Body& body) {
  {
    Timed t(c, kid);
    hipLaunchKernelGGL((each_kernel<Body>), flat_grid(c->geo.n), dim3(kBlock),
                       0, c->stream, c->geo, body);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

template <class Body>
int launch_span1(tea_hip_chunk* c, int kid, const Body& body, double* out) {
  {
    Timed t(c, kid);
    hipLaunchKernelGGL((span_kernel<1, Body>), span_grid(c), dim3(kBlock), 0,
                       c->stream, c->geo, body, c->partials, c->nspan,
                       c->ngroup == 1 ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  return finish_reduction(c, 1, out);
}

// Makes the second p buffer's cells outside the interior and the halo ring
// equal to p's (the fused kernel writes only the interior and the ring).
int sync_p_alt(tea_hip_chunk* c) {
  if (c->p_alt_synced) return TEA_HIP_OK;
  TEA_HIP_TRY(hipMemcpyAsync(c->p_alt, c->field[TEA_FIELD_P],
                             (size_t)c->geo.n * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
  c->p_alt_synced = true;
  return TEA_HIP_OK;
}

// ---- run-ahead of the fused CG iteration ----------------------------------
//
// The drop-in API returns p.w and r.r to the host, which forms alpha =
// rro / pw and beta = rrn / rro before the next pass can start
// (TeaLeaf/drivers/cg_driver.c:87-113): two host round trips per iteration,
// each costing the wait for the finisher's values, the host's arithmetic and
// a launch (about 13.5 us of a 272 us iteration at 4096^2).  The device
// forms the same scalars itself: the last block of each finisher runs the
// host's final tree passes over the grid's partials and divides
// (span_finish_kernel's run-ahead tail, core.hip), and the next pass of the
// iteration is already queued behind it, writing only spare buffers:
//   after cg_calc_w  (the r-step): r - alpha w into r_alt, and r.r;
//   after cg_calc_ur (the w-step): u + alpha p into u_alt, p = beta p + r
//     into p_alt with the U/P halo ring, w = A p into w_alt, p.w; then the
//     next r-step.
// The host's next call takes a pass over (exchanging the buffers) when its
// scalar is bitwise the one the device used (read back with the sums) and,
// for the w-step, the deferred work is exactly the u and p updates and the
// U/P ring the pass applied; anything else drops the run-ahead, which wrote
// nothing but spare buffers, device scalars and its own host slots.  The
// results are those of the iteration without it, bit for bit.  Each step's
// sum and scalar go to one of two host slot sets, alternating, so the next
// step is queued before the host has read the last one.

constexpr int kRaNone = 0, kRaR = 1, kRaWR = 2;
constexpr int kRaW = 0, kRaRStep = 1;  // slot kinds

int ra_len(const tea_hip_chunk* c) {
  return c->nspan == 1 ? 1 : (c->nspan + kWave - 1) / kWave;
}

bool ra_possible(const tea_hip_chunk* c) {
  return c->ra_enabled && c->fused && c->poll && c->pairs && c->ra_rro_known &&
         c->ra_fail < 3 && use_pair_blocks(c, 1) && use_pair_blocks(c, 2) &&
         ra_len(c) <= kWave * kWave;
}

double* ra_host_slot(tea_hip_chunk* c, int set, int kind) {
  return c->ra_host + 2 * (2 * set + kind);
}

double* ra_dev_slot(tea_hip_chunk* c, int set, int kind) {
  return c->ra_host_dev + 2 * (2 * set + kind);
}

int ensure_ra(tea_hip_chunk* c) {
  if (c->ra_host) return TEA_HIP_OK;
  const size_t bytes = (size_t)c->geo.n * sizeof(double);
  for (double** b : {&c->r_alt, &c->u_alt, &c->w_alt}) {
    if (hipMalloc(b, bytes) != hipSuccess) {
      *b = nullptr;
      set_error("hipMalloc of a run-ahead buffer failed");
      return TEA_HIP_ENOMEM;
    }
  }
  // [0..2] scalars, then the finisher's partials, then the blocks-done count.
  const size_t nd = 3 + (size_t)ra_len(c);
  if (hipMalloc(&c->ra_dev, (nd + 1) * sizeof(double)) != hipSuccess) {
    c->ra_dev = nullptr;
    set_error("hipMalloc of the run-ahead scalars failed");
    return TEA_HIP_ENOMEM;
  }
  c->ra_count = reinterpret_cast<unsigned*>(c->ra_dev + nd);
  TEA_HIP_TRY(hipMemsetAsync(c->ra_count, 0, sizeof(unsigned), c->stream));
  double* host = nullptr;
  if (hipHostMalloc(&host, 8 * sizeof(double),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    set_error("hipHostMalloc of the run-ahead slots failed");
    return TEA_HIP_ENOMEM;
  }
  TEA_HIP_TRY(hipHostGetDevicePointer((void**)&c->ra_host_dev, host, 0));
  c->ra_host = host;
  return TEA_HIP_OK;
}

// The finisher of a run-ahead pass (into the device's partials) with its
// scalar (into host slot (set, kind)).
int ra_finish(tea_hip_chunk* c, int set, int kind, double rro, bool from_dev) {
  if (c->ra_dirty) {
    // A dropped run-ahead may still be queued and write these slots.
    TEA_HIP_TRY(hipStreamSynchronize(c->stream));
    c->ra_dirty = false;
  }
  arm_host_values(ra_host_slot(c, set, kind), 2);
  const RaTail tail{rro, from_dev ? 1 : 0, kind == kRaW ? 0 : 1, c->ra_dev,
                    ra_dev_slot(c, set, kind), c->ra_count};
  return launch_span_finish_ra(c, c->ra_dev + 3, tail);
}

// The r-step behind a w-step whose p.w the slot set's ratio turns into
// alpha: r_alt = r - alpha w, r.r, beta.
int ra_launch_r(tea_hip_chunk* c, int set, const double* w, double rro, bool from_dev) {
  {
    Timed t(c, K_CG_FUSED_R);
    hipLaunchKernelGGL(fused_r1_kernel, pair_grid(c), dim3(kBlock), 0, c->stream, c->geo,
                       c->field[TEA_FIELD_R], c->r_alt, w, 0.0, c->ra_dev, c->sub, c->nblk);
  }
  TEA_HIP_TRY(hipGetLastError());
  return ra_finish(c, set, kRaRStep, rro, from_dev);
}

// Waits for a slot's sum and scalar.
int ra_wait(tea_hip_chunk* c, int set, int kind, double* sum, double* ratio) {
  const double* h = ra_host_slot(c, set, kind);
  TEA_RETURN_IF(wait_host_values(c, h, 2));
  *sum = h[0];
  *ratio = h[1];
  return TEA_HIP_OK;
}

void ra_mismatch(tea_hip_chunk* c) {
  c->ra = kRaNone;
  c->ra_dirty = true;
  c->ra_fail += 1;
}

template <class T>
void swap_ptr(T*& a, T*& b) {
  T* t = a;
  a = b;
  b = t;
}

}  // namespace

int ra_discard(tea_hip_chunk* c) {
  if (c->ra != kRaNone) {
    c->ra = kRaNone;
    c->ra_dirty = true;
  }
  return TEA_HIP_OK;
}

int apply_pending(tea_hip_chunk* c) {
  TEA_RETURN_IF(apply_cg_pending(c));
  return flush_held(c);
}

int apply_cg_pending(tea_hip_chunk* c) {
  if (c->ra) TEA_RETURN_IF(ra_discard(c));
  if (c->hold) TEA_RETURN_IF(apply_mp_hold(c));
  if (!c->pend_u && !c->pend_p && !c->pend_halo) return TEA_HIP_OK;
  double** f = c->field;
  const bool halo_u = c->pend_u;  // the deferred ring update includes U
  if (c->pend_u || c->pend_p) {
    ApplyPendingBody body{f[TEA_FIELD_U], f[TEA_FIELD_P], f[TEA_FIELD_R],
                          c->pend_alpha, c->pend_beta, c->pend_u ? 1 : 0,
                          c->pend_p ? 1 : 0, c->geo};
    TEA_RETURN_IF(launch_each(c, K_CG_APPLY_PENDING, body));
    if (c->pend_p) c->p_alt_synced = false;
  }
  const bool halo = c->pend_halo;
  c->pend_u = c->pend_p = c->pend_halo = false;
  if (halo) {
    static const int ext[4] = {-1, -1, -1, -1};
    int fields[TEA_HALO_NUM] = {0};
    fields[TEA_HALO_U] = halo_u ? 1 : 0;
    fields[TEA_HALO_P] = 1;
    TEA_RETURN_IF(launch_halos(c, 1, ext, fields));
    c->p_alt_synced = false;
  }
  return TEA_HIP_OK;
}

}  // namespace tea

using namespace tea;

extern "C" {

int tea_hip_cg_init(tea_hip_chunk* c, int coefficient, double rx, double ry,
                    double* rro) {
  TEA_ENTER(c);
  TEA_CHECK_ARG(rro != nullptr, "null rro");
  TEA_CHECK_ARG(coefficient == TEA_CONDUCTIVITY ||
                    coefficient == TEA_RECIP_CONDUCTIVITY,
                "unknown conductivity model");
  TEA_RETURN_IF(apply_pending(c));
  c->faces_ok_u = c->faces_ok_p = false;
  double** f = c->field;
  const Geo g = c->geo;
  TEA_RETURN_IF(launch_each(
      c, K_CG_INIT_U,
      InitUBody{f[TEA_FIELD_P], f[TEA_FIELD_R], f[TEA_FIELD_U], f[TEA_FIELD_W],
                f[TEA_FIELD_DENSITY], f[TEA_FIELD_ENERGY], coefficient, g.x,
                g.y}));
  TEA_RETURN_IF(launch_each(c, K_CG_INIT_K,
                            InitKBody{f[TEA_FIELD_W], f[TEA_FIELD_KX],
                                      f[TEA_FIELD_KY], rx, ry, g.x, g.y, g.h}));
  c->p_alt_synced = false;
  TEA_RETURN_IF(launch_span1(c, K_CG_INIT_OTHERS,
                             InitOthersBody{f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                                            f[TEA_FIELD_P], f[TEA_FIELD_R],
                                            f[TEA_FIELD_U], f[TEA_FIELD_W], g.x},
                             rro));
  // The host's first rro (TeaLeaf/drivers/cg_driver.c:45-52), for the
  // run-ahead's first alpha.
  c->ra_rro = *rro;
  c->ra_rro_known = true;
  c->ra_fail = 0;
  return TEA_HIP_OK;
}

int tea_hip_cg_calc_w(tea_hip_chunk* c, double* pw) {
  TEA_ENTER_RA(c);
  TEA_CHECK_ARG(pw != nullptr, "null pw");
  double** f = c->field;
  const Geo g = c->geo;
  if (c->ra == kRaWR) {
    // The w-step in flight applied exactly the deferred work: take it over.
    if (c->pend_u && c->pend_p && c->pend_halo && !c->held && c->pre_halo < 0 &&
        c->pend_alpha == c->ra_w_alpha && c->pend_beta == c->ra_beta) {
      swap_ptr(f[TEA_FIELD_P], c->p_alt);
      swap_ptr(f[TEA_FIELD_U], c->u_alt);
      swap_ptr(f[TEA_FIELD_W], c->w_alt);
      c->pend_u = c->pend_p = c->pend_halo = false;
      c->ra = kRaR;  // its r-step is queued behind it
      c->ra_commits += 1;
      c->ra_fail = 0;
      double sum = 0.0, alpha = 0.0;
      TEA_RETURN_IF(ra_wait(c, c->ra_set, kRaW, &sum, &alpha));
      c->ra_alpha = alpha;
      *pw = sum;
      return TEA_HIP_OK;
    }
    ra_mismatch(c);
  } else if (c->ra != kRaNone) {
    ra_mismatch(c);
  }
  TEA_FLUSH_HELD(c);
  if (!c->fused) {
    TEA_RETURN_IF(apply_pending(c));
    return launch_span1(c, K_CG_CALC_W,
                        CalcWBody{f[TEA_FIELD_P], f[TEA_FIELD_KX],
                                  f[TEA_FIELD_KY], f[TEA_FIELD_W], g.x},
                        pw);
  }
  const bool apply_p = c->pend_p;
  if (apply_p) TEA_RETURN_IF(sync_p_alt(c));
  FusedWArgs args{f[TEA_FIELD_P], c->p_alt,        f[TEA_FIELD_R],
                  f[TEA_FIELD_U], f[TEA_FIELD_W],  f[TEA_FIELD_KX],
                  f[TEA_FIELD_KY], c->pend_alpha,  c->pend_beta,
                  c->pend_u ? 1 : 0, c->pend_halo ? 1 : 0, f[TEA_FIELD_U], nullptr};
  c->pend_u = c->pend_p = c->pend_halo = false;
  int plen = c->nspan;
  if (use_pair_blocks(c, 1)) {
    const bool ahead = ra_possible(c);
    if (ahead) {
      // Start a run-ahead: the spare buffers take the cells outside the
      // interior (and the ring) that no pass writes.
      TEA_RETURN_IF(ensure_ra(c));
      const size_t bytes = (size_t)g.n * sizeof(double);
      TEA_HIP_TRY(hipMemcpyAsync(c->r_alt, f[TEA_FIELD_R], bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
      TEA_HIP_TRY(hipMemcpyAsync(c->u_alt, f[TEA_FIELD_U], bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
      TEA_HIP_TRY(hipMemcpyAsync(c->w_alt, f[TEA_FIELD_W], bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
    }
    {
      Timed t(c, K_CG_FUSED_W);
      if (apply_p)
        launch_fused_w1<true>(c, args);
      else
        launch_fused_w1<false>(c, args);
    }
    TEA_HIP_TRY(hipGetLastError());
    if (ahead) {
      const int set = c->ra_set ^ 1;
      TEA_RETURN_IF(ra_finish(c, set, kRaW, c->ra_rro, false));
      TEA_RETURN_IF(ra_launch_r(c, set, f[TEA_FIELD_W], c->ra_rro, false));
      if (apply_p) swap_ptr(f[TEA_FIELD_P], c->p_alt);
      c->ra = kRaR;
      c->ra_set = set;
      double sum = 0.0, alpha = 0.0;
      TEA_RETURN_IF(ra_wait(c, set, kRaW, &sum, &alpha));
      c->ra_alpha = alpha;
      *pw = sum;
      return TEA_HIP_OK;
    }
    TEA_RETURN_IF(launch_span_finish(c, 1, &plen, 0, true));
  } else {
    Timed t(c, K_CG_FUSED_W);
    const int sg = c->ngroup == 1 ? 1 : 0;
#define TEA_FUSED_W(AP, B)                                                   \
  hipLaunchKernelGGL((fused_w_kernel<AP, B>), span_grid(c), dim3(kBlock), 0, \
                     c->stream, g, args, c->partials, c->nspan, sg)
#define TEA_FUSED_W2(AP, B)                                                   \
  hipLaunchKernelGGL((fused_w2_kernel<AP, B>), span_grid(c), dim3(kBlock), 0, \
                     c->stream, g, args, c->partials, c->nspan, sg)
    if (c->pairs) {
      if (!apply_p) TEA_FUSED_W2(false, 2);
      else if (c->batch == 1) TEA_FUSED_W2(true, 1);
      else if (c->batch == 2) TEA_FUSED_W2(true, 2);
      else TEA_FUSED_W2(true, 4);
    } else {
      if (!apply_p) TEA_FUSED_W(false, 4);
      else if (c->batch == 2) TEA_FUSED_W(true, 2);
      else if (c->batch == 8) TEA_FUSED_W(true, 8);
      else TEA_FUSED_W(true, 4);
    }
#undef TEA_FUSED_W
#undef TEA_FUSED_W2
  }
  TEA_HIP_TRY(hipGetLastError());
  if (apply_p) {
    double* t = f[TEA_FIELD_P];
    f[TEA_FIELD_P] = c->p_alt;
    c->p_alt = t;
  }
  return finish_reduction_n(c, 1, plen, pw);
}

int tea_hip_cg_calc_ur(tea_hip_chunk* c, double alpha, double* rrn) {
  TEA_ENTER_RA(c);
  TEA_CHECK_ARG(rrn != nullptr, "null rrn");
  double** f = c->field;
  if (c->ra == kRaR) {
    // The r-step in flight used the device's alpha: take it over when the
    // host's is the same.
    if (alpha == c->ra_alpha && !c->pend_u && !c->pend_p && !c->pend_halo && !c->held &&
        c->pre_halo < 0) {
      swap_ptr(f[TEA_FIELD_R], c->r_alt);
      c->pend_u = true;
      c->pend_alpha = alpha;
      const int set = c->ra_set;
      c->ra = kRaNone;
      if (ra_possible(c)) {
        // Queue the next w-step (its beta from the device) and r-step
        // before waiting for this r.r.
        TEA_RETURN_IF(sync_p_alt(c));
        const int next = set ^ 1;
        FusedWArgs a{f[TEA_FIELD_P], c->p_alt,  f[TEA_FIELD_R], f[TEA_FIELD_U],
                     c->w_alt,       f[TEA_FIELD_KX], f[TEA_FIELD_KY], alpha,
                     0.0,            1,         1,              c->u_alt,
                     c->ra_dev + 1};
        {
          Timed t(c, K_CG_FUSED_W);
          launch_fused_w1<true>(c, a);
        }
        TEA_HIP_TRY(hipGetLastError());
        TEA_RETURN_IF(ra_finish(c, next, kRaW, 0.0, true));
        TEA_RETURN_IF(ra_launch_r(c, next, c->w_alt, 0.0, true));
        c->ra = kRaWR;
        c->ra_set = next;
        c->ra_w_alpha = alpha;
      }
      double sum = 0.0, beta = 0.0;
      TEA_RETURN_IF(ra_wait(c, set, kRaRStep, &sum, &beta));
      c->ra_beta = beta;
      c->ra_rro = sum;
      c->ra_commits += 1;
      c->ra_fail = 0;
      *rrn = sum;
      return TEA_HIP_OK;
    }
    ra_mismatch(c);
  } else if (c->ra != kRaNone) {
    ra_mismatch(c);
  }
  TEA_FLUSH_HELD(c);
  TEA_RETURN_IF(apply_pending(c));
  if (!c->fused)
    return launch_span1(c, K_CG_CALC_UR,
                        CalcURBody{f[TEA_FIELD_U], f[TEA_FIELD_R],
                                   f[TEA_FIELD_P], f[TEA_FIELD_W], alpha},
                        rrn);
  int rlen = c->nspan;
  if (use_pair_blocks(c, 2)) {
    {
      Timed t(c, K_CG_FUSED_R);
      hipLaunchKernelGGL(fused_r1_kernel, pair_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, f[TEA_FIELD_R], f[TEA_FIELD_R], f[TEA_FIELD_W],
                         alpha, nullptr, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    TEA_RETURN_IF(launch_span_finish(c, 1, &rlen, 0, true));
  } else {
    Timed t(c, K_CG_FUSED_R);
    if (c->pairs)
      hipLaunchKernelGGL(fused_r2_kernel, span_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, f[TEA_FIELD_R], f[TEA_FIELD_W],
                         alpha, c->partials, c->nspan, c->ngroup == 1 ? 1 : 0);
    else
      hipLaunchKernelGGL(fused_r_kernel, span_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, f[TEA_FIELD_R], f[TEA_FIELD_W],
                         alpha, c->partials, c->nspan, c->ngroup == 1 ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  TEA_RETURN_IF(finish_reduction_n(c, 1, rlen, rrn));
  c->pend_u = true;
  c->pend_alpha = alpha;
  c->ra_rro = *rrn;  // the host's next rro (TeaLeaf/drivers/cg_driver.c:113)
  return TEA_HIP_OK;
}

int tea_hip_cg_calc_p(tea_hip_chunk* c, double beta) {
  TEA_ENTER_RA(c);
  // Only deferred: a run-ahead w-step stays in flight for cg_calc_w.
  if (c->held || c->pre_halo >= 0) TEA_FLUSH_HELD(c);
  double** f = c->field;
  if (!c->fused || c->pend_p || c->pend_halo) TEA_RETURN_IF(apply_pending(c));
  if (!c->fused) {
    c->p_alt_synced = false;
    return launch_each(c, K_CG_CALC_P,
                       CalcPBody{f[TEA_FIELD_P], f[TEA_FIELD_R], beta, c->geo});
  }
  c->pend_p = true;
  c->pend_beta = beta;
  return TEA_HIP_OK;
}

int tea_hip_copy_u(tea_hip_chunk* c) {
  TEA_ENTER(c);
  TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  return launch_each(c, K_COPY_U,
                     CopyUBody{f[TEA_FIELD_U], f[TEA_FIELD_U0], c->geo});
}

int tea_hip_calculate_residual(tea_hip_chunk* c) {
  TEA_ENTER(c);
  TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  return launch_each(c, K_CALC_RESIDUAL,
                     ResidualBody{f[TEA_FIELD_U], f[TEA_FIELD_U0],
                                  f[TEA_FIELD_R], f[TEA_FIELD_KX],
                                  f[TEA_FIELD_KY], c->geo});
}

int tea_hip_calculate_2norm(tea_hip_chunk* c, const double* buffer,
                            double* norm) {
  TEA_ENTER_HELD(c);
  TEA_CHECK_ARG(norm != nullptr, "null norm");
  int fid = -1;
  for (int f = 0; f < TEA_NUM_FIELDS; ++f)
    if (buffer == c->field[f] && c->field_len[f] >= (size_t)c->geo.n) fid = f;
  TEA_CHECK_ARG(fid >= 0, "buffer is not an x*y field of this chunk");
  // Deferred work can move the field to its second buffer: name it by field.
  TEA_RETURN_IF(apply_pending(c));
  const double* b = c->field[fid];
  if (!use_pair_blocks(c, 1))
    return launch_span1(c, K_CALC_2NORM, NormBody{b}, norm);
  {
    Timed t(c, K_CALC_2NORM);
    hipLaunchKernelGGL(norm_pair_kernel, pair_grid(c), dim3(kBlock), 0,
                       c->stream, c->geo, b, c->sub, c->nblk);
  }
  TEA_HIP_TRY(hipGetLastError());
  int len = c->nspan;
  TEA_RETURN_IF(launch_span_finish(c, 1, &len, 0, true));
  return finish_reduction_n(c, 1, len, norm);
}

int tea_hip_finalise(tea_hip_chunk* c) {
  TEA_ENTER(c);
  TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  return launch_each(c, K_FINALISE,
                     FinaliseBody{f[TEA_FIELD_U], f[TEA_FIELD_DENSITY],
                                  f[TEA_FIELD_ENERGY], c->geo});
}

}  // extern "C"
