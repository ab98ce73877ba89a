// libteahip.so: mixed-precision CG (no counterpart in the reference, whose
// CG is FP64 throughout: TeaLeaf/c_kernels/sycl/cg.cpp).
//
// Storage: p, r, w and the face coefficients kx, ky in FP32; the solution u,
// the right-hand side u0 and the FP64 kx, ky of the plain path stay FP64.
// Arithmetic:
//   - w = A p in FP32, in the operand order of TeaLeaf/shared.h:59-63;
//   - p = beta p + r and r = r - alpha w in FP64, rounded once to FP32;
//   - u += alpha p in FP64;
//   - every dot product takes exact FP64 products of the FP32 values and
//     sums them with the reference's reduction tree (as the FP64 path);
//   - a true residual r = u0 - A u in FP64 (FP64 coefficients) replaces the
//     recursive one every refresh interval and at convergence.
// Bytes per cell per iteration: kernel 1 reads p, r (4+4), u (8), kx, ky
// (4+4) and writes p, w (4+4), u (8) = 40; kernel 2 reads r, w and writes r
// = 12; 52 in total against 88 for the FP64 iteration.
//
// Requires all four faces to be external (one chunk per GPU); the halo ring
// of u and p is kept current by the kernels themselves, as in the fused FP64
// iteration.

#include <hip/hip_runtime.h>

#include "tea_hip_impl.h"

namespace tea {

// p_new = beta p_old + r (r is the residual for CG, the preconditioned
// residual z for flexible CG); with DOT2 also reduces p.rdot.
template <int kBatch, bool DOT2>
__global__ __launch_bounds__(kBlock) void mp_w_kernel(Geo g, MpArgs a,
                                                      double* partials,
                                                      int nspan,
                                                      int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  const int x = g.x;
  const int xl = g.h, xr = g.x - g.h - 1, yb = g.h, yt = g.y - g.h - 1;
  const float* __restrict__ P = a.p_old;
  const float* __restrict__ R = a.r;
  double prod[DOT2 ? 2 : 1][kItems];
#pragma unroll
  for (int b0 = 0; b0 < kItems; b0 += kBatch) {
    int ci[kBatch], ck[kBatch], cj[kBatch];
    bool in[kBatch];
    float po[kBatch], pl[kBatch], pr[kBatch], pb[kBatch], pt[kBatch];
    float rc[kBatch], rl[kBatch], rr[kBatch], rb[kBatch], rt[kBatch];
    float kx0[kBatch], kx1[kBatch], ky0[kBatch], ky1[kBatch];
    double uc[kBatch];
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
      const int i = in[j] ? ci[j] : x + 1;
      po[j] = P[i];
      pl[j] = P[i - 1];
      pr[j] = P[i + 1];
      pb[j] = P[i - x];
      pt[j] = P[i + x];
      rc[j] = R[i];
      rl[j] = R[i - 1];
      rr[j] = R[i + 1];
      rb[j] = R[i - x];
      rt[j] = R[i + x];
      kx0[j] = a.kx[i];
      kx1[j] = a.kx[i + 1];
      ky0[j] = a.ky[i];
      ky1[j] = a.ky[i + x];
      uc[j] = a.apply_u ? a.u[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int i = ci[j], kk = ck[j], jj = cj[j];
      float pc, l, rgt, bot, top;
      if (a.apply_p) {
        pc = (float)(a.beta * (double)po[j] + (double)rc[j]);
        l = kk == xl ? pc : (float)(a.beta * (double)pl[j] + (double)rl[j]);
        rgt = kk == xr ? pc : (float)(a.beta * (double)pr[j] + (double)rr[j]);
        bot = jj == yb ? pc : (float)(a.beta * (double)pb[j] + (double)rb[j]);
        top = jj == yt ? pc : (float)(a.beta * (double)pt[j] + (double)rt[j]);
      } else {
        pc = po[j];
        l = pl[j];
        rgt = pr[j];
        bot = pb[j];
        top = pt[j];
      }
      const float diag = (1.0f + (kx1[j] + kx0[j])) + (ky1[j] + ky0[j]);
      const float s = (diag * pc - (kx1[j] * rgt + kx0[j] * l)) -
                      (ky1[j] * top + ky0[j] * bot);
      prod[0][b0 + j] = in[j] ? (double)s * (double)pc : 0.0;
      if (DOT2) prod[DOT2 ? 1 : 0][b0 + j] = in[j] ? (double)pc * (double)a.rdot[i] : 0.0;
      if (!in[j]) continue;
      a.w[i] = s;
      const bool L = kk == xl, Rr = kk == xr, B = jj == yb, T = jj == yt;
      if (a.apply_u) {
        const double un = uc[j] + a.alpha * (double)po[j];
        a.u[i] = un;
        if (L) a.u[i - 1] = un;
        if (Rr) a.u[i + 1] = un;
        if (B) a.u[i - x] = un;
        if (T) a.u[i + x] = un;
        if (L && B) a.u[i - 1 - x] = un;
        if (Rr && B) a.u[i + 1 - x] = un;
        if (L && T) a.u[i - 1 + x] = un;
        if (Rr && T) a.u[i + 1 + x] = un;
      }
      if (a.apply_p) {
        float* __restrict__ Q = a.p_new;
        Q[i] = pc;
        if (L) Q[i - 1] = pc;
        if (Rr) Q[i + 1] = pc;
        if (B) Q[i - x] = pc;
        if (T) Q[i + x] = pc;
        if (L && B) Q[i - 1 - x] = pc;
        if (Rr && B) Q[i + 1 - x] = pc;
        if (L && T) Q[i - 1 + x] = pc;
        if (Rr && T) Q[i + 1 + x] = pc;
      }
    }
  }
  span_reduce<DOT2 ? 2 : 1>(prod, partials, nspan, span, single_group != 0);
}

template __global__ void mp_w_kernel<4, true>(Geo, MpArgs, double*, int, int);

// r -= alpha w and r.r (and, with RW, r.w of the updated r).
template <bool RW>
__global__ __launch_bounds__(kBlock) void mp_r_kernel(
    Geo g, float* __restrict__ r, const float* __restrict__ w, double alpha,
    double* partials, int nspan, int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  float rv[kItems], wv[kItems];
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
  double prod[RW ? 2 : 1][kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const float rn = (float)((double)rv[k] - alpha * (double)wv[k]);
    prod[0][k] = in[k] ? (double)rn * (double)rn : 0.0;
    if (RW) prod[RW ? 1 : 0][k] = in[k] ? (double)rn * (double)wv[k] : 0.0;
    if (in[k]) r[base + k * kBlock + threadIdx.x] = rn;
  }
  span_reduce<RW ? 2 : 1>(prod, partials, nspan, span, single_group != 0);
}

// Two cells per lane (x even): the arithmetic of mp_w_kernel cell by cell,
// with 8-byte loads of the FP32 operands and 16-byte loads of u.  One pair's
// indices, loads and arithmetic, shared by the span-per-block and
// pair-per-thread kernels.
struct MpPair {
  int e0, kk, jj;
  bool in0, in1, any;
  float2 P0, Pt, Pb, R0, Rt, Rb, KX, KY, KYt, RD;
  double2 U;
  // Wave-edge neighbours: lane 0 holds p and r at e0 - 1, lane 63 p, r and
  // kx at e0 + 2 (the other lanes take theirs from adjacent lanes).
  float PN, RN, KXN;
};

__device__ inline void mp_pair_index(const Geo& g, int e, MpPair& s) {
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  s.e0 = e;
  s.jj = row;
  s.kk = e - row * g.x;
  s.in0 = e < g.n && interior(g, s.kk, row);
  s.in1 = e < g.n && interior(g, s.kk + 1, row);
  s.any = s.in0 || s.in1;
}

// A pair's own p, r and kx are also its neighbours' left/right operands (by
// shuffle), so they are read whenever the pair is in range; the rows above
// and below only for interior pairs.  x + 2 is an even index whose +-x rows
// are in range.  The wave-edge neighbours are issued with the rest, at
// ir - 1 and ir + 2, which are in range too.
template <bool DOT2>
__device__ inline void mp_pair_load(const Geo& g, const MpArgs& a, MpPair& s) {
  const int x = g.x;
  const float* __restrict__ P = a.p_old;
  const float* __restrict__ R = a.r;
  const int ic = s.e0 < g.n ? s.e0 : x + 2;
  const int ir = s.any ? s.e0 : x + 2;
  // Every load is issued unconditionally, r's and u's also where the step
  // does not use them (the first iteration only): loads under a branch were
  // each waited for inside it, before the next ones were issued.
  s.P0 = ld2(P, ic);
  s.Pt = ld2(P, ir + x);
  s.Pb = ld2(P, ir - x);
  s.R0 = ld2(R, ic);
  s.Rt = ld2(R, ir + x);
  s.Rb = ld2(R, ir - x);
  s.KX = ld2(a.kx, ic);
  s.KY = ld2(a.ky, ir);
  s.KYt = ld2(a.ky, ir + x);
  s.U = ld2_nt(a.u, ir);
  if (DOT2) s.RD = ld2(a.rdot, ir);
  // Lanes 0 and 63 share each load instruction, at their own addresses
  // (lane 0's kx value is not used).
  const int lane = threadIdx.x & (kWave - 1);
  if (lane == 0 || lane == kWave - 1) {
    const int ni = lane == 0 ? ir - 1 : ir + 2;
    s.PN = P[ni];
    s.RN = R[ni];
    s.KXN = a.kx[ni];
  }
}

// Computes and stores the pair (w, u, p_new and their halo ring) and returns
// the in-lane pair sums of p_new . w and, with DOT2, p_new . rdot.
template <bool DOT2>
__device__ inline void mp_pair_compute(const Geo& g, const MpArgs& a,
                                       const MpPair& s, int lane,
                                       double (&ps)[DOT2 ? 2 : 1]) {
  const int x = g.x;
  const int xl = g.h, xr = g.x - g.h - 1, yb = g.h, yt = g.y - g.h - 1;
  const double beta = a.beta;
  const int i = s.e0;
  float pl = __shfl_up(s.P0.y, 1), pr = __shfl_down(s.P0.x, 1);
  float rl = 0.0f, rr = 0.0f;
  if (a.apply_p) {
    rl = __shfl_up(s.R0.y, 1);
    rr = __shfl_down(s.R0.x, 1);
  }
  float kxr = __shfl_down(s.KX.x, 1);
  if (lane == 0) {
    pl = s.PN;
    if (a.apply_p) rl = s.RN;
  }
  if (lane == kWave - 1) {
    pr = s.PN;
    if (a.apply_p) rr = s.RN;
    kxr = s.KXN;
  }
  const int ka = s.kk, kb = s.kk + 1, row = s.jj;
  float ca, cb, la, ra, lb, rb2, ba, ta, bb, tb;
  if (a.apply_p) {
    ca = (float)(beta * (double)s.P0.x + (double)s.R0.x);
    cb = (float)(beta * (double)s.P0.y + (double)s.R0.y);
    la = ka == xl ? ca : (float)(beta * (double)pl + (double)rl);
    ra = ka == xr ? ca : cb;
    lb = kb == xl ? cb : ca;
    rb2 = kb == xr ? cb : (float)(beta * (double)pr + (double)rr);
    ba = row == yb ? ca : (float)(beta * (double)s.Pb.x + (double)s.Rb.x);
    ta = row == yt ? ca : (float)(beta * (double)s.Pt.x + (double)s.Rt.x);
    bb = row == yb ? cb : (float)(beta * (double)s.Pb.y + (double)s.Rb.y);
    tb = row == yt ? cb : (float)(beta * (double)s.Pt.y + (double)s.Rt.y);
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
  const float kxa0 = s.KX.x, kxa1 = s.KX.y, kxb0 = s.KX.y, kxb1 = kxr;
  const float diag_a = (1.0f + (kxa1 + kxa0)) + (s.KYt.x + s.KY.x);
  const float diag_b = (1.0f + (kxb1 + kxb0)) + (s.KYt.y + s.KY.y);
  const float sa = (diag_a * ca - (kxa1 * ra + kxa0 * la)) -
                   (s.KYt.x * ta + s.KY.x * ba);
  const float sb = (diag_b * cb - (kxb1 * rb2 + kxb0 * lb)) -
                   (s.KYt.y * tb + s.KY.y * bb);
  ps[0] = (s.in0 ? (double)sa * (double)ca : 0.0) +
          (s.in1 ? (double)sb * (double)cb : 0.0);
  if (DOT2)
    ps[DOT2 ? 1 : 0] = (s.in0 ? (double)ca * (double)s.RD.x : 0.0) +
                       (s.in1 ? (double)cb * (double)s.RD.y : 0.0);
  if (!s.any) return;
  st2(a.w, i, s.in0, s.in1, sa, sb);
  double ua = 0.0, ub = 0.0;
  if (a.apply_u) {
    ua = s.U.x + a.alpha * (double)s.P0.x;
    ub = s.U.y + a.alpha * (double)s.P0.y;
    st2_nt(a.u, i, s.in0, s.in1, ua, ub);
  }
  if (a.apply_p) st2_nt(a.p_new, i, s.in0, s.in1, ca, cb);
  // Halo ring of boundary-adjacent cells (rare; per cell).
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const bool inc = c == 0 ? s.in0 : s.in1;
    const int ic = i + c, kc = ka + c;
    const bool L = kc == xl, Rg = kc == xr, B = row == yb, T = row == yt;
    if (!inc || !(L | Rg | B | T)) continue;
    const float pv = c == 0 ? ca : cb;
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
      if (a.apply_u) a.u[ring[q]] = uv;
      if (a.apply_p) a.p_new[ring[q]] = pv;
    }
  }
}

// One span per block, 8 pairs per thread in batches of SB.
template <bool DOT2, int SB>
__global__ __launch_bounds__(kBlock) void mp_w2_kernel(Geo g, MpArgs a,
                                                       double* partials,
                                                       int nspan,
                                                       int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  const int lane = threadIdx.x & (kWave - 1);
  double psum[DOT2 ? 2 : 1][kSlots];
#pragma unroll
  for (int s0 = 0; s0 < kSlots; s0 += SB) {
    MpPair s[SB];
#pragma unroll
    for (int j = 0; j < SB; ++j)
      mp_pair_index(g, base + (s0 + j) * (2 * kBlock) + 2 * threadIdx.x, s[j]);
#pragma unroll
    for (int j = 0; j < SB; ++j) mp_pair_load<DOT2>(g, a, s[j]);
#pragma unroll
    for (int j = 0; j < SB; ++j) {
      double ps[DOT2 ? 2 : 1];
      mp_pair_compute<DOT2>(g, a, s[j], lane, ps);
      psum[0][s0 + j] = ps[0];
      if (DOT2) psum[DOT2 ? 1 : 0][s0 + j] = ps[DOT2 ? 1 : 0];
    }
  }
  span_reduce_pairs_n<DOT2 ? 2 : 1>(psum, partials, nspan, span, single_group != 0);
}

template __global__ void mp_w2_kernel<true, 4>(Geo, MpArgs, double*, int, int);

// One pair per thread: the block's 8-leaf subtrees go to sub (see
// block_subtree_pairs) and span_finish_kernel completes the spans.
template <bool DOT2>
__global__ __launch_bounds__(kBlock) void mp_w1_kernel(Geo g, MpArgs a,
                                                       double* sub, int nblk) {
  const int blk = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  MpPair s;
  mp_pair_index(g, blk * kPairBlock + 2 * threadIdx.x, s);
  mp_pair_load<DOT2>(g, a, s);
  double ps[DOT2 ? 2 : 1];
  mp_pair_compute<DOT2>(g, a, s, lane, ps);
  wave_subtree_pairs<DOT2 ? 2 : 1>(ps, sub, nblk, blk);
}

template __global__ void mp_w1_kernel<true>(Geo, MpArgs, double*, int);

// The reflection of an interior cell's new value v onto the depth-1 halo
// ring where the cell borders it (kc, rc its column and row).
template <class T>
__device__ inline void ring_store(T* f, int i, int kc, int rc, int x, int xl, int xr, int yb,
                                  int yt, T v) {
  const bool L = kc == xl, Rg = kc == xr, B = rc == yb, Tp = rc == yt;
  if (!(L | Rg | B | Tp)) return;
  if (L) f[i - 1] = v;
  if (Rg) f[i + 1] = v;
  if (B) f[i - x] = v;
  if (Tp) f[i + x] = v;
  if (L && B) f[i - 1 - x] = v;
  if (Rg && B) f[i + 1 - x] = v;
  if (L && Tp) f[i - 1 + x] = v;
  if (Rg && Tp) f[i + 1 + x] = v;
}

// Four cells per lane (x a multiple of 4): the arithmetic of mp_w_kernel cell
// by cell, with 16-byte loads of the FP32 operands and of u.  The
// neighbours across lanes move on the DPP wave shifts; lane 0 loads p and r
// at e0 - 1, lane 63 p, r and kx at e0 + 4.  The in-lane dot sums are
// ((v0 + v1) + (v2 + v3)), tree steps 1 and 2, and the waves store their
// sums for span_finish_kernel (wave_subtree_quads).  UM selects the u
// update: none, u += alpha p_old, or u = ((u + alpha_h p_h) + alpha_h2
// p_h2) + alpha p_old (two held updates and this step's, each rounded as
// when made one by one).  A workgroup of KS x 128 threads takes KS blocks
// one row apart (block c + S t, t = KS j .. KS j + KS - 1, S = x / 512;
// blocks past the grid return at once), which share their rows above and
// below in the CU's L1.
template <bool DOT2, int UM, int KS>
__global__ __launch_bounds__(kQuadThreads * KS) void mp_w4_kernel(Geo g, MpArgs a, double* sub,
                                                                  int nblk, int S) {
  int tid;
  const int blk = stacked_block<KS, kQuadThreads>(S, &tid);
  if (blk >= nblk) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int x = g.x;
  const int xl = g.h, xr = g.x - g.h - 1, yb = g.h, yt = g.y - g.h - 1;
  const float* __restrict__ P = a.p_old;
  const float* __restrict__ R = a.r;
  const int e = blk * kQuadBlock + 4 * tid;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * x;
  bool in[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) in[c] = e < g.n && interior(g, kk + c, row);
  const bool any = in[0] || in[1] || in[2] || in[3];
  // A quad's own p, r and kx are its neighbours' left/right operands too, so
  // they are read whenever it is in range; x + 4 is a safe aligned index
  // whose rows above and below are in range.
  const int ic = e < g.n ? e : x + 4;
  const int ir = any ? e : x + 4;
  // Every load is issued unconditionally, r's also where the step does not
  // use it (the first iteration only): loads under a branch were each
  // waited for inside it, before the next ones were issued.  u (and the
  // held p) are loaded as UM, a template argument, says.
  const float4 P0 = ld4(P, ic), Pt = ld4(P, ir + x), Pb = ld4(P, ir - x);
  const float4 R0 = ld4(R, ic), Rt = ld4(R, ir + x), Rb = ld4(R, ir - x);
  const float4 KX = ld4(a.kx, ic), KY = ld4(a.ky, ir), KYt = ld4(a.ky, ir + x);
  // u moves in a lane layout of its own: lane l of a wave updates cells
  // wb + 2l, +1 (A) and wb + 128 + 2l, +1 (B), wb the wave's first cell, so
  // each 16-byte load and store covers 1 KB of consecutive cells (32 bytes
  // per lane in the quad layout took two instructions over every other
  // 16 bytes).  The p_old values come from the quad lanes by permute.
  const int ua = blk * kQuadBlock + (tid / kWave) * (4 * kWave) + 2 * lane;
  const int ub = ua + 2 * kWave;
  const int rA = ua < g.n ? (int)g.dx.div((uint32_t)ua) : 0;
  const int kA = ua - rA * x;
  const int rB = ub < g.n ? (int)g.dx.div((uint32_t)ub) : 0;
  const int kB = ub - rB * x;
  const bool inA0 = ua < g.n && interior(g, kA, rA);
  const bool inA1 = ua < g.n && interior(g, kA + 1, rA);
  const bool inB0 = ub < g.n && interior(g, kB, rB);
  const bool inB1 = ub < g.n && interior(g, kB + 1, rB);
  double2 UA = {}, UB = {};
  float2 HA = {}, HB = {}, GA = {}, GB = {};
  if (UM > 0) {
    UA = ld2_nt(a.u, inA0 || inA1 ? ua : x + 2);
    UB = ld2_nt(a.u, inB0 || inB1 ? ub : x + 2);
  }
  if (UM == 3) {
    // p_h is the p_new buffer: each of its cells is read here by the wave
    // that overwrites it below, and by no other.
    HA = ld2(a.p_h, inA0 || inA1 ? ua : x + 2);
    HB = ld2(a.p_h, inB0 || inB1 ? ub : x + 2);
    GA = ld2(a.p_h2, inA0 || inA1 ? ua : x + 2);
    GB = ld2(a.p_h2, inB0 || inB1 ? ub : x + 2);
  }
  float4 RD = {};
  if (DOT2) RD = ld4(a.rdot, ir);
  float PN = 0.0f, RN = 0.0f, KXN = 0.0f;
  if (lane == 0 || lane == kWave - 1) {
    const int ni = lane == 0 ? ir - 1 : ir + 4;
    PN = P[ni];
    RN = R[ni];
    KXN = a.kx[ni];
  }
  float pl = from_prev_lane(P0.w), pr = from_next_lane(P0.x);
  float rl = 0.0f, rr = 0.0f;
  if (a.apply_p) {
    rl = from_prev_lane(R0.w);
    rr = from_next_lane(R0.x);
  }
  float kxr = from_next_lane(KX.x);
  if (lane == 0) {
    pl = PN;
    rl = RN;
  }
  if (lane == kWave - 1) {
    pr = PN;
    rr = RN;
    kxr = KXN;
  }
  const double beta = a.beta;
  // p_new of the quad's cells, of its left and right neighbours, and of the
  // cells below and above (a halo neighbour reflects the cell itself).
  // Vector values, not arrays: an array indexed by the loop counter (pc[c - 1])
  // was placed in LDS by the compiler before the loops were unrolled.
  fvec4 pc, pb, pt;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (a.apply_p) {
      pc[c] = (float)(beta * (double)q_at(P0, c) + (double)q_at(R0, c));
      pb[c] = row == yb ? pc[c] : (float)(beta * (double)q_at(Pb, c) + (double)q_at(Rb, c));
      pt[c] = row == yt ? pc[c] : (float)(beta * (double)q_at(Pt, c) + (double)q_at(Rt, c));
    } else {
      pc[c] = q_at(P0, c);
      pb[c] = q_at(Pb, c);
      pt[c] = q_at(Pt, c);
    }
  }
  // Left and right operands: the neighbouring cells' p_new, a reflection of
  // the cell itself next to the halo (p_new's halo is not written yet), and
  // across lanes the neighbour quads' values.
  fvec4 lft, rgt;
  if (a.apply_p) {
    lft[0] = kk == xl ? pc[0] : (float)(beta * (double)pl + (double)rl);
    rgt[3] = kk + 3 == xr ? pc[3] : (float)(beta * (double)pr + (double)rr);
  } else {
    lft[0] = pl;
    rgt[3] = pr;
  }
  const bool refl = a.apply_p != 0;
  lft[1] = refl && kk + 1 == xl ? pc[1] : pc[0];
  lft[2] = refl && kk + 2 == xl ? pc[2] : pc[1];
  lft[3] = refl && kk + 3 == xl ? pc[3] : pc[2];
  rgt[0] = refl && kk == xr ? pc[0] : pc[1];
  rgt[1] = refl && kk + 1 == xr ? pc[1] : pc[2];
  rgt[2] = refl && kk + 2 == xr ? pc[2] : pc[3];
  const fvec4 kx0 = {KX.x, KX.y, KX.z, KX.w}, kx1 = {KX.y, KX.z, KX.w, kxr};
  const fvec4 ky0 = {KY.x, KY.y, KY.z, KY.w}, ky1 = {KYt.x, KYt.y, KYt.z, KYt.w};
  fvec4 s;
  double prod[DOT2 ? 2 : 1][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float diag = (1.0f + (kx1[c] + kx0[c])) + (ky1[c] + ky0[c]);
    s[c] = (diag * pc[c] - (kx1[c] * rgt[c] + kx0[c] * lft[c])) -
           (ky1[c] * pt[c] + ky0[c] * pb[c]);
    prod[0][c] = in[c] ? (double)s[c] * (double)pc[c] : 0.0;
    if (DOT2) prod[DOT2 ? 1 : 0][c] = in[c] ? (double)pc[c] * (double)q_at(RD, c) : 0.0;
  }
  double v[DOT2 ? 2 : 1];
#pragma unroll
  for (int q = 0; q < (DOT2 ? 2 : 1); ++q)
    v[q] = (prod[q][0] + prod[q][1]) + (prod[q][2] + prod[q][3]);
  if (UM > 0) {
    // Cells 2l, 2l + 1 of each half-wave are components (x, y) or (z, w) of
    // quad lane l / 2 (A) or 32 + l / 2 (B).
    const int sa = lane >> 1, sb = kWave / 2 + (lane >> 1);
    const bool hi = (lane & 1) != 0;
    const float ax = __shfl(P0.x, sa), ay = __shfl(P0.y, sa), az = __shfl(P0.z, sa),
                aw = __shfl(P0.w, sa);
    const float bx = __shfl(P0.x, sb), by = __shfl(P0.y, sb), bz = __shfl(P0.z, sb),
                bw = __shfl(P0.w, sb);
    double u0[4] = {UA.x, UA.y, UB.x, UB.y};
    if (UM == 3) {
      u0[0] = u0[0] + a.alpha_h * (double)HA.x;
      u0[1] = u0[1] + a.alpha_h * (double)HA.y;
      u0[2] = u0[2] + a.alpha_h * (double)HB.x;
      u0[3] = u0[3] + a.alpha_h * (double)HB.y;
      u0[0] = u0[0] + a.alpha_h2 * (double)GA.x;
      u0[1] = u0[1] + a.alpha_h2 * (double)GA.y;
      u0[2] = u0[2] + a.alpha_h2 * (double)GB.x;
      u0[3] = u0[3] + a.alpha_h2 * (double)GB.y;
    }
    const double un[4] = {u0[0] + a.alpha * (double)(hi ? az : ax),
                          u0[1] + a.alpha * (double)(hi ? aw : ay),
                          u0[2] + a.alpha * (double)(hi ? bz : bx),
                          u0[3] + a.alpha * (double)(hi ? bw : by)};
    if (inA0 || inA1) st2_nt(a.u, ua, inA0, inA1, un[0], un[1]);
    if (inB0 || inB1) st2_nt(a.u, ub, inB0, inB1, un[2], un[3]);
    // Halo ring of boundary-adjacent cells (rare; per cell).
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bool inc = c == 0 ? inA0 : c == 1 ? inA1 : c == 2 ? inB0 : inB1;
      const int i = (c < 2 ? ua : ub) + (c & 1);
      const int kc = (c < 2 ? kA : kB) + (c & 1), rc = c < 2 ? rA : rB;
      if (inc) ring_store(a.u, i, kc, rc, x, xl, xr, yb, yt, un[c]);
    }
  }
  if (any) {
    st4(a.w, e, in, s);
    if (a.apply_p) {
      st4_nt(a.p_new, e, in, pc);
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (in[c]) ring_store(a.p_new, e + c, kk + c, row, x, xl, xr, yb, yt, pc[c]);
    }
  }
  wave_subtree_quads<DOT2 ? 2 : 1>(v, sub, nblk, blk, tid / kWave);
}


// r -= alpha w and r.r (with RW also r.w), two cells per lane.
template <bool RW>
__global__ __launch_bounds__(kBlock) void mp_r2_kernel(
    Geo g, float* __restrict__ r, const float* __restrict__ w, double alpha,
    double* partials, int nspan, int single_group) {
  const int span = xcd_span(blockIdx.x, gridDim.x);
  const int base = span * kSpan;
  float2 rv[kSlots], wv[kSlots];
  bool in0[kSlots], in1[kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const int e = base + k * (2 * kBlock) + 2 * threadIdx.x;
    const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
    const int kk = e - row * g.x;
    in0[k] = e < g.n && interior(g, kk, row);
    in1[k] = e < g.n && interior(g, kk + 1, row);
    const int i = (in0[k] || in1[k]) ? e : 0;
    rv[k] = ld2(r, i);
    wv[k] = ld2(w, i);
  }
  double psum[RW ? 2 : 1][kSlots];
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const float ra = (float)((double)rv[k].x - alpha * (double)wv[k].x);
    const float rb = (float)((double)rv[k].y - alpha * (double)wv[k].y);
    psum[0][k] = (in0[k] ? (double)ra * (double)ra : 0.0) +
                 (in1[k] ? (double)rb * (double)rb : 0.0);
    if (RW)
      psum[RW ? 1 : 0][k] = (in0[k] ? (double)ra * (double)wv[k].x : 0.0) +
                            (in1[k] ? (double)rb * (double)wv[k].y : 0.0);
    st2(r, base + k * (2 * kBlock) + 2 * threadIdx.x, in0[k], in1[k], ra, rb);
  }
  span_reduce_pairs_n<RW ? 2 : 1>(psum, partials, nspan, span, single_group != 0);
}

// mp_r2_kernel's arithmetic with one pair per thread.
template <bool RW>
__global__ __launch_bounds__(kBlock) void mp_r1_kernel(
    Geo g, float* __restrict__ r, const float* __restrict__ w, double alpha,
    double* sub, int nblk) {
  const int blk = blockIdx.x;
  const int e = blk * kPairBlock + 2 * threadIdx.x;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * g.x;
  const bool in0 = e < g.n && interior(g, kk, row);
  const bool in1 = e < g.n && interior(g, kk + 1, row);
  const int i = (in0 || in1) ? e : 0;
  const float2 rv = ld2(r, i), wv = ld2(w, i);
  const float ra = (float)((double)rv.x - alpha * (double)wv.x);
  const float rb = (float)((double)rv.y - alpha * (double)wv.y);
  double v[RW ? 2 : 1];
  v[0] = (in0 ? (double)ra * (double)ra : 0.0) + (in1 ? (double)rb * (double)rb : 0.0);
  if (RW)
    v[RW ? 1 : 0] = (in0 ? (double)ra * (double)wv.x : 0.0) +
                    (in1 ? (double)rb * (double)wv.y : 0.0);
  st2(r, e, in0, in1, ra, rb);
  wave_subtree_pairs<RW ? 2 : 1>(v, sub, nblk, blk);
}

// The same four cells per lane (x a multiple of 4).
template <bool RW>
__global__ __launch_bounds__(kQuadThreads) void mp_r4_kernel(
    Geo g, float* __restrict__ r, const float* __restrict__ w, double alpha,
    double* sub, int nblk) {
  const int blk = blockIdx.x;
  const int e = blk * kQuadBlock + 4 * threadIdx.x;
  const int row = e < g.n ? (int)g.dx.div((uint32_t)e) : 0;
  const int kk = e - row * g.x;
  bool in[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) in[c] = e < g.n && interior(g, kk + c, row);
  const bool any = in[0] || in[1] || in[2] || in[3];
  const int i = any ? e : 0;
  const float4 rv = ld4(r, i), wv = ld4(w, i);
  float rn[4];
  double prod[RW ? 2 : 1][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    rn[c] = (float)((double)q_at(rv, c) - alpha * (double)q_at(wv, c));
    prod[0][c] = in[c] ? (double)rn[c] * (double)rn[c] : 0.0;
    if (RW) prod[RW ? 1 : 0][c] = in[c] ? (double)rn[c] * (double)q_at(wv, c) : 0.0;
  }
  double v[RW ? 2 : 1];
#pragma unroll
  for (int q = 0; q < (RW ? 2 : 1); ++q)
    v[q] = (prod[q][0] + prod[q][1]) + (prod[q][2] + prod[q][3]);
  if (any) st4(r, e, in, rn);
  wave_subtree_quads<RW ? 2 : 1>(v, sub, nblk, blk, threadIdx.x / kWave);
}

// MpTrueResidualBody's arithmetic (below) with one cell pair per thread (x
// even): r = u0 - A u in FP64 with the FP64 coefficients, r rounded to FP32
// into r32, and the three sums (r32.r32, r.r, r32.w) as wave sums for
// span_finish_kernel.  u's left and right neighbours come from the adjacent
// lanes by DPP; lane 0 loads u at e - 1, lane 63 u and kx at e + 2.
__global__ __launch_bounds__(kBlock) void mp_residual1_kernel(
    Geo g, const double* __restrict__ u, const double* __restrict__ u0,
    const double* __restrict__ kx, const double* __restrict__ ky, float* __restrict__ r32,
    const float* __restrict__ w, double* sub, int nblk) {
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
  const double2 U = ld2(u, ic), Ut = ld2(u, ir + x), Ub = ld2(u, ir - x);
  const double2 KX = ld2(kx, ic), KY = ld2(ky, ir), KYt = ld2(ky, ir + x);
  const double2 B = ld2_nt(u0, ir);
  const float2 W = ld2(w, ir);
  double UN = 0.0, KXN = 0.0;
  if (lane == 0 || lane == kWave - 1) {
    const int ni = lane == 0 ? ir - 1 : ir + 2;
    UN = u[ni];
    KXN = kx[ni];
  }
  double ul = from_prev_lane(U.y), ur = from_next_lane(U.x), kxr = from_next_lane(KX.x);
  if (lane == 0) ul = UN;
  if (lane == kWave - 1) {
    ur = UN;
    kxr = KXN;
  }
  const double diag_a = (1.0 + (KX.y + KX.x)) + (KYt.x + KY.x);
  const double diag_b = (1.0 + (kxr + KX.y)) + (KYt.y + KY.y);
  const double aua = (diag_a * U.x - (KX.y * U.y + KX.x * ul)) - (KYt.x * Ut.x + KY.x * Ub.x);
  const double aub = (diag_b * U.y - (kxr * ur + KX.y * U.x)) - (KYt.y * Ut.y + KY.y * Ub.y);
  const double ra = B.x - aua, rb = B.y - aub;
  const float fa = (float)ra, fb = (float)rb;
  double v[3] = {(in0 ? (double)fa * (double)fa : 0.0) + (in1 ? (double)fb * (double)fb : 0.0),
                 (in0 ? ra * ra : 0.0) + (in1 ? rb * rb : 0.0),
                 (in0 ? (double)fa * (double)W.x : 0.0) + (in1 ? (double)fb * (double)W.y : 0.0)};
  if (any) st2(r32, e, in0, in1, fa, fb);
  wave_subtree_pairs<3>(v, sub, nblk, blk);
}

namespace {

// FP32 mirrors after the FP64 initialisation: coefficients and p on every
// cell (p's halo is already reflected), r on every cell; reduces
// (r32.r32, u0.u0) over the interior.
struct MpConvertBody {
  const double* r;
  const double* u0;
  __device__ void operator()(int i, int, int, double (&out)[2]) const {
    const float rv = (float)r[i];
    out[0] = (double)rv * (double)rv;
    out[1] = u0[i] * u0[i];
  }
};

__global__ __launch_bounds__(kBlock) void mp_convert_all_kernel(
    int n, const double* __restrict__ kx, const double* __restrict__ ky,
    const double* __restrict__ p, const double* __restrict__ r,
    float* __restrict__ kx32, float* __restrict__ ky32,
    float* __restrict__ p32a, float* __restrict__ p32b, float* __restrict__ p32c,
    float* __restrict__ r32) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  kx32[i] = (float)kx[i];
  ky32[i] = (float)ky[i];
  const float pv = (float)p[i];
  p32a[i] = pv;
  p32b[i] = pv;
  p32c[i] = pv;
  r32[i] = (float)r[i];
}

// u += alpha p (FP64 u, FP32 p) on the interior and its reflection on the
// depth-1 halo ring.
struct MpApplyUBody {
  double* u;
  const float* p;
  double alpha;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const int x = g.x;
    const double un = u[i] + alpha * (double)p[i];
    u[i] = un;
    const bool L = kk == g.h, R = kk == x - g.h - 1, B = jj == g.h,
               T = jj == g.y - g.h - 1;
    if (L) u[i - 1] = un;
    if (R) u[i + 1] = un;
    if (B) u[i - x] = un;
    if (T) u[i + x] = un;
    if (L && B) u[i - 1 - x] = un;
    if (R && B) u[i + 1 - x] = un;
    if (L && T) u[i - 1 + x] = un;
    if (R && T) u[i + 1 + x] = un;
  }
};

// r = u0 - A u in FP64 with the FP64 coefficients; stores r rounded to FP32
// and reduces (r32.r32, r.r).
struct MpTrueResidualBody {
  const double* u;
  const double* u0;
  const double* kx;
  const double* ky;
  float* r32;
  const float* w;
  int x;
  __device__ void operator()(int i, int, int, double (&out)[3]) const {
    const double kx0 = kx[i], kx1 = kx[i + 1], ky0 = ky[i], ky1 = ky[i + x];
    const double diag = (1.0 + (kx1 + kx0)) + (ky1 + ky0);
    const double au = (diag * u[i] - (kx1 * u[i + 1] + kx0 * u[i - 1])) -
                      (ky1 * u[i + x] + ky0 * u[i - x]);
    const double rt = u0[i] - au;
    const float rf = (float)rt;
    r32[i] = rf;
    out[0] = (double)rf * (double)rf;
    out[1] = rt * rt;
    out[2] = (double)rf * (double)w[i];
  }
};

int ensure_mp_buffers(tea_hip_chunk* c) {
  if (c->mp_r) return TEA_HIP_OK;
  const size_t bytes = (size_t)c->geo.n * sizeof(float);
  float** bufs[] = {&c->mp_p[0], &c->mp_p[1], &c->mp_p[2], &c->mp_r, &c->mp_w, &c->mp_kx,
                    &c->mp_ky};
  for (float** b : bufs) {
    if (hipMalloc(b, bytes) != hipSuccess) {
      *b = nullptr;
      set_error("hipMalloc of an FP32 CG buffer failed");
      return TEA_HIP_ENOMEM;
    }
    TEA_HIP_TRY(hipMemsetAsync(*b, 0, bytes, c->stream));
  }
  return TEA_HIP_OK;
}

}  // namespace

int apply_mp_hold(tea_hip_chunk* c) {
  const int n = c->hold;
  c->hold = 0;
  for (int i = 0; i < n; ++i) {
    Timed t(c, K_MP_APPLY_U);
    MpApplyUBody body{c->field[TEA_FIELD_U], c->mp_p[c->hold_buf[i]], c->hold_alpha[i],
                      c->geo};
    hipLaunchKernelGGL((each_kernel<MpApplyUBody>), flat_grid(c->geo.n), dim3(kBlock), 0,
                       c->stream, c->geo, body);
    TEA_HIP_TRY(hipGetLastError());
  }
  return TEA_HIP_OK;
}

int launch_mp_w4(tea_hip_chunk* c, MpArgs a, int kid) {
  const int cur = c->mp_cur, dst = mp_next(cur), prev = mp_next(dst);
  int um = a.apply_u ? 1 : 0;
  // Held updates are those of the two steps before this one: their p are
  // in p_new's buffer (the older) and in the buffer before p_old's.
  const bool ride = um == 1 && a.apply_p &&
                    (c->hold == 0 || (c->hold == 1 && c->hold_buf[0] == prev) ||
                     (c->hold == 2 && c->hold_buf[0] == dst && c->hold_buf[1] == prev));
  if (c->hold && !ride) TEA_RETURN_IF(apply_mp_hold(c));
  if (c->hold == 2) {
    um = 3;
    a.p_h = c->mp_p[dst];
    a.alpha_h = c->hold_alpha[0];
    a.p_h2 = c->mp_p[prev];
    a.alpha_h2 = c->hold_alpha[1];
    c->hold = 0;
  } else if (um == 1 && a.apply_p && c->hold_ok) {
    // Held: p_old's buffer is not written until two steps on.
    um = 0;
    c->hold_buf[c->hold] = cur;
    c->hold_alpha[c->hold] = a.alpha;
    ++c->hold;
  }
  Timed t(c, kid);
  // stack_factor(c) blocks one row apart per workgroup (stacked_block).
  const int S = stack_rows(c), ks = stack_factor(c);
  const dim3 grid = stacked_grid(c, ks), block(kQuadThreads * ks);
#define TEA_MP_W4(UMV)                                                                    \
  do {                                                                                    \
    if (ks == 1)                                                                          \
      hipLaunchKernelGGL((mp_w4_kernel<true, UMV, 1>), grid, block, 0, c->stream, c->geo, \
                         a, c->sub, c->nblk, S);                                          \
    else if (ks == 2)                                                                     \
      hipLaunchKernelGGL((mp_w4_kernel<true, UMV, 2>), grid, block, 0, c->stream, c->geo, \
                         a, c->sub, c->nblk, S);                                          \
    else                                                                                  \
      hipLaunchKernelGGL((mp_w4_kernel<true, UMV, 4>), grid, block, 0, c->stream, c->geo, \
                         a, c->sub, c->nblk, S);                                          \
  } while (0)
  if (um == 0)
    TEA_MP_W4(0);
  else if (um == 1)
    TEA_MP_W4(1);
  else
    TEA_MP_W4(3);
#undef TEA_MP_W4
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// r -= alpha w and the partials of r.r (and, with rw, of r.w in the second
// array) into c->partials (not waited for); *len is the number of partials
// left to finish in each array.
int launch_mp_r(tea_hip_chunk* c, double alpha, int* len, bool rw, bool finish) {
  *len = c->nspan;
  TEA_CHECK_ARG(finish || use_quad_blocks(c, 2), "launch_mp_r: unfinished sums need quads");
  if (use_quad_blocks(c, 2)) {
    {
      Timed t(c, K_MP_R);
      if (rw)
        hipLaunchKernelGGL(mp_r4_kernel<true>, quad_grid(c), dim3(kQuadThreads), 0, c->stream,
                           c->geo, c->mp_r, c->mp_w, alpha, c->sub, c->nblk);
      else
        hipLaunchKernelGGL(mp_r4_kernel<false>, quad_grid(c), dim3(kQuadThreads), 0, c->stream,
                           c->geo, c->mp_r, c->mp_w, alpha, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    if (!finish) return TEA_HIP_OK;
    return launch_span_finish(c, rw ? 2 : 1, len, 0, rw);
  }
  if (use_pair_blocks(c, 2)) {
    {
      Timed t(c, K_MP_R);
      if (rw)
        hipLaunchKernelGGL(mp_r1_kernel<true>, pair_grid(c), dim3(kBlock), 0, c->stream,
                           c->geo, c->mp_r, c->mp_w, alpha, c->sub, c->nblk);
      else
        hipLaunchKernelGGL(mp_r1_kernel<false>, pair_grid(c), dim3(kBlock), 0, c->stream,
                           c->geo, c->mp_r, c->mp_w, alpha, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    return launch_span_finish(c, rw ? 2 : 1, len, 0, rw);
  }
  Timed t(c, K_MP_R);
  const int sg = c->ngroup == 1 ? 1 : 0;
  if (c->pairs) {
    if (rw)
      hipLaunchKernelGGL(mp_r2_kernel<true>, span_grid(c), dim3(kBlock), 0, c->stream,
                         c->geo, c->mp_r, c->mp_w, alpha, c->partials, c->nspan, sg);
    else
      hipLaunchKernelGGL(mp_r2_kernel<false>, span_grid(c), dim3(kBlock), 0, c->stream,
                         c->geo, c->mp_r, c->mp_w, alpha, c->partials, c->nspan, sg);
  } else {
    if (rw)
      hipLaunchKernelGGL(mp_r_kernel<true>, span_grid(c), dim3(kBlock), 0, c->stream,
                         c->geo, c->mp_r, c->mp_w, alpha, c->partials, c->nspan, sg);
    else
      hipLaunchKernelGGL(mp_r_kernel<false>, span_grid(c), dim3(kBlock), 0, c->stream,
                         c->geo, c->mp_r, c->mp_w, alpha, c->partials, c->nspan, sg);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

}  // namespace tea

using namespace tea;

extern "C" {

int tea_hip_mpcg_init(tea_hip_chunk* c, int coefficient, double rx, double ry,
                      double* rr, double* bb) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(rr && bb, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  c->faces_ok_u = c->faces_ok_p = false;
  TEA_RETURN_IF(ensure_mp_buffers(c));
  // FP64 initialisation exactly as the plain path: kx, ky, u, w, r, p.
  double rro64 = 0.0;
  TEA_RETURN_IF(tea_hip_cg_init(c, coefficient, rx, ry, &rro64));
  static const int ext[4] = {-1, -1, -1, -1};
  int fields[TEA_HALO_NUM] = {0};
  fields[TEA_HALO_U] = fields[TEA_HALO_P] = 1;
  TEA_RETURN_IF(launch_halos(c, 1, ext, fields));
  TEA_RETURN_IF(tea_hip_copy_u(c));
  double** f = c->field;
  const Geo g = c->geo;
  {
    Timed t(c, K_MP_CONVERT);
    hipLaunchKernelGGL(mp_convert_all_kernel, flat_grid(g.n), dim3(kBlock), 0,
                       c->stream, g.n, f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                       f[TEA_FIELD_P], f[TEA_FIELD_R], c->mp_kx, c->mp_ky,
                       c->mp_p[0], c->mp_p[1], c->mp_p[2], c->mp_r);
    MpConvertBody body{f[TEA_FIELD_R], f[TEA_FIELD_U0]};
    hipLaunchKernelGGL((span_kernel<2, MpConvertBody>), span_grid(c),
                       dim3(kBlock), 0, c->stream, g, body, c->partials,
                       c->nspan, c->ngroup == 1 ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  c->mp_cur = 0;
  double out[2];
  TEA_RETURN_IF(finish_reduction(c, 2, out));
  *rr = out[0];
  *bb = out[1];
  return TEA_HIP_OK;
}

int tea_hip_mpcg_calc_w(tea_hip_chunk* c, double alpha, int apply_u,
                        double beta, int apply_p, double* pw, double* pr) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->mp_r != nullptr, "tea_hip_mpcg_init has not been called");
  TEA_CHECK_ARG(pw != nullptr && pr != nullptr, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD_KEEP_U(c);
  // The two-dot kernel of flexible CG with z = r: p = r + beta p_old,
  // w = A p, p.w and p.r.
  MpArgs a{c->mp_p[c->mp_cur], c->mp_p[mp_next(c->mp_cur)], c->mp_r,
           c->field[TEA_FIELD_U], c->mp_w, c->mp_kx, c->mp_ky, c->mp_r,
           alpha, beta, apply_u ? 1 : 0, apply_p ? 1 : 0};
  int len = c->nspan;
  if (use_pair_blocks(c, 1)) {
    if (c->quads) {
      TEA_RETURN_IF(launch_mp_w4(c, a, K_MP_W));
    } else {
      if (c->hold) TEA_RETURN_IF(apply_mp_hold(c));
      Timed t(c, K_MP_W);
      hipLaunchKernelGGL((mp_w1_kernel<true>), pair_grid(c), dim3(kBlock), 0,
                         c->stream, c->geo, a, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    TEA_RETURN_IF(launch_span_finish(c, 2, &len, 0, true));
  } else {
    if (c->hold) TEA_RETURN_IF(apply_mp_hold(c));
    Timed t(c, K_MP_W);
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
  if (apply_p) c->mp_cur = mp_next(c->mp_cur);
  double out[2];
  TEA_RETURN_IF(finish_reduction_n(c, 2, len, out));
  *pw = out[0];
  *pr = out[1];
  return TEA_HIP_OK;
}

int tea_hip_mpcg_calc_r(tea_hip_chunk* c, double alpha, double* rr, double* rw) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->mp_r != nullptr, "tea_hip_mpcg_init has not been called");
  TEA_CHECK_ARG(rr != nullptr && rw != nullptr, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD_KEEP_U(c);
  int len;
  TEA_RETURN_IF(launch_mp_r(c, alpha, &len, true));
  double out[2];
  TEA_RETURN_IF(finish_reduction_n(c, 2, len, out));
  *rr = out[0];
  *rw = out[1];
  return TEA_HIP_OK;
}

int tea_hip_mpcg_refresh(tea_hip_chunk* c, double alpha, int apply_u,
                         double* rr, double* rr_true, double* rw) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->mp_r != nullptr, "tea_hip_mpcg_init has not been called");
  TEA_CHECK_ARG(rr && rr_true && rw, "null outputs");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  double** f = c->field;
  const Geo g = c->geo;
  if (apply_u) {
    Timed t(c, K_MP_APPLY_U);
    MpApplyUBody body{f[TEA_FIELD_U], c->mp_p[c->mp_cur], alpha, g};
    hipLaunchKernelGGL((each_kernel<MpApplyUBody>), flat_grid(g.n),
                       dim3(kBlock), 0, c->stream, g, body);
  }
  double out[3];
  if (use_pair_blocks(c, 2)) {
    {
      Timed t(c, K_MP_REFRESH);
      hipLaunchKernelGGL(mp_residual1_kernel, pair_grid(c), dim3(kBlock), 0, c->stream, g,
                         f[TEA_FIELD_U], f[TEA_FIELD_U0], f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                         c->mp_r, c->mp_w, c->sub, c->nblk);
    }
    TEA_HIP_TRY(hipGetLastError());
    int len = c->nspan;
    TEA_RETURN_IF(launch_span_finish(c, 3, &len, 0, true));
    TEA_RETURN_IF(finish_reduction_n(c, 3, len, out));
  } else {
    {
      Timed t(c, K_MP_REFRESH);
      MpTrueResidualBody body{f[TEA_FIELD_U], f[TEA_FIELD_U0], f[TEA_FIELD_KX],
                              f[TEA_FIELD_KY], c->mp_r, c->mp_w, g.x};
      hipLaunchKernelGGL((span_kernel<3, MpTrueResidualBody>), span_grid(c),
                         dim3(kBlock), 0, c->stream, g, body, c->partials,
                         c->nspan, c->ngroup == 1 ? 1 : 0);
    }
    TEA_HIP_TRY(hipGetLastError());
    TEA_RETURN_IF(finish_reduction(c, 3, out));
  }
  *rr = out[0];
  *rr_true = out[1];
  *rw = out[2];
  return TEA_HIP_OK;
}

int tea_hip_mp_read(tea_hip_chunk* c, int which, float* host, size_t count) {
  TEA_CHECK_ARG(c != nullptr, "null chunk handle");
  TEA_CHECK_ARG(c->mp_r != nullptr, "tea_hip_mpcg_init has not been called");
  TEA_CHECK_ARG(count <= (size_t)c->geo.n, "count exceeds the field length");
  const float* src = which == TEA_MP_P    ? c->mp_p[c->mp_cur]
                     : which == TEA_MP_R  ? c->mp_r
                     : which == TEA_MP_W  ? c->mp_w
                     : which == TEA_MP_KX ? c->mp_kx
                     : which == TEA_MP_KY ? c->mp_ky
                                          : nullptr;
  TEA_CHECK_ARG(src != nullptr, "bad FP32 field id");
  TEA_HIP_TRY(hipSetDevice(c->device));
  TEA_FLUSH_HELD(c);
  TEA_HIP_TRY(hipMemcpyAsync(host, src, count * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
  return sync_chunk(c);
}

}  // extern "C"
