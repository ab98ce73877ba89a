// libteahip.so: the reference's Jacobi, Chebyshev and PPCG kernels
// (TeaLeaf/c_kernels/sycl/jacobi.cpp, cheby.cpp, ppcg.cpp, composed as
// TeaLeaf/c_kernels/sycl/kernel_interface.cpp:246-328 composes them).
//
// Each iteration entry point is one streaming pass in the default (fused)
// mode, with one cell pair per lane as in the fused CG kernels; consecutive
// Chebyshev steps and PPCG inner steps run two to a pass
// (cheby_two_step_kernel, ppcg_two_step_kernel):
//   Jacobi     reads u (5-point), u0, kx, ky; writes the new u into r's
//              buffer, the two fields' addresses are then exchanged, so the
//              old u is "r = u" without a copy; reduces sum |u_new - u|.
//              40 bytes per cell against the reference's 56 (copy 16 +
//              sweep 40).
//   Chebyshev  reads u (5-point), u0, kx, ky, p; writes w, r, p and the new
//              u into a spare buffer exchanged with u (the stencil still
//              reads the old u).  72 bytes per cell against 88.
//   PPCG inner reads sd (5-point), kx, ky, r, u; writes r, u and the new sd
//              into the spare buffer exchanged with sd.  64 bytes per cell
//              against 80.
// A double-buffered field's cells outside the interior are copied into the
// new buffer in the same pass, so every field holds exactly what the
// reference's in-place kernels leave.  Mode 0 (tea_hip_set_fused) and odd
// row lengths run one launch per reference kernel instead.  Every FP64
// expression keeps the reference's operand order; the library is compiled
// with -ffp-contract=off.

#include <hip/hip_runtime.h>

#include <utility>

#include "tea_hip_impl.h"

namespace tea {
namespace {

// The 5-point operator of TeaLeaf/shared.h:59-63 at one cell from its
// operand values c (cell), l, r, b, t (left, right, bottom, top) and face
// coefficients kx0 = kx[i], kx1 = kx[i+1], ky0 = ky[i], ky1 = ky[i+x].
__device__ inline double op5(double c, double l, double r, double b, double t,
                             double kx0, double kx1, double ky0, double ky1) {
  const double diag = (1.0 + (kx1 + kx0)) + (ky1 + ky0);
  return (diag * c - (kx1 * r + kx0 * l)) - (ky1 * t + ky0 * b);
}

// The Jacobi update of jacobi.cpp:88-91 from the same operands (r is the
// previous iterate).
__device__ inline double jacobi_cell(double u0, double l, double r, double b,
                                     double t, double kx0, double kx1,
                                     double ky0, double ky1) {
  return ((u0 + (kx1 * r + kx0 * l)) + (ky1 * t + ky0 * b)) /
         ((1.0 + (kx0 + kx1)) + (ky0 + ky1));
}

// ---- one launch per reference kernel --------------------------------------

struct JacobiInitBody {  // jacobi.cpp:7-54
  double* __restrict__ u;
  double* __restrict__ u0;
  const double* __restrict__ density;
  const double* __restrict__ energy;
  double* __restrict__ kx;
  double* __restrict__ ky;
  double rx, ry;
  int coefficient, x, y, h;
  __device__ void operator()(int i, int kk, int jj) const {
    if (kk > 0 && kk < x - 1 && jj > 0 && jj < y - 1) {
      const double v = energy[i] * density[i];
      u0[i] = v;
      u[i] = v;
    }
    if (jj >= h && jj < y - 1 && kk >= h && kk < x - 1) {
      const bool cond = coefficient == TEA_CONDUCTIVITY;
      const double dc = cond ? density[i] : 1.0 / density[i];
      const double dl = cond ? density[i - 1] : 1.0 / density[i - 1];
      const double dd = cond ? density[i - x] : 1.0 / density[i - x];
      kx[i] = rx * (dl + dc) / (2.0 * dl * dc);
      ky[i] = ry * (dd + dc) / (2.0 * dd * dc);
    }
  }
};

struct CopyAllBody {  // jacobi.cpp:120-137: r = u on every cell
  const double* __restrict__ src;
  double* __restrict__ dst;
  __device__ void operator()(int i, int, int) const { dst[i] = src[i]; }
};

struct JacobiBody {  // jacobi.cpp:57-117
  double* __restrict__ u;
  const double* __restrict__ u0;
  const double* __restrict__ r;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  int x;
  __device__ void operator()(int i, int, int, double (&out)[1]) const {
    const double v = jacobi_cell(u0[i], r[i - 1], r[i + 1], r[i - x], r[i + x],
                                 kx[i], kx[i + 1], ky[i], ky[i + x]);
    u[i] = v;
    out[0] = fabs(v - r[i]);
  }
};

struct ChebyInitBody {  // cheby.cpp:7-43
  double* __restrict__ p;
  double* __restrict__ r;
  const double* __restrict__ u;
  const double* __restrict__ u0;
  double* __restrict__ w;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  double theta;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const int x = g.x;
    const double s = op5(u[i], u[i - 1], u[i + 1], u[i - x], u[i + x], kx[i],
                         kx[i + 1], ky[i], ky[i + x]);
    w[i] = s;
    const double rv = u0[i] - s;
    r[i] = rv;
    p[i] = rv / theta;
  }
};

struct ChebyIterateBody {  // cheby.cpp:73-110
  double* __restrict__ p;
  double* __restrict__ r;
  const double* __restrict__ u;
  const double* __restrict__ u0;
  double* __restrict__ w;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  double alpha, beta;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const int x = g.x;
    const double s = op5(u[i], u[i - 1], u[i + 1], u[i - x], u[i + x], kx[i],
                         kx[i + 1], ky[i], ky[i + x]);
    w[i] = s;
    const double rv = u0[i] - s;
    r[i] = rv;
    p[i] = alpha * p[i] + beta * rv;
  }
};

struct AddBody {  // cheby.cpp:46-70 (u += p), ppcg.cpp:58 (u += sd)
  double* __restrict__ u;
  const double* __restrict__ d;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) u[i] = u[i] + d[i];
  }
};

struct DivideBody {  // ppcg.cpp:7-31: sd = r / theta
  double* __restrict__ sd;
  const double* __restrict__ r;
  double theta;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) sd[i] = r[i] / theta;
  }
};

struct PpcgCalcURBody {  // ppcg.cpp:34-66
  const double* __restrict__ sd;
  double* __restrict__ r;
  double* __restrict__ u;
  const double* __restrict__ kx;
  const double* __restrict__ ky;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (!interior(g, kk, jj)) return;
    const int x = g.x;
    const double sc = sd[i];
    r[i] = r[i] - op5(sc, sd[i - 1], sd[i + 1], sd[i - x], sd[i + x], kx[i],
                      kx[i + 1], ky[i], ky[i + x]);
    u[i] = u[i] + sc;
  }
};

struct PpcgCalcSDBody {  // ppcg.cpp:69-94
  double* __restrict__ sd;
  const double* __restrict__ r;
  double alpha, beta;
  Geo g;
  __device__ void operator()(int i, int kk, int jj) const {
    if (interior(g, kk, jj)) sd[i] = alpha * sd[i] + beta * r[i];
  }
};

// ---- one pair per lane --------------------------------------------------------
//
// Lane m of block b holds cells e = 512 b + 2m and e + 1 (x is even, so the
// pair never straddles a row).  The stencil operand's left and right
// neighbours come from the adjacent lanes; lanes 0 and 63 load theirs.

// A stencil operand at a pair: the pair itself, the rows above and below,
// and the cells left of e and right of e + 1.  Must be called by every lane
// of the wave (it shuffles).  Lane 0's and lane 63's extra loads are at
// ir - 1 and ir + 2, in range for an interior pair and for x + 2.
struct Operand {
  double2 c, t, b;
  double l, r;
};

__device__ inline Operand load_operand(const double* __restrict__ a,
                                       const PairAt& q, int x, int lane) {
  Operand o;
  o.c = ld2(a, q.ic);
  o.t = ld2(a, q.ir + x);
  o.b = ld2(a, q.ir - x);
  double edge = 0.0;
  if (lane == 0 || lane == kWave - 1) edge = a[lane == 0 ? q.ir - 1 : q.ir + 2];
  o.l = __shfl_up(o.c.y, 1);
  o.r = __shfl_down(o.c.x, 1);
  if (lane == 0) o.l = edge;
  if (lane == kWave - 1) o.r = edge;
  return o;
}

// Face coefficients of a pair: kx at e, e + 1 and e + 2, ky at e, e + 1
// and one row up.
struct Faces {
  double2 kx, ky, kyt;
  double kx2;
};

__device__ inline Faces load_faces(const double* __restrict__ kx,
                                   const double* __restrict__ ky,
                                   const PairAt& q, int x, int lane) {
  Faces f;
  f.kx = ld2(kx, q.ic);
  f.ky = ld2(ky, q.ir);
  f.kyt = ld2(ky, q.ir + x);
  double edge = 0.0;
  if (lane == kWave - 1) edge = kx[q.ir + 2];
  f.kx2 = __shfl_down(f.kx.x, 1);
  if (lane == kWave - 1) f.kx2 = edge;
  return f;
}

__device__ inline double op5_lo(const Operand& o, const Faces& f) {
  return op5(o.c.x, o.l, o.c.y, o.b.x, o.t.x, f.kx.x, f.kx.y, f.ky.x, f.kyt.x);
}

__device__ inline double op5_hi(const Operand& o, const Faces& f) {
  return op5(o.c.y, o.c.x, o.r, o.b.y, o.t.y, f.kx.y, f.kx2, f.ky.y, f.kyt.y);
}

// Stores both cells of a valid pair with one 16-byte streaming store.
__device__ inline void st2_all(double* p, const PairAt& q, double v0,
                               double v1) {
  if (q.valid) st2_nt(p, q.e, true, true, v0, v1);
}

template <int KS>
__global__ __launch_bounds__(kBlock * KS) void jacobi_pair_kernel(
    Geo g, const double* __restrict__ u_old, double* __restrict__ u_new,
    const double* __restrict__ u0, const double* __restrict__ kx,
    const double* __restrict__ ky, double* sub, int nblk, int S) {
  int tid;
  const int blk = stacked_block<KS, kBlock>(S, &tid);
  if (blk >= nblk) return;
  const int lane = threadIdx.x & (kWave - 1);
  const PairAt q = pair_at(g, blk * kPairBlock + 2 * tid);
  const Operand o = load_operand(u_old, q, g.x, lane);
  const Faces f = load_faces(kx, ky, q, g.x, lane);
  const double2 s = ld2_nt(u0, q.ir);
  const double n0 = q.in0 ? jacobi_cell(s.x, o.l, o.c.y, o.b.x, o.t.x, f.kx.x,
                                        f.kx.y, f.ky.x, f.kyt.x)
                          : o.c.x;
  const double n1 = q.in1 ? jacobi_cell(s.y, o.c.x, o.r, o.b.y, o.t.y, f.kx.y,
                                        f.kx2, f.ky.y, f.kyt.y)
                          : o.c.y;
  st2_all(u_new, q, n0, n1);
  double v[1] = {(q.in0 ? fabs(n0 - o.c.x) : 0.0) +
                 (q.in1 ? fabs(n1 - o.c.y) : 0.0)};
  wave_subtree_pairs<1>(v, sub, nblk, blk, tid / kWave);
}

__global__ __launch_bounds__(kBlock) void cheby_pair_kernel(
    Geo g, const double* __restrict__ u_old, double* __restrict__ u_new,
    const double* __restrict__ u0, double* __restrict__ p,
    double* __restrict__ r, double* __restrict__ w,
    const double* __restrict__ kx, const double* __restrict__ ky,
    double alpha, double beta) {
  const int lane = threadIdx.x & (kWave - 1);
  const PairAt q = pair_at(g, blockIdx.x * kPairBlock + 2 * threadIdx.x);
  const Operand o = load_operand(u_old, q, g.x, lane);
  const Faces f = load_faces(kx, ky, q, g.x, lane);
  const double2 s0 = ld2_nt(u0, q.ir);
  const double2 pv = ld2_nt(p, q.ir);
  if (!q.valid) return;
  const double w0 = op5_lo(o, f), w1 = op5_hi(o, f);
  const double r0 = s0.x - w0, r1 = s0.y - w1;
  const double p0 = alpha * pv.x + beta * r0, p1 = alpha * pv.y + beta * r1;
  st2_nt(w, q.e, q.in0, q.in1, w0, w1);
  st2_nt(r, q.e, q.in0, q.in1, r0, r1);
  st2_nt(p, q.e, q.in0, q.in1, p0, p1);
  st2_all(u_new, q, q.in0 ? o.c.x + p0 : o.c.x, q.in1 ? o.c.y + p1 : o.c.y);
}

__global__ __launch_bounds__(kBlock) void ppcg_pair_kernel(
    Geo g, const double* __restrict__ sd_old, double* __restrict__ sd_new,
    double* __restrict__ r, double* __restrict__ u,
    const double* __restrict__ kx, const double* __restrict__ ky,
    double alpha, double beta) {
  const int lane = threadIdx.x & (kWave - 1);
  const PairAt q = pair_at(g, blockIdx.x * kPairBlock + 2 * threadIdx.x);
  const Operand o = load_operand(sd_old, q, g.x, lane);
  const Faces f = load_faces(kx, ky, q, g.x, lane);
  const double2 rv = ld2_nt(r, q.ir);
  const double2 uv = ld2_nt(u, q.ir);
  if (!q.valid) return;
  const double r0 = rv.x - op5_lo(o, f), r1 = rv.y - op5_hi(o, f);
  st2_nt(r, q.e, q.in0, q.in1, r0, r1);
  st2_nt(u, q.e, q.in0, q.in1, uv.x + o.c.x, uv.y + o.c.y);
  st2_all(sd_new, q, q.in0 ? alpha * o.c.x + beta * r0 : o.c.x,
          q.in1 ? alpha * o.c.y + beta * r1 : o.c.y);
}

template <class Body>
int launch_each_cell(tea_hip_chunk* c, int kid, const Body& body) {
  {
    Timed t(c, kid);
    hipLaunchKernelGGL((each_kernel<Body>), flat_grid(c->geo.n), dim3(kBlock),
                       0, c->stream, c->geo, body);
  }
  TEA_HIP_TRY(hipGetLastError());
  return TEA_HIP_OK;
}

// The fused Chebyshev and PPCG passes need the x*y spare buffer.
int ensure_spare(tea_hip_chunk* c) {
  if (c->spare) return TEA_HIP_OK;
  if (hipMalloc(&c->spare, (size_t)c->geo.n * sizeof(double)) != hipSuccess) {
    c->spare = nullptr;
    set_error("hipMalloc of the spare field buffer failed");
    return TEA_HIP_ENOMEM;
  }
  return TEA_HIP_OK;
}

// Fused passes run on even row lengths in the default mode.
bool fused_pairs(const tea_hip_chunk* c) { return c->fused && c->pairs; }

// ---- two PPCG inner steps in one pass ---------------------------------------
//
// Steps k and k+1 (ppcg.cpp:34-94) with the depth-1 reflective SD update
// between them (local_halos.cpp:7-102).  A block owns a tile of kTileY rows
// x kTileX columns of the interior; each thread kRows rows of one column
// (one cell at 16 x 64 with 1024 threads, which ran 10-15% faster than
// 8 x 128 with 256 threads and four rows each) and at most one cell of the
// margin below.  Every global read is issued before the first barrier.
//   1. sd_k on the tile and two cells around it is staged in LDS;
//   2. sd_(k+1) = a0 sd_k + b0 (r_k - A sd_k) is formed on the tile and one
//      cell around it (the tile's r_(k+1) stays in registers);
//   3. cells of that margin on the halo ring take their mirror's value, as
//      the halo update gives them;
//   4. step k+1 for the tile's cells: u in place, r and sd_(k+2) into second
//      buffers (neighbouring tiles still read the old sd and r);
//   5. tiles on the edge of the interior write the second buffers' cells
//      outside it: for sd what the halo update gives, r copied.
// sd_k and r_k are read once (the margins from L2), u, r and sd written once:
// 64 bytes per cell for two steps against 128 for two single passes.  When
// cg_calc_ur's u += alpha p is still pending (the first pass of a PPCG outer
// step), the pass also reads p and adds alpha p to u before the steps' sd.  Every
// value uses the single-step kernel's expressions, so the result is bitwise
// that of two single steps with the halo update between them.
constexpr int kTileX = 64;
constexpr int kTileY = 16;
constexpr int kTSBlock = 1024;                     // threads per tile
constexpr int kRows = kTileY * kTileX / kTSBlock;  // rows per thread
constexpr int kS0W = kTileX + 4, kS0H = kTileY + 4;
constexpr int kS1W = kTileX + 2, kS1H = kTileY + 2;
constexpr int kMargin = 2 * kS1W + 2 * kTileY;     // step-1 margin cells
constexpr int kMarginPer = (kMargin + kTSBlock - 1) / kTSBlock;
static_assert(kTileX % kWave == 0 && kTSBlock % kTileX == 0, "tile shape");
static_assert(kMarginPer == 1, "one margin cell per thread");

// The tile of this block.  The hardware deals blocks to the 8 XCDs in
// turn (block b to XCD b % 8); XCD k takes the k-th eighth of the tiles in
// row-major order, so a tile's neighbours, whose rows and boundary cache
// lines it re-reads, sit in the same L2.
__device__ inline int tile_of_block(int nt) {
  const int b = blockIdx.x, k = b & 7, per = nt >> 3, rem = nt & 7;
  return k * per + min(k, rem) + (b >> 3);
}

// Position (a, b) in the step-1 area of margin slot q: rows 0 and kS1H-1
// over all columns, then columns 0 and kS1W-1 over the rows between.
__device__ inline void margin_pos(int q, int& a, int& b) {
  if (q < 2 * kS1W) {
    a = q < kS1W ? 0 : kS1H - 1;
    b = q < kS1W ? q : q - kS1W;
  } else {
    const int e = q - 2 * kS1W;
    a = 1 + (e >> 1);
    b = (e & 1) ? kS1W - 1 : 0;
  }
}

__global__ __launch_bounds__(kTSBlock) void ppcg_two_step_kernel(
    Geo g, const double* __restrict__ sd_old, double* __restrict__ sd_new,
    const double* __restrict__ r, double* __restrict__ r_new,
    double* __restrict__ u, const double* __restrict__ kx,
    const double* __restrict__ ky, const double* __restrict__ p_pend,
    double alpha_pend, double a0, double b0, double a1, double b1,
    int tiles_x, int tiles_y, int pre_halo) {
  // sd_k (rows r0-2.., columns c0-2..) until step 2 is formed, then, in the
  // same storage, sd_(k+1) (rows r0-1.., columns c0-1..).
  __shared__ double lds[kS0H * kS0W];
  double (*S0)[kS0W] = reinterpret_cast<double (*)[kS0W]>(lds);
  double (*S1)[kS1W] = reinterpret_cast<double (*)[kS1W]>(lds);
  const int x = g.x, y = g.y, h = g.h;
  const int xe = x - h, ye = y - h;  // ends of the interior (exclusive)
  const int tile = tile_of_block(tiles_x * tiles_y);
  const int txi = tile % tiles_x, tyi = tile / tiles_x;
  const int c0 = h + txi * kTileX, r0 = h + tyi * kTileY;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const bool edge = txi == 0 || tyi == 0 || txi == tiles_x - 1 || tyi == tiles_y - 1;

  // 1. All global reads are issued before the block waits once: for the
  // tile's own column (column c0 + t) r, u, kx (through the right boundary
  // face; lane 63 also reads its right neighbour's) and ky (cells and the
  // row above); for the margin cells r, kx and ky; sd_k for the staging
  // area.
  const int tc = t % kTileX, tr = (t / kTileX) * kRows;
  const int kk = c0 + tc;
  double rk1[kRows], uo[kRows], kx0[kRows], kxe[kRows], kyv[kRows + 1];
#pragma unroll
  for (int m = 0; m < kRows; ++m) rk1[m] = uo[m] = kx0[m] = kxe[m] = kyv[m] = 0.0;
  kyv[kRows] = 0.0;
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    if (jj < ye && kk <= xe) kx0[m] = kx[kk + jj * x];
    if (kk < xe && jj < ye) {
      const int i = kk + jj * x;
      rk1[m] = r[i];
      uo[m] = u[i];
      // u += alpha p of cg_calc_ur, when still pending (cg.cpp:228).
      if (p_pend != nullptr) uo[m] = uo[m] + alpha_pend * p_pend[i];
      if (lane == kWave - 1) kxe[m] = kx[i + 1];
      kyv[m] = ky[i];
      kyv[m + 1] = ky[i + x];
    }
  }
  double mr[kMarginPer], mkx0[kMarginPer], mkx1[kMarginPer], mky0[kMarginPer],
      mky1[kMarginPer];
  int mpos[kMarginPer];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n) {
    const int q = t + n * kTSBlock;
    int a = 0, b = 0;
    margin_pos(q, a, b);
    const int jj = r0 - 1 + a, kc = c0 - 1 + b;
    mr[n] = mkx0[n] = mkx1[n] = mky0[n] = mky1[n] = 0.0;
    mpos[n] = -1;
    if (q < kMargin && jj >= h && jj < ye && kc >= h && kc < xe) {
      const int i = kc + jj * x;
      mr[n] = r[i];
      mkx0[n] = kx[i];
      mkx1[n] = kx[i + 1];
      mky0[n] = ky[i];
      mky1[n] = ky[i + x];
      mpos[n] = a * kS1W + b;
    }
  }
  // Stage sd_k (cells outside the field are never used: zero).
  constexpr int kStage = (kS0H * kS0W + kTSBlock - 1) / kTSBlock;
  double st[kStage];
#pragma unroll
  for (int n = 0; n < kStage; ++n) {
    const int q = t + n * kTSBlock;
    const int a = q / kS0W, b = q - a * kS0W;
    const int jj = r0 - 2 + a, kc = c0 - 2 + b;
    st[n] = (q < kS0H * kS0W && jj >= 0 && jj < y && kc >= 0 && kc < x)
                ? sd_old[kc + jj * x]
                : 0.0;
  }
#pragma unroll
  for (int n = 0; n < kStage; ++n) {
    const int q = t + n * kTSBlock;
    if (q < kS0H * kS0W) lds[q] = st[n];
  }
  __syncthreads();
  // A held-back depth-1 update of the staged field before step k: its ring
  // cells take their mirror's value (double mirror at the corners).  The
  // staging area reaches the ring in edge tiles and in tiles next to a
  // partial last tile.
  if (pre_halo && (c0 - 2 <= h - 1 || c0 + kTileX + 1 >= xe ||
                   r0 - 2 <= h - 1 || r0 + kTileY + 1 >= ye)) {
    for (int q = t; q < kS0H * kS0W; q += kTSBlock) {
      const int a = q / kS0W, b = q - a * kS0W;
      const int jj = r0 - 2 + a, kc = c0 - 2 + b;
      const bool ring_box = jj >= h - 1 && jj <= ye && kc >= h - 1 && kc <= xe;
      const bool inner = jj >= h && jj < ye && kc >= h && kc < xe;
      if (!ring_box || inner) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      S0[a][b] = S0[mj - r0 + 2][mk - c0 + 2];
    }
    __syncthreads();
  }

  // 2. sd_(k+1) on the tile's own cells (rk1 becomes r_(k+1)) ...
  double sdk[kRows], s1own[kRows];
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    double kx1 = __shfl_down(kx0[m], 1);
    if (lane == kWave - 1) kx1 = kxe[m];
    sdk[m] = s1own[m] = 0.0;
    if (kk < xe && jj < ye) {
      const double c = S0[tr + m + 2][tc + 2];
      const double rn = rk1[m] - op5(c, S0[tr + m + 2][tc + 1], S0[tr + m + 2][tc + 3],
                                     S0[tr + m + 1][tc + 2], S0[tr + m + 3][tc + 2], kx0[m],
                                     kx1, kyv[m], kyv[m + 1]);
      sdk[m] = c;
      rk1[m] = rn;
      s1own[m] = a0 * c + b0 * rn;
    }
  }
  // ... and on the one-cell margin: rows -1 and kTileY over columns
  // -1 .. kTileX, then columns -1 and kTileX over the tile's rows.  Margin
  // cells outside the interior stay zero here; step 3 mirrors them.
  double s1m[kMarginPer];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n) {
    const int p = mpos[n] < 0 ? 0 : mpos[n];
    const int a = p / kS1W, b = p - a * kS1W;
    const double c = S0[a + 1][b + 1];
    const double rn = mr[n] - op5(c, S0[a + 1][b], S0[a + 1][b + 2], S0[a][b + 1],
                                  S0[a + 2][b + 1], mkx0[n], mkx1[n], mky0[n],
                                  mky1[n]);
    s1m[n] = a0 * c + b0 * rn;
  }
  __syncthreads();  // every read of sd_k is done: S1 replaces it
#pragma unroll
  for (int m = 0; m < kRows; ++m) S1[tr + m + 1][tc + 1] = s1own[m];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n)
    if (mpos[n] >= 0) lds[mpos[n]] = s1m[n];
  __syncthreads();

  // 3. Ring cells of the margin take their mirror's value (double mirror at
  // the corners, as the halo update's left/right-then-bottom/top order gives).
  if (edge) {
    for (int q = t; q < kS1H * kS1W; q += kTSBlock) {
      const int a = q / kS1W, b = q - a * kS1W;
      const int jj = r0 - 1 + a, kc = c0 - 1 + b;
      const bool ring_box = jj >= h - 1 && jj <= ye && kc >= h - 1 && kc <= xe;
      const bool inner = jj >= h && jj < ye && kc >= h && kc < xe;
      if (!ring_box || inner) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      S1[a][b] = S1[mj - r0 + 1][mk - c0 + 1];
    }
    __syncthreads();
  }

  // 4. Step k+1 on the tile's own cells.
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    double kx1 = __shfl_down(kx0[m], 1);
    if (lane == kWave - 1) kx1 = kxe[m];
    if (kk < xe && jj < ye) {
      const int i = kk + jj * x;
      const double c = S1[tr + m + 1][tc + 1];
      const double rn = rk1[m] - op5(c, S1[tr + m + 1][tc], S1[tr + m + 1][tc + 2], S1[tr + m][tc + 1],
                                     S1[tr + m + 2][tc + 1], kx0[m], kx1, kyv[m],
                                     kyv[m + 1]);
      r_new[i] = rn;
      u[i] = (uo[m] + sdk[m]) + c;
      sd_new[i] = a1 * c + b1 * rn;
    }
  }

  // 5. The second buffers' cells outside the interior next to this tile.
  // The halo update copies whole columns h-1 and x-h from their inner
  // neighbours, then whole rows y-h and h-1: a cell takes the value at its
  // mirror (each coordinate on the ring moved one cell inwards), from
  // sd_(k+1) where that is interior, else the unchanged old value.  r
  // outside the interior is copied.
  if (edge) {
    const int ka = txi == 0 ? 0 : c0;
    const int kb = txi == tiles_x - 1 ? x : min(c0 + kTileX, xe);
    const int ja = tyi == 0 ? 0 : r0;
    const int jb = tyi == tiles_y - 1 ? y : min(r0 + kTileY, ye);
    const int w = kb - ka;
    for (int q = t; q < w * (jb - ja); q += kTSBlock) {
      const int jj = ja + q / w, kc = ka + q % w;
      if (jj >= h && jj < ye && kc >= h && kc < xe) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      const bool from_new = mj >= h && mj < ye && mk >= h && mk < xe;
      sd_new[kc + jj * x] =
          from_new ? S1[mj - r0 + 1][mk - c0 + 1] : sd_old[mk + mj * x];
      r_new[kc + jj * x] = r[kc + jj * x];
    }
  }
}

// Two Chebyshev steps (cheby.cpp:46-110) with the depth-1 reflective U
// update between them, as ppcg_two_step_kernel does for PPCG: u_k staged
// with two cells around the tile, u_(k+1) = u_k + p_(k+1) with
// p_(k+1) = a0 p_k + b0 (u0 - A u_k) formed on the tile and one cell around
// it, the ring cells of that margin mirrored, then step k+1 for the tile's
// cells.  u and p go to second buffers (neighbouring tiles still read the
// old ones); w and r, written only at a tile's own cells and never read,
// stay in place and end as the second step's.  72 bytes per cell for two
// steps against 144 for two single passes.
__global__ __launch_bounds__(kTSBlock) void cheby_two_step_kernel(
    Geo g, const double* __restrict__ u_old, double* __restrict__ u_new,
    const double* __restrict__ u0, const double* __restrict__ p_old,
    double* __restrict__ p_new, double* __restrict__ r, double* __restrict__ w,
    const double* __restrict__ kx, const double* __restrict__ ky, double a0,
    double b0, double a1, double b1, int tiles_x, int tiles_y, int pre_halo) {
  __shared__ double lds[kS0H * kS0W];
  double (*S0)[kS0W] = reinterpret_cast<double (*)[kS0W]>(lds);
  double (*S1)[kS1W] = reinterpret_cast<double (*)[kS1W]>(lds);
  const int x = g.x, y = g.y, h = g.h;
  const int xe = x - h, ye = y - h;
  const int tile = tile_of_block(tiles_x * tiles_y);
  const int txi = tile % tiles_x, tyi = tile / tiles_x;
  const int c0 = h + txi * kTileX, r0 = h + tyi * kTileY;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const bool edge = txi == 0 || tyi == 0 || txi == tiles_x - 1 || tyi == tiles_y - 1;

  // 1. The tile's own column (u0, p_k, kx through the right boundary face,
  // ky with the row above), the margin cells' u0, p_k, kx and ky, and u_k
  // for the staging area, before one wait.
  const int tc = t % kTileX, tr = (t / kTileX) * kRows;
  const int kk = c0 + tc;
  double bv[kRows], pk[kRows], kx0[kRows], kxe[kRows], kyv[kRows + 1];
#pragma unroll
  for (int m = 0; m < kRows; ++m) bv[m] = pk[m] = kx0[m] = kxe[m] = kyv[m] = 0.0;
  kyv[kRows] = 0.0;
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    if (jj < ye && kk <= xe) kx0[m] = kx[kk + jj * x];
    if (kk < xe && jj < ye) {
      const int i = kk + jj * x;
      bv[m] = u0[i];
      pk[m] = p_old[i];
      if (lane == kWave - 1) kxe[m] = kx[i + 1];
      kyv[m] = ky[i];
      kyv[m + 1] = ky[i + x];
    }
  }
  double mb[kMarginPer], mp[kMarginPer], mkx0[kMarginPer], mkx1[kMarginPer],
      mky0[kMarginPer], mky1[kMarginPer];
  int mpos[kMarginPer];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n) {
    const int q = t + n * kTSBlock;
    int a = 0, b = 0;
    margin_pos(q, a, b);
    const int jj = r0 - 1 + a, kc = c0 - 1 + b;
    mb[n] = mp[n] = mkx0[n] = mkx1[n] = mky0[n] = mky1[n] = 0.0;
    mpos[n] = -1;
    if (q < kMargin && jj >= h && jj < ye && kc >= h && kc < xe) {
      const int i = kc + jj * x;
      mb[n] = u0[i];
      mp[n] = p_old[i];
      mkx0[n] = kx[i];
      mkx1[n] = kx[i + 1];
      mky0[n] = ky[i];
      mky1[n] = ky[i + x];
      mpos[n] = a * kS1W + b;
    }
  }
  constexpr int kStage = (kS0H * kS0W + kTSBlock - 1) / kTSBlock;
  double st[kStage];
#pragma unroll
  for (int n = 0; n < kStage; ++n) {
    const int q = t + n * kTSBlock;
    const int a = q / kS0W, b = q - a * kS0W;
    const int jj = r0 - 2 + a, kc = c0 - 2 + b;
    st[n] = (q < kS0H * kS0W && jj >= 0 && jj < y && kc >= 0 && kc < x)
                ? u_old[kc + jj * x]
                : 0.0;
  }
#pragma unroll
  for (int n = 0; n < kStage; ++n) {
    const int q = t + n * kTSBlock;
    if (q < kS0H * kS0W) lds[q] = st[n];
  }
  __syncthreads();
  // A held-back depth-1 update of the staged field before step k: its ring
  // cells take their mirror's value (double mirror at the corners).  The
  // staging area reaches the ring in edge tiles and in tiles next to a
  // partial last tile.
  if (pre_halo && (c0 - 2 <= h - 1 || c0 + kTileX + 1 >= xe ||
                   r0 - 2 <= h - 1 || r0 + kTileY + 1 >= ye)) {
    for (int q = t; q < kS0H * kS0W; q += kTSBlock) {
      const int a = q / kS0W, b = q - a * kS0W;
      const int jj = r0 - 2 + a, kc = c0 - 2 + b;
      const bool ring_box = jj >= h - 1 && jj <= ye && kc >= h - 1 && kc <= xe;
      const bool inner = jj >= h && jj < ye && kc >= h && kc < xe;
      if (!ring_box || inner) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      S0[a][b] = S0[mj - r0 + 2][mk - c0 + 2];
    }
    __syncthreads();
  }

  // 2. u_(k+1) on the tile's own cells (pk becomes p_(k+1)) ...
  double u1own[kRows];
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    double kx1 = from_next_lane(kx0[m]);
    if (lane == kWave - 1) kx1 = kxe[m];
    u1own[m] = 0.0;
    if (kk < xe && jj < ye) {
      const double c = S0[tr + m + 2][tc + 2];
      const double wv = op5(c, S0[tr + m + 2][tc + 1], S0[tr + m + 2][tc + 3], S0[tr + m + 1][tc + 2],
                            S0[tr + m + 3][tc + 2], kx0[m], kx1, kyv[m], kyv[m + 1]);
      const double rv = bv[m] - wv;
      pk[m] = a0 * pk[m] + b0 * rv;
      u1own[m] = c + pk[m];
    }
  }
  // ... and on the one-cell margin (as in ppcg_two_step_kernel).
  double u1m[kMarginPer];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n) {
    const int q = mpos[n] < 0 ? 0 : mpos[n];
    const int a = q / kS1W, b = q - a * kS1W;
    const double c = S0[a + 1][b + 1];
    const double wv = op5(c, S0[a + 1][b], S0[a + 1][b + 2], S0[a][b + 1],
                          S0[a + 2][b + 1], mkx0[n], mkx1[n], mky0[n], mky1[n]);
    const double rv = mb[n] - wv;
    u1m[n] = c + (a0 * mp[n] + b0 * rv);
  }
  __syncthreads();  // every read of u_k is done: u_(k+1) replaces it
#pragma unroll
  for (int m = 0; m < kRows; ++m) S1[tr + m + 1][tc + 1] = u1own[m];
#pragma unroll
  for (int n = 0; n < kMarginPer; ++n)
    if (mpos[n] >= 0) lds[mpos[n]] = u1m[n];
  __syncthreads();

  // 3. Ring cells of the margin take their mirror's value.
  if (edge) {
    for (int q = t; q < kS1H * kS1W; q += kTSBlock) {
      const int a = q / kS1W, b = q - a * kS1W;
      const int jj = r0 - 1 + a, kc = c0 - 1 + b;
      const bool ring_box = jj >= h - 1 && jj <= ye && kc >= h - 1 && kc <= xe;
      const bool inner = jj >= h && jj < ye && kc >= h && kc < xe;
      if (!ring_box || inner) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      S1[a][b] = S1[mj - r0 + 1][mk - c0 + 1];
    }
    __syncthreads();
  }

  // 4. Step k+1 on the tile's own cells.
#pragma unroll
  for (int m = 0; m < kRows; ++m) {
    const int jj = r0 + tr + m;
    double kx1 = from_next_lane(kx0[m]);
    if (lane == kWave - 1) kx1 = kxe[m];
    if (kk < xe && jj < ye) {
      const int i = kk + jj * x;
      const double c = S1[tr + m + 1][tc + 1];
      const double wv = op5(c, S1[tr + m + 1][tc], S1[tr + m + 1][tc + 2], S1[tr + m][tc + 1],
                            S1[tr + m + 2][tc + 1], kx0[m], kx1, kyv[m], kyv[m + 1]);
      const double rv = bv[m] - wv;
      const double pv = a1 * pk[m] + b1 * rv;
      w[i] = wv;
      r[i] = rv;
      p_new[i] = pv;
      u_new[i] = c + pv;
    }
  }

  // 5. The second buffers' cells outside the interior next to this tile: u
  // as the halo update leaves it (see ppcg_two_step_kernel), p copied.
  if (edge) {
    const int ka = txi == 0 ? 0 : c0;
    const int kb = txi == tiles_x - 1 ? x : min(c0 + kTileX, xe);
    const int ja = tyi == 0 ? 0 : r0;
    const int jb = tyi == tiles_y - 1 ? y : min(r0 + kTileY, ye);
    const int wd = kb - ka;
    for (int q = t; q < wd * (jb - ja); q += kTSBlock) {
      const int jj = ja + q / wd, kc = ka + q % wd;
      if (jj >= h && jj < ye && kc >= h && kc < xe) continue;
      const int mj = jj == h - 1 ? h : (jj == ye ? ye - 1 : jj);
      const int mk = kc == h - 1 ? h : (kc == xe ? xe - 1 : kc);
      const bool from_new = mj >= h && mj < ye && mk >= h && mk < xe;
      u_new[kc + jj * x] =
          from_new ? S1[mj - r0 + 1][mk - c0 + 1] : u_old[mk + mj * x];
      p_new[kc + jj * x] = p_old[kc + jj * x];
    }
  }
}

// The second spare buffer, allocated on first use.
int ensure_spare2(tea_hip_chunk* c) {
  if (c->spare2) return TEA_HIP_OK;
  if (hipMalloc(&c->spare2, (size_t)c->geo.n * sizeof(double)) != hipSuccess) {
    c->spare2 = nullptr;
    set_error("hipMalloc of the second spare field buffer failed");
    return TEA_HIP_ENOMEM;
  }
  return TEA_HIP_OK;
}

// One Chebyshev step: the single pass (even rows) or one launch per
// reference kernel.
int cheby_step(tea_hip_chunk* c, double alpha, double beta) {
  double** f = c->field;
  const Geo g = c->geo;
  if (!fused_pairs(c)) {
    TEA_RETURN_IF(launch_each_cell(
        c, K_CHEBY_ITERATE,
        ChebyIterateBody{f[TEA_FIELD_P], f[TEA_FIELD_R], f[TEA_FIELD_U],
                         f[TEA_FIELD_U0], f[TEA_FIELD_W], f[TEA_FIELD_KX],
                         f[TEA_FIELD_KY], alpha, beta, g}));
    return launch_each_cell(c, K_CHEBY_CALC_U,
                            AddBody{f[TEA_FIELD_U], f[TEA_FIELD_P], g});
  }
  TEA_RETURN_IF(ensure_spare(c));
  {
    Timed t(c, K_CHEBY_FUSED);
    hipLaunchKernelGGL(cheby_pair_kernel, pair_grid(c), dim3(kBlock), 0,
                       c->stream, g, f[TEA_FIELD_U], c->spare, f[TEA_FIELD_U0],
                       f[TEA_FIELD_P], f[TEA_FIELD_R], f[TEA_FIELD_W],
                       f[TEA_FIELD_KX], f[TEA_FIELD_KY], alpha, beta);
  }
  TEA_HIP_TRY(hipGetLastError());
  std::swap(f[TEA_FIELD_U], c->spare);
  return TEA_HIP_OK;
}

// Two Chebyshev steps with the U halo update between them, in one pass.
int cheby_two_steps(tea_hip_chunk* c, double a0, double b0, double a1,
                    double b1, bool pre_halo) {
  double** f = c->field;
  const Geo g = c->geo;
  TEA_RETURN_IF(ensure_spare(c));
  TEA_RETURN_IF(ensure_spare2(c));
  const int tiles_x = (g.x - 2 * g.h + kTileX - 1) / kTileX;
  const int tiles_y = (g.y - 2 * g.h + kTileY - 1) / kTileY;
  {
    Timed t(c, K_CHEBY_TWO_STEP);
    hipLaunchKernelGGL(cheby_two_step_kernel, dim3(tiles_x * tiles_y),
                       dim3(kTSBlock), 0, c->stream, g, f[TEA_FIELD_U], c->spare,
                       f[TEA_FIELD_U0], f[TEA_FIELD_P], c->spare2, f[TEA_FIELD_R],
                       f[TEA_FIELD_W], f[TEA_FIELD_KX], f[TEA_FIELD_KY], a0, b0,
                       a1, b1, tiles_x, tiles_y, pre_halo ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  std::swap(f[TEA_FIELD_U], c->spare);
  std::swap(f[TEA_FIELD_P], c->spare2);
  c->p_alt_synced = false;
  if (pre_halo) c->pre_halo = -1;
  return TEA_HIP_OK;
}

// One inner step: the single pass (even rows) or one launch per reference
// kernel; the new sd goes to the second buffer in the single pass.
int ppcg_step(tea_hip_chunk* c, double alpha, double beta) {
  double** f = c->field;
  const Geo g = c->geo;
  if (!fused_pairs(c)) {
    TEA_RETURN_IF(launch_each_cell(
        c, K_PPCG_CALC_UR,
        PpcgCalcURBody{f[TEA_FIELD_SD], f[TEA_FIELD_R], f[TEA_FIELD_U],
                       f[TEA_FIELD_KX], f[TEA_FIELD_KY], g}));
    return launch_each_cell(
        c, K_PPCG_CALC_SD,
        PpcgCalcSDBody{f[TEA_FIELD_SD], f[TEA_FIELD_R], alpha, beta, g});
  }
  TEA_RETURN_IF(ensure_spare(c));
  {
    Timed t(c, K_PPCG_FUSED);
    hipLaunchKernelGGL(ppcg_pair_kernel, pair_grid(c), dim3(kBlock), 0,
                       c->stream, g, f[TEA_FIELD_SD], c->spare, f[TEA_FIELD_R],
                       f[TEA_FIELD_U], f[TEA_FIELD_KX], f[TEA_FIELD_KY], alpha,
                       beta);
  }
  TEA_HIP_TRY(hipGetLastError());
  std::swap(f[TEA_FIELD_SD], c->spare);
  return TEA_HIP_OK;
}

// Two inner steps with the SD halo update between them, in one pass.
int ppcg_two_steps(tea_hip_chunk* c, double a0, double b0, double a1,
                   double b1, bool pend_u, bool pre_halo) {
  double** f = c->field;
  const Geo g = c->geo;
  TEA_RETURN_IF(ensure_spare(c));
  TEA_RETURN_IF(ensure_spare2(c));
  const int tiles_x = (g.x - 2 * g.h + kTileX - 1) / kTileX;
  const int tiles_y = (g.y - 2 * g.h + kTileY - 1) / kTileY;
  {
    Timed t(c, K_PPCG_TWO_STEP);
    hipLaunchKernelGGL(ppcg_two_step_kernel, dim3(tiles_x * tiles_y),
                       dim3(kTSBlock), 0, c->stream, g, f[TEA_FIELD_SD], c->spare,
                       f[TEA_FIELD_R], c->spare2, f[TEA_FIELD_U], f[TEA_FIELD_KX],
                       f[TEA_FIELD_KY], pend_u ? f[TEA_FIELD_P] : nullptr,
                       c->pend_alpha, a0, b0, a1, b1, tiles_x, tiles_y,
                       pre_halo ? 1 : 0);
  }
  TEA_HIP_TRY(hipGetLastError());
  std::swap(f[TEA_FIELD_SD], c->spare);
  std::swap(f[TEA_FIELD_R], c->spare2);
  if (pend_u) c->pend_u = false;
  if (pre_halo) c->pre_halo = -1;
  return TEA_HIP_OK;
}

}  // namespace

int apply_pre_halo(tea_hip_chunk* c) {
  if (c->pre_halo < 0) return TEA_HIP_OK;
  static const int ext[4] = {-1, -1, -1, -1};
  int fields[TEA_HALO_NUM] = {0};
  fields[c->pre_halo] = 1;
  c->pre_halo = -1;
  return launch_halos(c, 1, ext, fields);
}

int flush_held(tea_hip_chunk* c) {
  const int kind = c->held;
  if (kind == kHeldNone && c->pre_halo < 0) return TEA_HIP_OK;
  TEA_RETURN_IF(apply_cg_pending(c));
  TEA_RETURN_IF(apply_pre_halo(c));
  if (kind == kHeldNone) return TEA_HIP_OK;
  const bool halo = c->held_halo;
  c->held = kHeldNone;
  c->held_halo = false;
  if (kind == kHeldPpcg)
    TEA_RETURN_IF(ppcg_step(c, c->held_alpha, c->held_beta));
  else
    TEA_RETURN_IF(cheby_step(c, c->held_alpha, c->held_beta));
  if (!halo) return TEA_HIP_OK;
  static const int ext[4] = {-1, -1, -1, -1};
  int fields[TEA_HALO_NUM] = {0};
  fields[kind == kHeldPpcg ? TEA_HALO_SD : TEA_HALO_U] = 1;
  return launch_halos(c, 1, ext, fields);
}

}  // namespace tea

using namespace tea;

extern "C" {

int tea_hip_jacobi_init(tea_hip_chunk* c, int coefficient, double rx,
                        double ry) {
  TEA_ENTER(c);
  TEA_CHECK_ARG(coefficient == TEA_CONDUCTIVITY ||
                    coefficient == TEA_RECIP_CONDUCTIVITY,
                "unknown coefficient");
  TEA_RETURN_IF(apply_pending(c));
  c->faces_ok_u = c->faces_ok_p = false;
  double** f = c->field;
  const Geo g = c->geo;
  return launch_each_cell(
      c, K_JACOBI_INIT,
      JacobiInitBody{f[TEA_FIELD_U], f[TEA_FIELD_U0], f[TEA_FIELD_DENSITY],
                     f[TEA_FIELD_ENERGY], f[TEA_FIELD_KX], f[TEA_FIELD_KY], rx,
                     ry, coefficient, g.x, g.y, g.h});
}

int tea_hip_jacobi_iterate(tea_hip_chunk* c, double* error) {
  TEA_ENTER(c);
  TEA_CHECK_ARG(error != nullptr, "null error");
  TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  const Geo g = c->geo;
  if (!(c->fused && use_pair_blocks(c, 1))) {
    TEA_RETURN_IF(launch_each_cell(c, K_JACOBI_COPY_U,
                                   CopyAllBody{f[TEA_FIELD_U], f[TEA_FIELD_R]}));
    {
      Timed t(c, K_JACOBI_ITERATE);
      hipLaunchKernelGGL(
          (span_kernel<1, JacobiBody>), span_grid(c), dim3(kBlock), 0,
          c->stream, g,
          JacobiBody{f[TEA_FIELD_U], f[TEA_FIELD_U0], f[TEA_FIELD_R],
                     f[TEA_FIELD_KX], f[TEA_FIELD_KY], g.x},
          c->partials, c->nspan, c->ngroup == 1 ? 1 : 0);
    }
    TEA_HIP_TRY(hipGetLastError());
    return finish_reduction(c, 1, error);
  }
  {
    Timed t(c, K_JACOBI_FUSED);
    // Workgroups of stack_factor(c) blocks one row apart (tea_hip_impl.h).
    const int ks = stack_factor(c), S = stack_rows(c);
    const dim3 grid = stacked_grid(c, ks), block(kBlock * ks);
    if (ks == 1)
      hipLaunchKernelGGL(jacobi_pair_kernel<1>, grid, block, 0, c->stream, g, f[TEA_FIELD_U],
                         f[TEA_FIELD_R], f[TEA_FIELD_U0], f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                         c->sub, c->nblk, S);
    else if (ks == 2)
      hipLaunchKernelGGL(jacobi_pair_kernel<2>, grid, block, 0, c->stream, g, f[TEA_FIELD_U],
                         f[TEA_FIELD_R], f[TEA_FIELD_U0], f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                         c->sub, c->nblk, S);
    else
      hipLaunchKernelGGL(jacobi_pair_kernel<4>, grid, block, 0, c->stream, g, f[TEA_FIELD_U],
                         f[TEA_FIELD_R], f[TEA_FIELD_U0], f[TEA_FIELD_KX], f[TEA_FIELD_KY],
                         c->sub, c->nblk, S);
  }
  TEA_HIP_TRY(hipGetLastError());
  // The new u is in r's buffer; the old u becomes r (the reference's copy).
  std::swap(f[TEA_FIELD_U], f[TEA_FIELD_R]);
  int len = c->nspan;
  TEA_RETURN_IF(launch_span_finish(c, 1, &len, 0, true));
  return finish_reduction_n(c, 1, len, error);
}

int tea_hip_cheby_init(tea_hip_chunk* c, double theta) {
  TEA_ENTER(c);
  TEA_RETURN_IF(apply_pending(c));
  c->faces_ok_u = c->faces_ok_p = false;
  double** f = c->field;
  return launch_each_cell(
      c, K_CHEBY_INIT,
      ChebyInitBody{f[TEA_FIELD_P], f[TEA_FIELD_R], f[TEA_FIELD_U],
                    f[TEA_FIELD_U0], f[TEA_FIELD_W], f[TEA_FIELD_KX],
                    f[TEA_FIELD_KY], theta, c->geo});
}

int tea_hip_cheby_iterate(tea_hip_chunk* c, double alpha, double beta) {
  TEA_ENTER_HELD(c);
  if (c->fused) {
    // Held back until the next step: with the U halo update between them
    // (all faces external), both steps run as one pass.
    if (c->held == kHeldCheby && c->held_halo) {
      c->held = kHeldNone;
      c->held_halo = false;
      TEA_RETURN_IF(apply_cg_pending(c));
      const bool pre = c->pre_halo == TEA_HALO_U;
      if (!pre) TEA_RETURN_IF(apply_pre_halo(c));
      return cheby_two_steps(c, c->held_alpha, c->held_beta, alpha, beta, pre);
    }
    // A held-back U update before this step stays held: the pass that runs
    // the step reflects the ring it stages.
    if (c->pre_halo == TEA_HALO_U && !c->held) {
      TEA_RETURN_IF(apply_cg_pending(c));
    } else {
      TEA_RETURN_IF(apply_pending(c));
    }
    c->held = kHeldCheby;
    c->held_halo = false;
    c->held_alpha = alpha;
    c->held_beta = beta;
    return TEA_HIP_OK;
  }
  TEA_RETURN_IF(apply_pending(c));
  return cheby_step(c, alpha, beta);
}

int tea_hip_ppcg_init(tea_hip_chunk* c, double theta) {
  TEA_ENTER(c);
  // sd = r / theta reads neither u nor p: u += alpha p, left pending by
  // cg_calc_ur, stays pending for the first inner pass to apply.
  if (c->pend_p || c->pend_halo) TEA_RETURN_IF(apply_pending(c));
  double** f = c->field;
  return launch_each_cell(
      c, K_PPCG_INIT, DivideBody{f[TEA_FIELD_SD], f[TEA_FIELD_R], theta, c->geo});
}

int tea_hip_ppcg_inner_iteration(tea_hip_chunk* c, double alpha,
                                 double beta) {
  TEA_ENTER_HELD(c);
  if (c->fused) {
    // Held back until the next inner step: with the SD halo update between
    // them (all faces external), both steps run as one pass.
    if (c->held == kHeldPpcg && c->held_halo) {
      c->held = kHeldNone;
      c->held_halo = false;
      // u += alpha p left pending by cg_calc_ur precedes both steps; the
      // pass adds it to u before their sd (the same roundings).
      const bool pend_u = c->pend_u && !c->pend_p && !c->pend_halo;
      if (!pend_u) TEA_RETURN_IF(apply_cg_pending(c));
      const bool pre = c->pre_halo == TEA_HALO_SD;
      if (!pre) TEA_RETURN_IF(apply_pre_halo(c));
      return ppcg_two_steps(c, c->held_alpha, c->held_beta, alpha, beta, pend_u,
                            pre);
    }
    // The step is held with u += alpha p and a held-back SD update before
    // it still pending (the pass that runs the step applies both first);
    // any other deferred work is applied now.
    if (c->held || c->pend_p || c->pend_halo ||
        (c->pre_halo >= 0 && c->pre_halo != TEA_HALO_SD))
      TEA_RETURN_IF(apply_pending(c));
    c->held = kHeldPpcg;
    c->held_halo = false;
    c->held_alpha = alpha;
    c->held_beta = beta;
    return TEA_HIP_OK;
  }
  TEA_RETURN_IF(apply_pending(c));
  return ppcg_step(c, alpha, beta);
}

}  // extern "C"
