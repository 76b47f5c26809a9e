// ADI half-step for the diagonally dominant regime: each tridiagonal line is
// solved by a converged Jacobi (Neumann-series) iteration over a tile of the
// line held in registers, instead of a sequential sweep per line.
//
// Every row of an ADI half-step system (src/alternatingDirectionSolver.cpp:8-134)
// has Re(diag) >= 1/h, h = dz/2: the material terms k*beta and the transparent
// boundary terms only add non-negative real parts (Im e >= 0 because Re(kd) is
// clamped at 0).  The two off-diagonals are -i/(2 k d^2), so the Jacobi
// iteration x <- D^-1 (b - O x) contracts by rho = 2|off|/|diag| <= h/(k d^2)
// per sweep.  With rho <= 5e-3 (the 2048^2 / 20 nm grid: rho = 4.2e-4) K sweeps
// give an error below rho^K relative.  For complex128 fields K is chosen so that
// rho^K < 1e-17: the result is the exact solution to FP64 rounding, like the
// reference's Thomas sweep.  Complex64 fields stop at rho^K < 1e-10, below the
// rounding of their stored values.  Less dominant systems use the sequential kernel.
//
// Tiles.  A work-group holds TL lines x TP positions; each lane holds ADI_C
// consecutive positions of one line in registers.  The K-sweep iterate at a
// position depends only on the K positions either side, so the K positions at
// each tile edge are a halo, recomputed by the neighbouring tile, and every
// other position comes out exactly as from an untiled sweep.  A sweep updates
// a lane's positions from its own registers; only the first and last positions
// cross lanes, through LDS slots indexed by lane (conflict-free).
//  - y-implicit lines are contiguous in memory: one wavefront per line tile
//    (TL = 4, TP = 256), so the exchange needs no work-group barrier.
//  - x-implicit lines run across memory (stride N): 16 lines side by side make
//    every load 16 consecutive elements (TL = 16, TP = 128, a barrier per sweep).
// Work-groups of neighbouring line groups run on one XCD, where the explicit
// direction's neighbour lines they share are still in that XCD's L2.
#include <type_traits>

#include "pxg_tridiag.hpp"

namespace pxg {

constexpr int ADI_C = 4;            // consecutive positions per lane
constexpr int ADI_HALO_MAX = 8;     // largest sweep count (and halo) of the tile path
constexpr int ADI_MASK_WORDS = 16;

// A tile of TL_ lines x TP_ positions.  The shapes built (index = the plan's tile
// choice, pxg_plan_set_adi_tiles): x-implicit (16, 128) (8, 256) (16, 256)
// (32, 64) (32, 128) (8, 128); y-implicit (4, 256) (1, 256) (2, 256) (8, 256)
// (2, 512) (1, 2048).  Every shape gives the same bits (each position sees the
// same K-sweep iterate).  A tile as long as the line or longer covers it whole:
// it needs no halo, and its lines are not cut into segments.
template <bool IMPX, int TL_, int TP_>
struct AdiTile {
  static constexpr int TL = TL_;              // lines per tile
  static constexpr int TP = TP_;              // positions per line
  static constexpr int G = TP / ADI_C;        // lanes per line
  static constexpr int LANES = TL * G;
  static constexpr int DG = IMPX ? TL : 1;    // lane distance of the next position group
  static constexpr bool WAVE_LOCAL = !IMPX && G == 64;  // a line tile is one wavefront
  static_assert(LANES <= 1024 && LANES % 64 == 0, "work-group of whole wavefronts");
};

// Complex product and squared modulus with every rounding spelled out (one
// product rounded, then one fused multiply-add per real part), so that every
// instantiation of the kernel (each tile shape) gives the same bits: left to
// the compiler, a*b + c*d may fuse either product.
__device__ __forceinline__ cx<double> cmul_fixed(cx<double> a, cx<double> b)
{
#pragma clang fp contract(off)
  return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
}

__device__ __forceinline__ double norm2_fixed(cx<double> a)
{
#pragma clang fp contract(off)
  return fma(a.x, a.x, a.y * a.y);
}

// Transparent boundary at the ends of the implicit line: diag correction from
// the cell's value and its inward neighbour along the line.
__device__ __forceinline__ cx<double> tbc_diag(cx<double> diag, cx<double> u, cx<double> uin, double dimp)
{
  cx<double> e;
  if (tbc_exp_ikd(u, uin, 1e-6, e)) diag = diag - cx<double>{-0.5 * e.y / dimp, 0.5 * e.x / dimp};
  return diag;
}

// Transparent boundary at the first and last line (explicit direction): rhs
// correction from the cell's value and its neighbour in the next line inward.
__device__ __forceinline__ cx<double> tbc_rhs(cx<double> rhs, cx<double> u, cx<double> uin, double dexp)
{
  cx<double> e;
  if (tbc_exp_ikd(u, uin, 1e-6, e)) {
    const cx<double> q = cmul_fixed(cx<double>{-0.5 * e.y, 0.5 * e.x}, u);
    rhs = rhs + cx<double>{q.x / dexp, q.y / dexp};
  }
  return rhs;
}

template <class R>
__device__ __forceinline__ cx<double> ldd(const cx<R> *f, size_t i)
{
  return cvt<double>(f[i]);
}

// Launch constants of one half-step (one field): computed once on the host.
struct AdiCoef {
  double k;           // wavenumber
  double z;           // height of the material evaluation
  double ih;          // 1/h, h = dz/2
  double dimp, dexp;  // k d^2 of the implicit and explicit directions
  double idimp, iexp; // their inverses
  double hexp;        // 0.5 / dexp
  double o;           // off-diagonal i*o, o = -0.5 / dimp
  double ddy;         // k dy^2 (y-implicit diagonal of the reference)
};

// Linear work-group index -> tile index such that the work-groups the hardware
// places on one XCD (index mod 8) take a contiguous range of tiles.
__device__ __forceinline__ int xcd_tile(int b, int total)
{
  const int q = total >> 3, r = total & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <class R, bool IMPX, int TL, int TP>
__global__ __launch_bounds__((AdiTile<IMPX, TL, TP>::LANES)) void k_adi_jacobi(TriArgs<R> a, AdiCoef c,
                                                                            const cx<R> *__restrict__ prev,
                                                                            cx<R> *__restrict__ next, int iters,
                                                                            int halo, int nseg)
{
#pragma clang fp contract(off)  // fused operations below are written out
  using T = AdiTile<IMPX, TL, TP>;
  constexpr int C = ADI_C;
  __shared__ cx<double> edge[2][2][T::LANES];  // [buffer][first, last position][lane]
  __shared__ uint64_t mask[ADI_MASK_WORDS];
  const int N = a.n;
  const int tid = threadIdx.x;
  const int l = IMPX ? tid % T::TL : tid / T::G;
  const int g = IMPX ? tid / T::TL : tid % T::G;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int seg = tile % nseg, grp = tile / nseg;
  const int j0 = seg * (T::TP - 2 * halo) - halo;  // position of tile slot 0
  const int jb = j0 + g * C;                        // this lane's first position
  const int line = grp * T::TL + l;
  const double k = c.k, z = c.z;
  const unsigned st = a.first_step;
  const int nw = a.mat.mode == MAT_CSG ? (a.mat.n_parts + 63) >> 6 : 0;
  if (a.mat.mode == MAT_CSG) {
    const int l0 = grp * T::TL, l1 = min(N - 1, l0 + T::TL - 1);
    const int p0 = max(0, j0), p1 = min(N - 1, j0 + T::TP - 1);
    if (IMPX)
      csg_cull(a.mat, mask, nw, coord(a.xmin, p0, a.dx), coord(a.xmin, p1, a.dx), coord(a.ymin, l0, a.dy),
               coord(a.ymin, l1, a.dy), z, z);
    else
      csg_cull(a.mat, mask, nw, coord(a.xmin, l0, a.dx), coord(a.xmin, l1, a.dx), coord(a.ymin, p0, a.dy),
               coord(a.ymin, p1, a.dy), z, z);
  }
  // Clamped cell indices: every load below stays inside the field, and cells
  // outside the grid (dead) are masked when the coefficients are formed.  The
  // field loads follow the material evaluation: issued before it, they held
  // 24-48 registers through the CSG walk and cut occupancy to 3-4 waves.
  const int ln = min(line, N - 1), lm = max(ln - 1, 0), lp = min(ln + 1, N - 1);
  auto at = [&](int jj, int li) -> size_t {
    return IMPX ? static_cast<size_t>(jj) * N + li : static_cast<size_t>(li) * N + jj;
  };
  bool live[C];
  int jc[C];
#pragma unroll
  for (int m = 0; m < C; ++m) {
    live[m] = line < N && jb + m >= 0 && jb + m < N;
    jc[m] = min(max(jb + m, 0), N - 1);
  }
  // material of the lane's cells
  double dd[C], bb[C];
#pragma unroll
  for (int m = 0; m < C; ++m) dd[m] = bb[m] = 0.0;
  if (a.mat.mode == MAT_CSG) {
    double xs[C], ys[C];
#pragma unroll
    for (int m = 0; m < C; ++m) {
      xs[m] = IMPX ? coord(a.xmin, jc[m], a.dx) : coord(a.xmin, ln, a.dx);
      ys[m] = IMPX ? coord(a.ymin, ln, a.dy) : coord(a.ymin, jc[m], a.dy);
    }
    int e[C], eb[C];
    csg_eval_x2<C, false>(a.mat, mask, nw, xs, ys, z, z, e, eb);
#pragma unroll
    for (int m = 0; m < C; ++m) part_values(a.mat, e[m], dd[m], bb[m]);
  } else if (a.mat.mode == MAT_PLANES) {
    const int pi = a.mat.z_invariant ? 0 : static_cast<int>(2 * st + (IMPX ? 0 : 1)) - a.mat.first_plane;
    const double *pl = a.mat.planes + static_cast<size_t>(pi) * a.mat.plane_stride;
#pragma unroll
    for (int m = 0; m < C; ++m) {
      const size_t idx = at(jc[m], ln);
      dd[m] = pl[2 * idx];
      bb[m] = pl[2 * idx + 1];
    }
  }

  cx<R> ru[C], rm[C], rp[C];
#pragma unroll
  for (int m = 0; m < C; ++m) {
    ru[m] = prev[at(jc[m], ln)];
    rm[m] = prev[at(jc[m], lm)];
    rp[m] = prev[at(jc[m], lp)];
  }

  // Rows of the lane's cells: diag and rhs (src/alternatingDirectionSolver.cpp:8-134)
  cx<double> dg[C], rh[C];
#pragma unroll
  for (int m = 0; m < C; ++m) {
    if (IMPX || (a.quirks & PXG_FIX_ADI_Y_DIAG)) {
      dg[m] = {fma(bb[m], k, c.ih), fma(dd[m], k, c.idimp)};
    } else {
      // 1/h + i / (k dy^2 + k beta + i k delta)
      const double p = fma(bb[m], k, c.ddy), q = dd[m] * k, r = recip(fma(p, p, q * q));
      dg[m] = {fma(q, r, c.ih), p * r};
    }
    const cx<double> u = cvt<double>(ru[m]);
    cx<double> un{0.0, 0.0};
    if (line > 0) un = cvt<double>(rm[m]);
    if (line < N - 1) un = un + cvt<double>(rp[m]);
    const cx<double> t = cmul_fixed(cx<double>{c.ih, -c.iexp}, u);
    rh[m] = {fma(-un.y, c.hexp, t.x), fma(un.x, c.hexp, t.y)};
  }
  // Transparent boundary rows, after the loaded values are consumed: the few
  // lanes holding boundary cells read their neighbours again.
  // (a work-group-uniform test first: only tiles at the grid's edges hold such cells)
  const bool edge_tile = j0 <= 0 || j0 + T::TP >= N || grp == 0 || (grp + 1) * T::TL >= N;
  if (a.use_tbc && edge_tile) {
#pragma unroll
    for (int m = 0; m < C; ++m) {
      const int j = jb + m;
      if (!live[m] || (j != 0 && j != N - 1 && line != 0 && line != N - 1)) continue;
      const cx<double> u = ldd(prev, at(j, ln));
      if (j == N - 1) dg[m] = tbc_diag(dg[m], u, ldd(prev, at(N - 2, ln)), c.dimp);
      if (j == 0) dg[m] = tbc_diag(dg[m], u, ldd(prev, at(1, ln)), c.dimp);
      if (line == N - 1) rh[m] = tbc_rhs(rh[m], u, ldd(prev, at(j, lm)), c.dexp);
      if (line == 0) rh[m] = tbc_rhs(rh[m], u, ldd(prev, at(j, lp)), c.dexp);
    }
  }
  // x0 = D^-1 b and the sweep factor w = D^-1 (i o): x <- x0 - w (x_left + x_right);
  // dead cells (outside the grid) stay exactly zero: x0 = 0 and w = 0
  cx<double> x0[C], w[C], x[C];
#pragma unroll
  for (int m = 0; m < C; ++m) {
    const double r = recip(norm2_fixed(dg[m]));
    const double lr = live[m] ? r : 0.0, lo = live[m] ? c.o * r : 0.0;
    const cx<double> inv = {dg[m].x * lr, -dg[m].y * lr};
    x0[m] = x[m] = cmul_fixed(rh[m], inv);
    w[m] = {dg[m].y * lo, dg[m].x * lo};  // i o conj(diag) / |diag|^2
  }

  // One Jacobi sweep; B selects the LDS buffer (alternating buffers need one
  // sync per sweep: a lane cannot rewrite a buffer before every lane has
  // passed the sync that follows the reads of it).
  auto sweep = [&](auto bsel) {
    constexpr int B = decltype(bsel)::value;
    edge[B][0][tid] = x[0];
    edge[B][1][tid] = x[C - 1];
    lds_sync<T::WAVE_LOCAL>();
    const cx<double> xl = g > 0 ? edge[B][1][tid - T::DG] : cx<double>{0.0, 0.0};
    const cx<double> xr = g < T::G - 1 ? edge[B][0][tid + T::DG] : cx<double>{0.0, 0.0};
    cx<double> left = xl;
#pragma unroll
    for (int m = 0; m < C; ++m) {
      const cx<double> right = m < C - 1 ? x[m + 1] : xr;
      const cx<double> sum = left + right;
      left = x[m];
      // x0 - w sum, as two fused steps per component
      x[m].x = fma(-w[m].x, sum.x, fma(w[m].y, sum.y, x0[m].x));
      x[m].y = fma(-w[m].x, sum.y, fma(-w[m].y, sum.x, x0[m].y));
    }
  };
  for (int it = 0; it + 1 < iters; it += 2) {
    sweep(std::integral_constant<int, 0>{});
    sweep(std::integral_constant<int, 1>{});
  }
  if (iters & 1) sweep(std::integral_constant<int, 0>{});
#pragma unroll
  for (int m = 0; m < C; ++m) {
    const int t = g * C + m;  // tile slot
    if (live[m] && t >= halo && t < T::TP - halo) next[at(jb + m, line)] = cvt<R>(x[m]);
  }
}

// Sweeps needed for rho^K < tol, or 0 when the system is not dominant enough
// for the tile halo (the caller then uses the sequential kernel).
int adi_jacobi_iters(double dz, double k, double d, double tol)
{
  const double rho = (dz / 2.0) / (k * d * d);
  if (!(rho > 0.0) || rho > 5e-3) return 0;
  int K = 1;
  double e = rho;
  while (e > tol && K <= ADI_HALO_MAX) {
    e *= rho;
    ++K;
  }
  return K <= ADI_HALO_MAX ? K : 0;
}

template <class R, bool IMPX, int TL, int TP>
static hipError_t launch_tiles(const TriArgs<R> &a, const cx<R> *prev, cx<R> *next, int iters, hipStream_t s)
{
  using T = AdiTile<IMPX, TL, TP>;
  // a tile that covers the whole line needs no halo: nothing lies beyond its ends
  const int halo = T::TP >= a.n ? 0 : iters;
  const int stride = T::TP - 2 * halo;
  const int nseg = (a.n + stride - 1) / stride;
  const int ngrp = (a.n + T::TL - 1) / T::TL;
  // the same expressions as the reference's coefficients, once per launch
  AdiCoef c;
  c.k = a.khost[0];
  const double h = a.dz / 2.0;
  c.z = IMPX ? coord(a.zmin, a.first_step, a.dz) : coord(a.zmin, a.first_step, a.dz) + h;
  const double ddx = c.k * a.dx * a.dx, ddy = c.k * a.dy * a.dy;
  c.dimp = IMPX ? ddx : ddy;
  c.dexp = IMPX ? ddy : ddx;
  c.ih = 1.0 / h;
  c.idimp = 1.0 / c.dimp;
  c.iexp = 1.0 / c.dexp;
  c.hexp = 0.5 / c.dexp;
  c.o = -0.5 / c.dimp;
  c.ddy = ddy;
  hipLaunchKernelGGL((k_adi_jacobi<R, IMPX, TL, TP>), dim3(nseg * ngrp), dim3(T::LANES), 0, s, a, c, prev, next,
                     iters, halo, nseg);
  return hipGetLastError();
}

int adi_tile_count() { return 6; }

template <class R>
hipError_t launch_adi_jacobi(const TriArgs<R> &a, bool implicit_x, const cx<R> *prev, cx<R> *next,
                             int iters, hipStream_t s)
{
  if (iters < 1 || iters > ADI_HALO_MAX) return hipErrorInvalidValue;
  if (implicit_x) {
    switch (a.adi_tile) {
      case 0: return launch_tiles<R, true, 16, 128>(a, prev, next, iters, s);
      case 1: return launch_tiles<R, true, 8, 256>(a, prev, next, iters, s);
      case 2: return launch_tiles<R, true, 16, 256>(a, prev, next, iters, s);
      case 3: return launch_tiles<R, true, 32, 64>(a, prev, next, iters, s);
      case 4: return launch_tiles<R, true, 32, 128>(a, prev, next, iters, s);
      case 5: return launch_tiles<R, true, 8, 128>(a, prev, next, iters, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (a.adi_tile) {
    case 0: return launch_tiles<R, false, 4, 256>(a, prev, next, iters, s);
    case 1: return launch_tiles<R, false, 1, 256>(a, prev, next, iters, s);
    case 2: return launch_tiles<R, false, 2, 256>(a, prev, next, iters, s);
    case 3: return launch_tiles<R, false, 8, 256>(a, prev, next, iters, s);
    case 4: return launch_tiles<R, false, 2, 512>(a, prev, next, iters, s);
    case 5: return launch_tiles<R, false, 1, 2048>(a, prev, next, iters, s);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_adi_jacobi<float>(const TriArgs<float> &, bool, const cx<float> *, cx<float> *, int,
                                             hipStream_t);
template hipError_t launch_adi_jacobi<double>(const TriArgs<double> &, bool, const cx<double> *, cx<double> *,
                                              int, hipStream_t);

}  // namespace pxg
