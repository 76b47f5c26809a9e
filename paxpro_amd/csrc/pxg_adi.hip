// ADI half-step for the diagonally dominant regime: each tridiagonal line is
// solved by a converged Jacobi (Neumann-series) iteration inside a tile held in
// LDS, instead of a sequential sweep per line.
//
// Every row of an ADI half-step system (src/alternatingDirectionSolver.cpp:8-134)
// has Re(diag) >= 1/h, h = dz/2: the material terms k*beta and the transparent
// boundary terms only add non-negative real parts (Im e >= 0 because Re(kd) is
// clamped at 0).  The two off-diagonals are -i/(2 k d^2), so the Jacobi
// iteration x <- D^-1 (b - O x) contracts by rho = 2|off|/|diag| <= h/(k d^2)
// per sweep.  With rho <= 5e-3 (config 2: rho = 4.2e-4) K sweeps give an error
// below rho^K relative, and K is chosen so that rho^K < 1e-17: the result is the
// exact solution to FP64 rounding, like the reference's Thomas sweep.  A tile
// covers 8 lines x 128 positions; the 8 positions at each tile edge are halo,
// recomputed by the neighbouring tile, and K <= 8 keeps truncation at the tile
// edge below the same bound.  Less dominant systems use the sequential kernel.
//
// Lanes map to cells so that global accesses coalesce: along y (contiguous) in
// the y-implicit pass, across the 8 consecutive lines (rows) in the x-implicit
// pass.
#include "pxg_tridiag.hpp"

namespace pxg {

constexpr int ADI_LINES = 8;
constexpr int ADI_TILE = 128;  // positions per line held by a work-group
constexpr int ADI_HALO = 8;    // recomputed positions at each tile edge
constexpr int ADI_MASK_WORDS = 16;

__device__ __forceinline__ bool tbc_e(cx<double> u_out, cx<double> u_in, cx<double> &e)
{
  if (!(sqrt(norm2(u_in)) > 1e-6)) return false;
  const cx<double> ratio = cdiv(u_out, u_in);
  double kr = atan2(ratio.y, ratio.x);
  const double amp = hypot(ratio.x, ratio.y);
  if (kr < 0.0) kr = 0.0;
  double s, c;
  sincos(kr, &s, &c);
  e = {amp * c, amp * s};
  return true;
}

template <class R>
__device__ __forceinline__ cx<double> ldd(const cx<R> *f, size_t i)
{
  return cvt<double>(f[i]);
}

template <class R, bool IMPX>
__global__ __launch_bounds__(ADI_LINES *ADI_TILE) void k_adi_jacobi(TriArgs<R> a, const cx<R> *__restrict__ prev,
                                                                   cx<R> *__restrict__ next, int iters)
{
  __shared__ cx<double> xs[2][ADI_LINES][ADI_TILE];
  __shared__ uint64_t mask[ADI_MASK_WORDS];
  const int N = a.n;
  const int tid = threadIdx.x;
  // lane -> (line l, tile position t); coalesced direction fastest
  const int l = IMPX ? tid % ADI_LINES : tid / ADI_TILE;
  const int t = IMPX ? tid / ADI_LINES : tid % ADI_TILE;
  const int line = blockIdx.y * ADI_LINES + l;
  const int j0 = blockIdx.x * (ADI_TILE - 2 * ADI_HALO) - ADI_HALO;
  const int j = j0 + t;                      // position along the line
  const bool live = line < N && j >= 0 && j < N;
  const double k = a.kvec[0];
  const double h = a.dz / 2.0;
  const unsigned st = a.first_step;
  const double z = IMPX ? coord(a.zmin, st, a.dz) : coord(a.zmin, st, a.dz) + h;
  const double ddx = k * a.dx * a.dx, ddy = k * a.dy * a.dy;
  const double dimp = IMPX ? ddx : ddy, dexp = IMPX ? ddy : ddx;
  const cx<double> off = {0.0, -0.5 / dimp};
  const int nw = a.mat.mode == MAT_CSG ? (a.mat.n_parts + 63) >> 6 : 0;
  if (a.mat.mode == MAT_CSG) {
    const int l0 = blockIdx.y * ADI_LINES, l1 = min(N - 1, l0 + ADI_LINES - 1);
    const int p0 = max(0, j0), p1 = min(N - 1, j0 + ADI_TILE - 1);
    if (IMPX)
      csg_cull(a.mat, mask, nw, coord(a.xmin, p0, a.dx), coord(a.xmin, p1, a.dx), coord(a.ymin, l0, a.dy),
               coord(a.ymin, l1, a.dy), z, z);
    else
      csg_cull(a.mat, mask, nw, coord(a.xmin, l0, a.dx), coord(a.xmin, l1, a.dx), coord(a.ymin, p0, a.dy),
               coord(a.ymin, p1, a.dy), z, z);
  }
  // element index of (line + dl, position jj)
  auto at = [&](int jj, int dl) -> size_t {
    const int ln = line + dl;
    return IMPX ? static_cast<size_t>(jj) * N + ln : static_cast<size_t>(ln) * N + jj;
  };
  cx<double> inv{0.0, 0.0}, rhs{0.0, 0.0}, x{0.0, 0.0};
  if (live) {
    double d = 0.0, be = 0.0;
    if (a.mat.mode == MAT_CSG) {
      const double xx = IMPX ? coord(a.xmin, j, a.dx) : coord(a.xmin, line, a.dx);
      const double yy = IMPX ? coord(a.ymin, line, a.dy) : coord(a.ymin, j, a.dy);
      part_values(a.mat, csg_eval(a.mat, mask, nw, xx, yy, z), d, be);
    } else if (a.mat.mode == MAT_PLANES) {
      const int pi = a.mat.z_invariant ? 0 : static_cast<int>(2 * st + (IMPX ? 0 : 1)) - a.mat.first_plane;
      const double *pl = a.mat.planes + static_cast<size_t>(pi) * a.mat.plane_stride;
      const size_t idx = at(j, 0);
      d = pl[2 * idx];
      be = pl[2 * idx + 1];
    }
    cx<double> diag;
    if (IMPX || (a.quirks & PXG_FIX_ADI_Y_DIAG))
      diag = {1.0 / h + be * k, 1.0 / dimp + d * k};
    else
      diag = cx<double>{1.0 / h, 0.0} + cdiv(cx<double>{0.0, 1.0}, cx<double>{ddy + be * k, d * k});
    const cx<double> u = ldd(prev, at(j, 0));
    rhs = cx<double>{1.0 / h, -1.0 / dexp} * u;
    if (line > 0) {
      const cx<double> un = ldd(prev, at(j, -1));
      rhs = rhs + cx<double>{-0.5 * un.y / dexp, 0.5 * un.x / dexp};
    }
    if (line < N - 1) {
      const cx<double> un = ldd(prev, at(j, +1));
      rhs = rhs + cx<double>{-0.5 * un.y / dexp, 0.5 * un.x / dexp};
    }
    if (a.use_tbc) {
      cx<double> e;
      if (j == N - 1 && tbc_e(u, ldd(prev, at(N - 2, 0)), e))
        diag = diag - cx<double>{-0.5 * e.y / dimp, 0.5 * e.x / dimp};
      if (j == 0 && tbc_e(u, ldd(prev, at(1, 0)), e))
        diag = diag - cx<double>{-0.5 * e.y / dimp, 0.5 * e.x / dimp};
      if (line == N - 1 && tbc_e(u, ldd(prev, at(j, -1)), e)) {
        const cx<double> q = cx<double>{-0.5 * e.y, 0.5 * e.x} * u;
        rhs = rhs + cx<double>{q.x / dexp, q.y / dexp};
      }
      if (line == 0 && tbc_e(u, ldd(prev, at(j, +1)), e)) {
        const cx<double> q = cx<double>{-0.5 * e.y, 0.5 * e.x} * u;
        rhs = rhs + cx<double>{q.x / dexp, q.y / dexp};
      }
    }
    inv = cdiv(cx<double>{1.0, 0.0}, diag);
    x = rhs * inv;
  }
  xs[0][l][t] = x;
  __syncthreads();
  int cur = 0;
  for (int it = 0; it < iters; ++it) {
    cx<double> xl{0.0, 0.0}, xr{0.0, 0.0};
    if (t > 0 && j - 1 >= 0) xl = xs[cur][l][t - 1];
    if (t < ADI_TILE - 1 && j + 1 < N) xr = xs[cur][l][t + 1];
    const cx<double> xn = live ? (rhs - off * (xl + xr)) * inv : cx<double>{0.0, 0.0};
    xs[cur ^ 1][l][t] = xn;
    cur ^= 1;
    __syncthreads();
  }
  if (live && t >= ADI_HALO && t < ADI_TILE - ADI_HALO) next[at(j, 0)] = cvt<R>(xs[cur][l][t]);
}

// Sweeps needed for rho^K < 1e-17, or 0 when the system is not dominant
// enough for the tile halo (the caller then uses the sequential kernel).
int adi_jacobi_iters(double dz, double k, double d)
{
  const double rho = (dz / 2.0) / (k * d * d);
  if (!(rho > 0.0) || rho > 5e-3) return 0;
  int K = 1;
  double e = rho;
  while (e > 1e-17 && K <= ADI_HALO) {
    e *= rho;
    ++K;
  }
  return K <= ADI_HALO ? K : 0;
}

template <class R>
hipError_t launch_adi_jacobi(const TriArgs<R> &a, bool implicit_x, const cx<R> *prev, cx<R> *next,
                             int iters, hipStream_t s)
{
  const int interior = ADI_TILE - 2 * ADI_HALO;
  const dim3 grid((a.n + interior - 1) / interior, (a.n + ADI_LINES - 1) / ADI_LINES);
  if (implicit_x)
    hipLaunchKernelGGL((k_adi_jacobi<R, true>), grid, dim3(ADI_LINES * ADI_TILE), 0, s, a, prev, next, iters);
  else
    hipLaunchKernelGGL((k_adi_jacobi<R, false>), grid, dim3(ADI_LINES * ADI_TILE), 0, s, a, prev, next, iters);
  return hipGetLastError();
}

template hipError_t launch_adi_jacobi<float>(const TriArgs<float> &, bool, const cx<float> *, cx<float> *, int,
                                             hipStream_t);
template hipError_t launch_adi_jacobi<double>(const TriArgs<double> &, bool, const cx<double> *, cx<double> *,
                                              int, hipStream_t);

}  // namespace pxg
