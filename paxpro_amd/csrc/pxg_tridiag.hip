// Tridiagonal kernels: the Crank-Nicolson march (CrankNicholson::solveStep,
// src/crancNicholson.cpp:17-148), the ADI half-steps (src/alternatingDirectionSolver.cpp)
// and the batched form of ThomasAlgorithm::solve (include/thomasAlgorithm.tpp:2-32).
//
// CN: one work-group owns one field for the whole run (a persistent march over
// all z-steps, no launch per step).  Each step builds the complex tridiagonal
// system row by row in FP64 and solves it with a partitioned method: P lanes
// each eliminate a contiguous chunk of about n/P rows so that every interior unknown is
// expressed through the chunk's first and last unknowns; those 2P unknowns form
// a tridiagonal system solved by parallel cyclic reduction in LDS; the lanes
// then recover their interiors.  The arithmetic is Gaussian elimination without
// pivoting, as in the reference's Thomas sweep, reordered for parallelism.
//
// ADI: one lane per tridiagonal line.  The x-implicit pass walks x-lines whose
// consecutive lines are consecutive in memory (one lane per system, interleaved
// layout, coalesced); the y-implicit pass walks contiguous y-lines.  The global
// Nx*Ny system of the reference has zero couplings between lines
// (src/alternatingDirectionSolver.cpp:49-52,106-109), so per-line sweeps are the
// same elimination.
#include "pxg_tridiag.hpp"

namespace pxg {

constexpr int CN_MAX_P = 512;
constexpr int CN_REG_ROWS = 4;  // chunk rows held in registers by k_cn_reg
constexpr int TRI_MASK_WORDS = 16;

// P lanes own balanced chunks: the first n mod P chunks have n/P + 1 rows,
// the others n/P.  P is the smallest power of two that brings the chunks to
// at most 4 rows (the register-resident march), capped at CN_MAX_P, and at
// most n/2 so that every chunk has distinct first and last rows.  m is the
// largest chunk.  (Rounding P down instead left most sizes with chunks of 5-8
// rows, outside the register march: 12.4 us/step at 584 cells against 8.1 at
// 1024.)
CnPartition cn_partition(int n)
{
  int P = 1;
  while (P < CN_MAX_P && P * 2 <= n / 2 && (n + P - 1) / P > CN_REG_ROWS) P *= 2;
  return {P, (n + P - 1) / P};
}

__host__ __device__ static inline size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

static size_t cn_field_scratch(int n, int esz, const CnPartition &cp)
{
  return align256(static_cast<size_t>(n) * esz) + 3 * align256(static_cast<size_t>(cp.m) * cp.P * 16);
}

size_t tridiag_scratch_bytes(const pxg_problem &p, bool mixed)
{
  const int esz = mixed ? 8 : 16;
  if (p.kind == PXG_CN) return align256(cn_field_scratch(p.nx, esz, cn_partition(p.nx))) * p.batch;
  if (p.kind == PXG_ADI) {
    const size_t cells = static_cast<size_t>(p.nx) * p.nx;
    return (align256(cells * esz) + 2 * align256(cells * 16)) * p.batch;
  }
  return 0;
}

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ cx<double> cmul(cx<double> a, cx<double> b) { return a * b; }

// Transparent-boundary factor of the CN and sequential ADI rows: the closed
// form tbc_exp_ikd (no logarithm or exponential).
__device__ __forceinline__ bool tbc_factor(cx<double> u_out, cx<double> u_in, double thresh, cx<double> &e)
{
  return tbc_exp_ikd(u_out, u_in, thresh, e);
}

__device__ __forceinline__ void material_2d(const MatDev &m, const uint64_t *mask, int nw,
                                            unsigned plane, int i, double x, double z, double &d,
                                            double &b)
{
  if (m.mode == MAT_CSG) {
    part_values(m, csg_eval(m, mask, nw, x, 0.0, z), d, b);
  } else if (m.mode == MAT_CURVED) {
    curved_material(m.wg, x, z, d, b);
  } else if (m.mode == MAT_PLANES) {
    const int p = m.z_invariant ? 0 : static_cast<int>(plane) - m.first_plane;
    const double *pl = m.planes + static_cast<size_t>(p) * m.plane_stride;
    d = pl[2 * i];
    b = pl[2 * i + 1];
  } else {
    d = 0.0;
    b = 0.0;
  }
}

template <class R>
__device__ __forceinline__ cx<double> ld(const cx<R> *f, int i)
{
  return cvt<double>(f[i]);
}

// Row i of the CN system with the default ParaxialEquation (F=G=H=1, J=0):
// diag = 1 + i rho/2 + (r/2)(beta + i delta), off-diagonals -i rho/4,
// rhs = i rho/4 (u_{i-1} + u_{i+1}) - i rho/2 u_i + (1 - (r/2)(beta0 + i delta0)) u_i,
// transparent boundary rows at i = 0 and N-1 (src/crancNicholson.cpp:30-89,107-148).
// ul, u, ur are the previous field at i-1, i, i+1 (zero outside); uin is the
// inner neighbour of a boundary row (i = 1 or N-2), used only at i = 0, N-1.
// The row from the material of the step (d1, b1) and of the previous step (d0, b0).
__device__ __forceinline__ void cn_row_m(int N, int bc, int i, double rho, double r, double d1, double b1,
                                         double d0, double b0, cx<double> ul, cx<double> u, cx<double> ur,
                                         cx<double> uin, cx<double> &ai, cx<double> &bi, cx<double> &ci,
                                         cx<double> &di)
{
  const cx<double> off = {0.0, -0.25 * rho};
  ai = i > 0 ? off : cx<double>{0.0, 0.0};
  ci = i < N - 1 ? off : cx<double>{0.0, 0.0};
  bi = {1.0 + 0.5 * (b1 * r), 0.5 * rho + 0.5 * (d1 * r)};
  cx<double> v = (ul + ur) * cx<double>{0.0, 0.25 * rho};
  v = v - cx<double>{-0.5 * rho * u.y, 0.5 * rho * u.x};
  v = v + cx<double>{1.0 - 0.5 * (b0 * r), -0.5 * (d0 * r)} * u;
  di = v;
  if (bc == PXG_BC_TRANSPARENT && N >= 2 && (i == 0 || i == N - 1)) {
    cx<double> e;
    if (tbc_factor(u, uin, 1e-16, e)) {
      const cx<double> q = cx<double>{0.0, 0.25 * rho} * e;
      bi = bi - q;
      di = di + q * u;
    }
  }
}

// Row i with a ParaxialEquation given as host-sampled planes (CnJob.eq): the
// terms of src/crancNicholson.cpp:41-88 in the reference's order, e1 the values
// at z, e0 at z - dz.  The off-diagonals are the reference's symmetric sub
// diagonal, sub[i] = -i rho/4 H-(i) G(i) coupling rows i and i+1; the boundary
// rows take H-, G (and H-', G') of row N-1, the last row of the reference's
// loop (:146-147).
__device__ __forceinline__ void cn_row_eq(int N, int bc, int i, double rho, double r, double d1, double b1,
                                          double d0, double b0, const double *e1, const double *e0,
                                          cx<double> ul, cx<double> u, cx<double> ur, cx<double> uin,
                                          cx<double> &ai, cx<double> &bi, cx<double> &ci, cx<double> &di)
{
  const double *q = e1 + 5 * i, *qp = e0 + 5 * i;
  const double fv = q[0], gv = q[1], hp = q[2], hm = q[3], jv = q[4];
  const double fp = qp[0], gp = qp[1], hpp = qp[2], hmp = qp[3], jp = qp[4];
  const double delta = d1 - jv, deltaP = d0 - jp;
  auto sub = [&](int row) { return cx<double>{0.0, ((-0.25 * rho) * e1[5 * row + 3]) * e1[5 * row + 1]}; };
  ai = i > 0 ? sub(i - 1) : cx<double>{0.0, 0.0};
  ci = i < N - 1 ? sub(i) : cx<double>{0.0, 0.0};
  // diag = F + 0.25 (H+ + H-) G i rho + 0.5 (beta r + i delta r)
  bi = cx<double>{fv * 1.0, 0.0} + cx<double>{0.0, 0.25 * (hp + hm) * gv * rho} +
       cx<double>{0.5 * (b1 * r), 0.5 * (delta * r)};
  // rhs = (u_{i-1} H-' G' + u_{i+1} H+' G') i rho/4 - u i rho/4 (H+' + H-') G
  //       + (F' - 0.5 (beta' r + i delta' r)) u
  cx<double> v = scale(scale(ul, hmp), gp);
  v = v + scale(scale(ur, hpp), gp);
  v = v * cx<double>{0.0, 0.25 * rho};
  v = v - u * cx<double>{0.0, 0.25 * rho * (hpp + hmp) * gv};
  v = v + cx<double>{1.0 * fp - 0.5 * (b0 * r), -0.5 * (deltaP * r)} * u;
  di = v;
  if (bc == PXG_BC_TRANSPARENT && N >= 2 && (i == 0 || i == N - 1)) {
    cx<double> e;
    if (tbc_factor(u, uin, 1e-16, e)) {
      const double *ql = e1 + 5 * (N - 1), *qlp = e0 + 5 * (N - 1);
      bi = bi - cx<double>{0.0, 0.25 * rho * ql[3] * ql[1]} * e;
      di = di + cx<double>{0.0, 0.25 * rho * qlp[3] * qlp[1]} * e * u;
    }
  }
}

__device__ __forceinline__ void cn_row_v(const CnJob &J, int i, unsigned st, double z, double rho, double r,
                                         const uint64_t *mask, int nw, cx<double> ul, cx<double> u,
                                         cx<double> ur, cx<double> uin, cx<double> &ai, cx<double> &bi,
                                         cx<double> &ci, cx<double> &di)
{
  const double x = coord(J.xmin, i, J.dx);
  double d1, b1, d0, b0;
  material_2d(J.mat, mask, nw, st, i, x, z, d1, b1);
  material_2d(J.mat, mask, nw, st - 1, i, x, z - J.dz, d0, b0);
  if (J.eq) {
    const size_t stride = 5 * static_cast<size_t>(J.nx);
    const int p1 = J.eq_inv ? 0 : static_cast<int>(st) - J.eq_first;
    const int p0 = J.eq_inv ? 0 : static_cast<int>(st) - 1 - J.eq_first;
    cn_row_eq(J.nx, J.bc, i, rho, r, d1, b1, d0, b0, J.eq + p1 * stride, J.eq + p0 * stride, ul, u, ur, uin, ai,
              bi, ci, di);
    return;
  }
  cn_row_m(J.nx, J.bc, i, rho, r, d1, b1, d0, b0, ul, u, ur, uin, ai, bi, ci, di);
}

template <class R>
__device__ __forceinline__ void cn_row(const CnJob &J, const cx<R> *prev, int i, unsigned st,
                                       double z, double rho, double r, const uint64_t *mask, int nw,
                                       cx<double> &ai, cx<double> &bi, cx<double> &ci,
                                       cx<double> &di)
{
  const int N = J.nx;
  const cx<double> u = ld(prev, i);
  const cx<double> ul = i > 0 ? ld(prev, i - 1) : cx<double>{0.0, 0.0};
  const cx<double> ur = i < N - 1 ? ld(prev, i + 1) : cx<double>{0.0, 0.0};
  const cx<double> uin = (i == 0 && N >= 2) ? ld(prev, 1) : ((i == N - 1 && N >= 2) ? ld(prev, N - 2) : u);
  cn_row_v(J, i, st, z, rho, r, mask, nw, ul, u, ur, uin, ai, bi, ci, di);
}

// Rows inside the absorbing layer of npix cells at either edge.
__device__ __forceinline__ bool absorbing(int i, int N, unsigned npix)
{
  return i < static_cast<int>(npix) || N - 1 - i < static_cast<int>(npix);
}

__device__ __noinline__ double absorb_factor(int i, int N, unsigned npix, double inv)
{
  double f = 1.0;
  if (i < static_cast<int>(npix)) f *= exp(-static_cast<double>(npix - i - 1) * inv);
  const int j = N - 1 - i;
  if (j < static_cast<int>(npix)) f *= exp(-static_cast<double>(npix - j - 1) * inv);
  return f;
}

// One row a x_{i-1} + b x_i + c x_{i+1} = d of the reduced system.
struct TriRow {
  cx<double> a, b, c, d;
};

// Parallel cyclic reduction of the M-row reduced system; lane p owns rows 2p
// and 2p+1 (r0, r1) and receives their solution (x0, x1).  Rows are kept
// normalised to a unit diagonal, so a level needs one reciprocal per row and
// only the neighbours' (a, c, d) from LDS (ra, rc, rd, cap entries each); a
// lane keeps its own rows in registers.  When 2M <= cap the levels alternate
// between two halves of the arrays (one barrier per level), otherwise they
// update in place (two).  Every lane of the work-group calls it; the barriers
// order LDS only.
__device__ __forceinline__ void cn_pcr(cx<double> *ra, cx<double> *rc, cx<double> *rd, int cap, int M, int p,
                                       bool active, const TriRow &r0, const TriRow &r1, cx<double> &x0,
                                       cx<double> &x1)
{
  const cx<double> zero{0.0, 0.0}, one{1.0, 0.0};
  const bool dbuf = 2 * M <= cap;
  cx<double> a[2], c[2], d[2];
  if (active) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const TriRow &r = q ? r1 : r0;
      const cx<double> ib = crecip(r.b);
      a[q] = r.a * ib;
      c[q] = r.c * ib;
      d[q] = r.d * ib;
      ra[2 * p + q] = a[q];
      rc[2 * p + q] = c[q];
      rd[2 * p + q] = d[q];
    }
  }
  lds_barrier();
  int base = 0;  // offset of the half holding the current level
  for (int h = 1; h < M; h *= 2) {
    cx<double> na[2], nc[2], nd[2];
    if (active) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int row = 2 * p + q;
        cx<double> am = zero, cm = zero, dm = zero, ap = zero, cp = zero, dp = zero;
        if (row - h >= 0) {
          am = ra[base + row - h];
          cm = rc[base + row - h];
          dm = rd[base + row - h];
        }
        if (row + h < M) {
          ap = ra[base + row + h];
          cp = rc[base + row + h];
          dp = rd[base + row + h];
        }
        // eliminate x_{row-h} and x_{row+h} with the (unit-diagonal) rows there
        const cx<double> ib = crecip(one - a[q] * cm - c[q] * ap);
        na[q] = (zero - a[q] * am) * ib;
        nc[q] = (zero - c[q] * cp) * ib;
        nd[q] = (d[q] - a[q] * dm - c[q] * dp) * ib;
      }
    }
    if (dbuf) {
      base = M - base;
    } else {
      lds_barrier();
    }
    if (active) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        a[q] = na[q];
        c[q] = nc[q];
        d[q] = nd[q];
        ra[base + 2 * p + q] = a[q];
        rc[base + 2 * p + q] = c[q];
        rd[base + 2 * p + q] = d[q];
      }
    }
    lds_barrier();
  }
  x0 = d[0];
  x1 = d[1];
}

// ---------------------------------------------------------------- CN march
// One work-group per job (blockIdx.x); jobs may differ in size, grid,
// wavenumber and material.  Lanes p >= job.P take part in barriers only.
template <class R>
__global__ __launch_bounds__(CN_MAX_P) void k_cn_jobs(const CnJob *__restrict__ jobs, cx<R> *fields,
                                                      char *scratch, cx<R> *sol, unsigned first,
                                                      unsigned nsteps)
{
  // two halves of 2 * CN_MAX_P rows each: PCR levels alternate between them
  __shared__ cx<double> ra[4 * CN_MAX_P], rc[4 * CN_MAX_P], rd[4 * CN_MAX_P];
  __shared__ uint64_t mask[TRI_MASK_WORDS];
  // a copy: the descriptor stays in registers instead of being re-read after
  // every barrier (barriers order global memory, so a reference reloads it)
  const CnJob J = jobs[blockIdx.x];
  const int p = threadIdx.x, N = J.nx, P = J.P, m = J.m;
  const bool active = p < P;
  char *base = scratch + J.scoff;
  cx<R> *bufA = fields + J.foff;
  cx<R> *bufB = reinterpret_cast<cx<R> *>(base);
  const size_t cst = align256(static_cast<size_t>(m) * P * 16);
  cx<double> *ca = reinterpret_cast<cx<double> *>(base + align256(static_cast<size_t>(N) * sizeof(cx<R>)));
  cx<double> *cc = reinterpret_cast<cx<double> *>(reinterpret_cast<char *>(ca) + cst);
  cx<double> *cd = reinterpret_cast<cx<double> *>(reinterpret_cast<char *>(cc) + cst);
  const double k = J.k;
  const double rho = J.dz / (k * J.dx * J.dx);
  const double r = k * J.dz;
  // balanced chunk [s, e] of lane p (cn_partition)
  const int len0 = N / P, rem = N % P;
  const int s = p * len0 + min(p, rem);
  const int e = active ? s + len0 + (p < rem ? 1 : 0) - 1 : s - 1;
  const int M = 2 * P;
  const int nw = J.mat.mode == MAT_CSG ? (J.mat.n_parts + 63) >> 6 : 0;

  for (unsigned it = 0; it < nsteps; ++it) {
    const unsigned st = first + it;
    const cx<R> *prev = (it & 1) ? bufB : bufA;
    cx<R> *next = (it & 1) ? bufA : bufB;
    const double z = coord(J.zmin, st, J.dz);
    if (J.mat.mode == MAT_CSG)
      csg_cull(J.mat, mask, nw, J.xmin, coord(J.xmin, N - 1, J.dx), 0.0, 0.0, z - J.dz, z);

    TriRow r0{}, r1{};
    if (active) {
      // chunk elimination: rows s+1..e expressed through x_s (and x_{e+1})
      cx<double> as, bs, cs, ds;
      cn_row(J, prev, s, st, z, rho, r, mask, nw, as, bs, cs, ds);
      cx<double> ah{}, ch{}, dh{};
      for (int i = s + 1; i <= e; ++i) {
        cx<double> ai, bi, ci, di;
        cn_row(J, prev, i, st, z, rho, r, mask, nw, ai, bi, ci, di);
        if (i == s + 1) {
          const cx<double> ib = crecip(bi);
          ah = ai * ib;
          ch = ci * ib;
          dh = di * ib;
        } else {
          const cx<double> iden = crecip(bi - ai * ch);
          const cx<double> ahn = (cx<double>{0.0, 0.0} - ai * ah) * iden;
          ch = ci * iden;
          dh = (di - ai * dh) * iden;
          ah = ahn;
        }
        const int j = i - s;
        ca[j * P + p] = ah;
        cc[j * P + p] = ch;
        cd[j * P + p] = dh;
      }
      // last row of the chunk is a reduced equation: ah x_s + x_e + ch x_{e+1} = dh
      r1 = {ah, {1.0, 0.0}, ch, dh};
      // backward: x_i = dt_i - at_i x_s - ct_i x_e, base at i = e: (0, 0, -1)
      cx<double> at{0.0, 0.0}, ct{-1.0, 0.0}, dt{0.0, 0.0};
      for (int i = e - 1; i >= s + 1; --i) {
        const int j = i - s;
        const cx<double> a0 = ca[j * P + p], c0 = cc[j * P + p], d0 = cd[j * P + p];
        dt = d0 - c0 * dt;
        at = a0 - c0 * at;
        ct = cx<double>{0.0, 0.0} - c0 * ct;
        ca[j * P + p] = at;
        cc[j * P + p] = ct;
        cd[j * P + p] = dt;
      }
      // first row of the chunk: a_s x_{s-1} + (b_s - c_s at) x_s - c_s ct x_e = d_s - c_s dt
      r0 = {as, bs - cs * at, cx<double>{0.0, 0.0} - cs * ct, ds - cs * dt};
    }
    cx<double> xs, xe;
    cn_pcr(ra, rc, rd, 4 * CN_MAX_P, M, p, active, r0, r1, xs, xe);
    if (active) {
      cx<R> *sc = (sol && J.soff != ~static_cast<size_t>(0) && static_cast<int>(st) < J.nz)
                      ? sol + J.soff + static_cast<size_t>(st) * N
                      : nullptr;
      for (int i = s; i <= e; ++i) {
        cx<double> x;
        if (i == s) x = xs;
        else if (i == e) x = xe;
        else {
          const int j = i - s;
          x = cd[j * P + p] - ca[j * P + p] * xs - cc[j * P + p] * xe;
        }
        if (absorbing(i, N, J.absorb_npix)) x = scale(x, absorb_factor(i, N, J.absorb_npix, J.absorb_inv));
        const cx<R> xr = cvt<R>(x);
        next[i] = xr;
        if (sc) sc[i] = xr;
      }
    }
    __syncthreads();
  }
  if ((nsteps & 1) && active)
    for (int i = s; i <= e; ++i) bufA[i] = bufB[i];
}

// CN march with the field in registers.  For jobs whose chunks have at most
// MM rows (N <= MM * CN_MAX_P): lane p keeps its rows of the field and its
// forward-elimination coefficients in registers for the whole run; only the
// chunk-edge values of the field (for the neighbouring rows) and the reduced
// system go through LDS, and the field goes back to memory once, at the end.
// The elimination is the same as in k_cn_jobs (Gaussian elimination without
// pivoting over the chunk, PCR over the 2P chunk ends); the interior is
// recovered by back substitution x_i = dh_i - ah_i x_s - ch_i x_{i+1}.
template <class R, int MM>
__global__ __launch_bounds__(CN_MAX_P) void k_cn_reg(const CnJob *__restrict__ jobs, cx<R> *fields, cx<R> *sol,
                                                     unsigned first, unsigned nsteps)
{
  // two halves of 2 * CN_MAX_P rows each: PCR levels alternate between them
  __shared__ cx<double> ra[4 * CN_MAX_P], rc[4 * CN_MAX_P], rd[4 * CN_MAX_P];
  __shared__ cx<double> edge_lo[CN_MAX_P], edge_hi[CN_MAX_P];
  __shared__ uint64_t mask[TRI_MASK_WORDS];
  // The scalars of the descriptor the step loop uses are copied: barriers
  // order global memory, so fields read through the reference would be
  // reloaded after each one.  The material stays behind the reference (a full
  // copy spills scalar registers into vector ones).
  const CnJob &J = jobs[blockIdx.x];
  const int p = threadIdx.x, N = J.nx, P = J.P, bc = J.bc;
  const bool active = p < P;
  cx<R> *fld = fields + J.foff;
  const double k = J.k, dz = J.dz, dx = J.dx, xmin = J.xmin, zmin = J.zmin;
  const double rho = dz / (k * dx * dx);
  const double r = k * dz;
  const unsigned npix = J.absorb_npix;
  const double ainv = J.absorb_inv;
  const size_t soff = J.soff;
  const int nz = J.nz;
  const int len0 = N / P, rem = N % P;
  const int s = p * len0 + min(p, rem);
  const int len = active ? len0 + (p < rem ? 1 : 0) : 0;  // 2 <= len <= MM for active lanes
  const int M = 2 * P;
  const int nw = J.mat.mode == MAT_CSG ? (J.mat.n_parts + 63) >> 6 : 0;
  const cx<double> zero{0.0, 0.0};

  cx<double> u[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) u[j] = j < len ? cvt<double>(fld[s + j]) : zero;
  // z-invariant planes: the lane's material once, not a global load per row
  // and step on the step's critical path
  const bool fixed_mat = J.mat.mode == MAT_PLANES && J.mat.z_invariant;
  // absorber factors of the lane's rows (1 outside the layer), once per call
  double af[MM];
  bool lane_absorbs = false;
#pragma unroll
  for (int j = 0; j < MM; ++j) {
    af[j] = 1.0;
    if (j < len && absorbing(s + j, N, npix)) {
      af[j] = absorb_factor(s + j, N, npix, ainv);
      lane_absorbs = true;
    }
  }
  double fd[MM], fb[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) {
    fd[j] = fb[j] = 0.0;
    if (fixed_mat && j < len) {
      fd[j] = J.mat.planes[2 * (s + j)];
      fb[j] = J.mat.planes[2 * (s + j) + 1];
    }
  }
  // row i = s + j of the step
  auto row = [&](int j, unsigned st, double z, cx<double> ul_, cx<double> u_, cx<double> ur_, cx<double> uin,
                 cx<double> &ai, cx<double> &bi, cx<double> &ci, cx<double> &di) {
    double d1 = fd[j], b1 = fb[j], d0 = fd[j], b0 = fb[j];
    if (!fixed_mat) {
      const double x = coord(xmin, s + j, dx);
      material_2d(J.mat, mask, nw, st, s + j, x, z, d1, b1);
      material_2d(J.mat, mask, nw, st - 1, s + j, x, z - dz, d0, b0);
    }
    cn_row_m(N, bc, s + j, rho, r, d1, b1, d0, b0, ul_, u_, ur_, uin, ai, bi, ci, di);
  };

  for (unsigned it = 0; it < nsteps; ++it) {
    const unsigned st = first + it;
    const double z = coord(zmin, st, dz);
    if (J.mat.mode == MAT_CSG)
      csg_cull(J.mat, mask, nw, xmin, coord(xmin, N - 1, dx), 0.0, 0.0, z - dz, z);
    // chunk-edge values for the neighbouring lanes
    cx<double> ulast = u[0], u2last = u[0];
#pragma unroll
    for (int j = 1; j < MM; ++j) {
      if (j == len - 1) ulast = u[j];
      if (j == len - 2) u2last = u[j];
    }
    if (active) {
      edge_lo[p] = u[0];
      edge_hi[p] = ulast;
    }
    lds_barrier();
    const cx<double> ul = (active && p > 0) ? edge_hi[p - 1] : zero;
    const cx<double> ur = (active && p < P - 1) ? edge_lo[p + 1] : zero;

    cx<double> ah[MM], ch[MM], dh[MM];
    TriRow r0{}, r1{};
    if (active) {
      // first row of the chunk
      cx<double> as, bs, cs, ds;
      row(0, st, z, ul, u[0], u[1], u[1], as, bs, cs, ds);
      // forward elimination of rows s+1 .. s+len-1
#pragma unroll
      for (int j = 1; j < MM; ++j) {
        if (j < len) {
          const int i = s + j;
          const cx<double> un = (j + 1 < MM && j + 1 < len) ? u[j + 1 < MM ? j + 1 : 0] : ur;
          const cx<double> uin = (i == N - 1) ? u2last : u[j];
          cx<double> ai, bi, ci, di;
          row(j, st, z, u[j - 1], u[j], un, uin, ai, bi, ci, di);
          if (j == 1) {
            const cx<double> ib = crecip(bi);
            ah[j] = ai * ib;
            ch[j] = ci * ib;
            dh[j] = di * ib;
          } else {
            const cx<double> iden = crecip(bi - ai * ch[j - 1]);
            ah[j] = (zero - ai * ah[j - 1]) * iden;
            ch[j] = ci * iden;
            dh[j] = (di - ai * dh[j - 1]) * iden;
          }
        }
      }
      // the chunk's last row is a reduced equation: ah x_s + x_e + ch x_{e+1} = dh
      cx<double> ahl = zero, chl = zero, dhl = zero;
#pragma unroll
      for (int j = 1; j < MM; ++j)
        if (j == len - 1) {
          ahl = ah[j];
          chl = ch[j];
          dhl = dh[j];
        }
      r1 = {ahl, {1.0, 0.0}, chl, dhl};
      // backward: row s+1 through x_s and x_e (running values, base (0, 0, -1) at e)
      cx<double> at = zero, ct{-1.0, 0.0}, dt = zero;
#pragma unroll
      for (int j = MM - 2; j >= 1; --j) {
        if (j <= len - 2) {
          dt = dh[j] - ch[j] * dt;
          at = ah[j] - ch[j] * at;
          ct = zero - ch[j] * ct;
        }
      }
      r0 = {as, bs - cs * at, zero - cs * ct, ds - cs * dt};
    }
    cx<double> xs, xe;
    cn_pcr(ra, rc, rd, 4 * CN_MAX_P, M, p, active, r0, r1, xs, xe);
    if (active) {
      // back substitution from the chunk's end
      cx<double> xn = xe;
#pragma unroll
      for (int j = MM - 1; j >= 1; --j) {
        if (j == len - 1) {
          u[j] = xe;
        } else if (j < len - 1) {
          u[j] = dh[j] - ah[j] * xs - ch[j] * xn;
          xn = u[j];
        }
      }
      u[0] = xs;
      cx<R> *sc = (sol && soff != ~static_cast<size_t>(0) && static_cast<int>(st) < nz)
                      ? sol + soff + static_cast<size_t>(st) * N
                      : nullptr;
#pragma unroll
      for (int j = 0; j < MM; ++j) {
        if (j < len) {
          if (lane_absorbs) u[j] = scale(u[j], af[j]);
          if (sc) sc[s + j] = cvt<R>(u[j]);
        }
      }
    }
    lds_barrier();  // edge arrays and the reduced system are reused next step
  }
#pragma unroll
  for (int j = 0; j < MM; ++j)
    if (j < len) fld[s + j] = cvt<R>(u[j]);
}

// CN march for a z-invariant material (z-invariant planes, or none): every
// row of the system is the same at every step except the transparent-boundary
// terms, a change of the diagonal at rows 0 and N-1 (cn_row_m).  So the system
// without them, A0, is factored once per call, with k_cn_reg's elimination
// (chunks, then PCR over the chunk ends): the local factors and, per PCR level
// and reduced row, the row's (a, c) and the reciprocal of its new diagonal, all
// kept in registers.  A step then carries only its right-hand side through the
// factors, y = A0^-1 d, with no complex reciprocal on its path, and applies the
// boundary terms by the Woodbury identity, A = A0 + U D U^T with U = [e0, eN],
// D = diag(d0, dN):
//   x = y - W (I + D G)^-1 D U^T y,  W = A0^-1 U (two solves, once), G = U^T W.
// Lanes are limited to CN_FIXED_P so that the PCR factors of every level fit in
// registers (N <= 1024 with k_cn_reg's partition).
constexpr int CN_FIXED_P = 256;
constexpr int CN_FIXED_LEVELS = 9;  // PCR levels for up to 2 * CN_FIXED_P reduced rows

template <class R, int MM>
__global__ __launch_bounds__(CN_FIXED_P) void k_cn_fixed(const CnJob *__restrict__ jobs, cx<R> *fields, cx<R> *sol,
                                                         unsigned first, unsigned nsteps)
{
  __shared__ cx<double> ra[4 * CN_FIXED_P], rc[4 * CN_FIXED_P], rd[4 * CN_FIXED_P];
  __shared__ cx<double> edge_lo[CN_FIXED_P], edge_hi[CN_FIXED_P];
  __shared__ cx<double> wb[8];
  const CnJob &J = jobs[blockIdx.x];
  const int p = threadIdx.x, N = J.nx, P = J.P, bc = J.bc;
  const bool active = p < P;
  cx<R> *fld = fields + J.foff;
  const double k = J.k, dz = J.dz, dx = J.dx;
  const double rho = dz / (k * dx * dx);
  const double r = k * dz;
  const unsigned npix = J.absorb_npix;
  const double ainv = J.absorb_inv;
  const size_t soff = J.soff;
  const int nz = J.nz;
  const int len0 = N / P, rem = N % P;
  const int s = p * len0 + min(p, rem);
  const int len = active ? len0 + (p < rem ? 1 : 0) : 0;  // 2 <= len <= MM for active lanes
  const int M = 2 * P;
  const bool dbuf = 2 * M <= 4 * CN_FIXED_P;
  const bool tbc = bc == PXG_BC_TRANSPARENT && N >= 2;
  const cx<double> zero{0.0, 0.0}, one{1.0, 0.0};

  cx<double> u[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) u[j] = j < len ? cvt<double>(fld[s + j]) : zero;
  double af[MM], fd[MM], fb[MM];
  bool lane_absorbs = false;
#pragma unroll
  for (int j = 0; j < MM; ++j) {
    af[j] = 1.0;
    fd[j] = fb[j] = 0.0;
    if (j < len) {
      if (absorbing(s + j, N, npix)) {
        af[j] = absorb_factor(s + j, N, npix, ainv);
        lane_absorbs = true;
      }
      if (J.mat.mode == MAT_PLANES) {
        fd[j] = J.mat.planes[2 * (s + j)];
        fb[j] = J.mat.planes[2 * (s + j) + 1];
      }
    }
  }
  // rows of A0 (cn_row_m without the boundary terms)
  const cx<double> off = {0.0, -0.25 * rho};
  auto a_of = [&](int j) { return s + j > 0 ? off : zero; };
  auto c_of = [&](int j) { return s + j < N - 1 ? off : zero; };
  auto b_of = [&](int j) { return cx<double>{1.0 + 0.5 * (fb[j] * r), 0.5 * rho + 0.5 * (fd[j] * r)}; };

  // local factors of the chunk rows 1 .. len-1 and the reduced rows' coefficients
  cx<double> iden[MM], ah[MM], ch[MM];
  cx<double> ra0 = zero, rb0 = one, rc0 = zero, ra1 = zero, rc1 = zero;
#pragma unroll
  for (int j = 0; j < MM; ++j) iden[j] = ah[j] = ch[j] = zero;
  if (active) {
#pragma unroll
    for (int j = 1; j < MM; ++j) {
      if (j < len) {
        const cx<double> ai = a_of(j), bi = b_of(j), ci = c_of(j);
        if (j == 1) {
          iden[j] = crecip(bi);
          ah[j] = ai * iden[j];
          ch[j] = ci * iden[j];
        } else {
          iden[j] = crecip(bi - ai * ch[j - 1]);
          ah[j] = (zero - ai * ah[j - 1]) * iden[j];
          ch[j] = ci * iden[j];
        }
      }
    }
#pragma unroll
    for (int j = 1; j < MM; ++j)
      if (j == len - 1) {
        ra1 = ah[j];
        rc1 = ch[j];
      }
    cx<double> at = zero, ct{-1.0, 0.0};
#pragma unroll
    for (int j = MM - 2; j >= 1; --j) {
      if (j <= len - 2) {
        at = ah[j] - ch[j] * at;
        ct = zero - ch[j] * ct;
      }
    }
    ra0 = a_of(0);
    rb0 = b_of(0) - c_of(0) * at;
    rc0 = zero - c_of(0) * ct;
  }
  // PCR factors: the normalisation of the two reduced rows, then per level the
  // rows' (a, c) before the level and the reciprocal of the level's diagonal
  cx<double> nb[2];
  cx<double> la[CN_FIXED_LEVELS][2], lc[CN_FIXED_LEVELS][2], lb[CN_FIXED_LEVELS][2];
  {
    cx<double> a[2] = {zero, zero}, c[2] = {zero, zero};
    nb[0] = nb[1] = one;
    if (active) {
      nb[0] = crecip(rb0);
      nb[1] = crecip(one);
      a[0] = ra0 * nb[0];
      c[0] = rc0 * nb[0];
      a[1] = ra1 * nb[1];
      c[1] = rc1 * nb[1];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        ra[2 * p + q] = a[q];
        rc[2 * p + q] = c[q];
      }
    }
    lds_barrier();
    int base = 0;
#pragma unroll
    for (int lv = 0; lv < CN_FIXED_LEVELS; ++lv) {
      const int h = 1 << lv;
      if (h >= M) break;
      cx<double> na[2], nc[2];
      if (active) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int row = 2 * p + q;
          cx<double> am = zero, cm = zero, ap = zero, cp = zero;
          if (row - h >= 0) {
            am = ra[base + row - h];
            cm = rc[base + row - h];
          }
          if (row + h < M) {
            ap = ra[base + row + h];
            cp = rc[base + row + h];
          }
          const cx<double> ib = crecip(one - a[q] * cm - c[q] * ap);
          la[lv][q] = a[q];
          lc[lv][q] = c[q];
          lb[lv][q] = ib;
          na[q] = (zero - a[q] * am) * ib;
          nc[q] = (zero - c[q] * cp) * ib;
        }
      }
      if (dbuf) {
        base = M - base;
      } else {
        lds_barrier();
      }
      if (active) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          a[q] = na[q];
          c[q] = nc[q];
          ra[base + 2 * p + q] = a[q];
          rc[base + 2 * p + q] = c[q];
        }
      }
      lds_barrier();
    }
  }

  // y = A0^-1 d for the lane's rows (every lane calls it: it holds barriers)
  auto solve = [&](const cx<double> (&d)[MM], cx<double> (&y)[MM]) {
    cx<double> dh[MM];
#pragma unroll
    for (int j = 0; j < MM; ++j) dh[j] = zero;
    cx<double> d0 = zero, d1 = zero;
    if (active) {
#pragma unroll
      for (int j = 1; j < MM; ++j) {
        if (j < len) dh[j] = (j == 1) ? d[j] * iden[j] : (d[j] - a_of(j) * dh[j - 1]) * iden[j];
      }
      cx<double> dt = zero;
#pragma unroll
      for (int j = MM - 2; j >= 1; --j)
        if (j <= len - 2) dt = dh[j] - ch[j] * dt;
      cx<double> dl = zero;
#pragma unroll
      for (int j = 1; j < MM; ++j)
        if (j == len - 1) dl = dh[j];
      d0 = (d[0] - c_of(0) * dt) * nb[0];
      d1 = dl * nb[1];
      rd[2 * p] = d0;
      rd[2 * p + 1] = d1;
    }
    lds_barrier();
    int base = 0;
#pragma unroll
    for (int lv = 0; lv < CN_FIXED_LEVELS; ++lv) {
      const int h = 1 << lv;
      if (h >= M) break;
      cx<double> nd[2];
      if (active) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int row = 2 * p + q;
          const cx<double> dm = row - h >= 0 ? rd[base + row - h] : zero;
          const cx<double> dp = row + h < M ? rd[base + row + h] : zero;
          nd[q] = ((q ? d1 : d0) - la[lv][q] * dm - lc[lv][q] * dp) * lb[lv][q];
        }
      }
      if (dbuf) {
        base = M - base;
      } else {
        lds_barrier();
      }
      if (active) {
        d0 = nd[0];
        d1 = nd[1];
        rd[base + 2 * p] = d0;
        rd[base + 2 * p + 1] = d1;
      }
      lds_barrier();
    }
    if (active) {
      cx<double> xn = d1;
#pragma unroll
      for (int j = MM - 1; j >= 1; --j) {
        if (j == len - 1) {
          y[j] = d1;
        } else if (j < len - 1) {
          y[j] = dh[j] - ah[j] * d0 - ch[j] * xn;
          xn = y[j];
        }
      }
      y[0] = d0;
    }
  };

  // Woodbury columns W = A0^-1 [e0, eN] and G = U^T W
  cx<double> w0[MM], wn[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) w0[j] = wn[j] = zero;
  cx<double> g00 = zero, g01 = zero, g10 = zero, g11 = zero;
  if (tbc) {
    cx<double> e[MM];
#pragma unroll
    for (int j = 0; j < MM; ++j) e[j] = (active && j < len && s + j == 0) ? one : zero;
    solve(e, w0);
#pragma unroll
    for (int j = 0; j < MM; ++j) e[j] = (active && j < len && s + j == N - 1) ? one : zero;
    solve(e, wn);
    if (active) {
#pragma unroll
      for (int j = 0; j < MM; ++j) {
        if (j < len && s + j == 0) {
          wb[0] = w0[j];
          wb[1] = wn[j];
        }
        if (j < len && s + j == N - 1) {
          wb[2] = w0[j];
          wb[3] = wn[j];
        }
      }
    }
    lds_barrier();
    g00 = wb[0];
    g01 = wb[1];
    g10 = wb[2];
    g11 = wb[3];
    lds_barrier();
  }

  for (unsigned it = 0; it < nsteps; ++it) {
    const unsigned st = first + it;
    cx<double> ulast = u[0], u2last = u[0];
#pragma unroll
    for (int j = 1; j < MM; ++j) {
      if (j == len - 1) ulast = u[j];
      if (j == len - 2) u2last = u[j];
    }
    if (active) {
      edge_lo[p] = u[0];
      edge_hi[p] = ulast;
    }
    lds_barrier();
    const cx<double> ul = (active && p > 0) ? edge_hi[p - 1] : zero;
    const cx<double> ur = (active && p < P - 1) ? edge_lo[p + 1] : zero;
    // right-hand side (cn_row_m) and the boundary rows' diagonal changes
    cx<double> d[MM], y[MM];
    cx<double> dl0 = zero, dlN = zero;
#pragma unroll
    for (int j = 0; j < MM; ++j) {
      d[j] = y[j] = zero;
      if (active && j < len) {
        const cx<double> ulj = j == 0 ? ul : u[j > 0 ? j - 1 : 0];
        const cx<double> urj = (j + 1 < len) ? u[j + 1 < MM ? j + 1 : 0] : ur;
        cx<double> v = (ulj + urj) * cx<double>{0.0, 0.25 * rho};
        v = v - cx<double>{-0.5 * rho * u[j].y, 0.5 * rho * u[j].x};
        v = v + cx<double>{1.0 - 0.5 * (fb[j] * r), -0.5 * (fd[j] * r)} * u[j];
        const int i = s + j;
        if (tbc && (i == 0 || i == N - 1)) {
          const cx<double> uin = i == 0 ? u[j + 1 < MM ? j + 1 : 0] : u2last;
          cx<double> e;
          if (tbc_factor(u[j], uin, 1e-16, e)) {
            const cx<double> q = cx<double>{0.0, 0.25 * rho} * e;
            v = v + q * u[j];
            if (i == 0) dl0 = zero - q;
            if (i == N - 1) dlN = zero - q;
          }
        }
        d[j] = v;
      }
    }
    solve(d, y);
    if (tbc) {
      // x = y - W (I + D G)^-1 D [y0; yN]
      if (active) {
#pragma unroll
        for (int j = 0; j < MM; ++j) {
          if (j < len && s + j == 0) {
            wb[4] = y[j];
            wb[6] = dl0;
          }
          if (j < len && s + j == N - 1) {
            wb[5] = y[j];
            wb[7] = dlN;
          }
        }
      }
      lds_barrier();
      const cx<double> y0 = wb[4], yn = wb[5], e0 = wb[6], en = wb[7];
      const cx<double> m00 = one + e0 * g00, m01 = e0 * g01, m10 = en * g10, m11 = one + en * g11;
      const cx<double> r0 = e0 * y0, r1 = en * yn;
      const cx<double> idet = crecip(m00 * m11 - m01 * m10);
      const cx<double> c0 = (m11 * r0 - m01 * r1) * idet, c1 = (m00 * r1 - m10 * r0) * idet;
#pragma unroll
      for (int j = 0; j < MM; ++j) y[j] = y[j] - w0[j] * c0 - wn[j] * c1;
    }
    if (active) {
      cx<R> *sc = (sol && soff != ~static_cast<size_t>(0) && static_cast<int>(st) < nz)
                      ? sol + soff + static_cast<size_t>(st) * N
                      : nullptr;
#pragma unroll
      for (int j = 0; j < MM; ++j) {
        if (j < len) {
          u[j] = lane_absorbs ? scale(y[j], af[j]) : y[j];
          if (sc) sc[s + j] = cvt<R>(u[j]);
        }
      }
    }
    lds_barrier();  // edge arrays, the reduced system and wb are reused next step
  }
#pragma unroll
  for (int j = 0; j < MM; ++j)
    if (j < len) fld[s + j] = cvt<R>(u[j]);
}

// CN march for chunks of 5..MM rows (N <= MM * CN_MAX_P).  The field stays in
// registers as in k_cn_reg; the elimination coefficients of the chunk rows go
// to the job's scratch (coalesced [row][lane]) and are read back once, by the
// back substitution.  The reduced equation of the chunk's first row needs the
// backward running values (at, ct, dt) of k_cn_reg; here they come from the
// forward pass: the backward recurrence v_j = h_j - ch_j v_{j+1} is affine, so
// its composition over rows 1..len-2 is carried forward as v_1 = P + Q v
// (P <- P + Q h_j, Q <- -Q ch_j), with no second pass over the rows.
constexpr int CN_BIG_ROWS = 16;

template <class R, int MM>
__global__ __launch_bounds__(CN_MAX_P) void k_cn_big(const CnJob *__restrict__ jobs, cx<R> *fields, char *scratch,
                                                     cx<R> *sol, unsigned first, unsigned nsteps)
{
  // two halves of 2 * CN_MAX_P rows each: PCR levels alternate between them
  __shared__ cx<double> ra[4 * CN_MAX_P], rc[4 * CN_MAX_P], rd[4 * CN_MAX_P];
  __shared__ cx<double> edge_lo[CN_MAX_P], edge_hi[CN_MAX_P];
  __shared__ uint64_t mask[TRI_MASK_WORDS];
  const CnJob &J = jobs[blockIdx.x];
  const int p = threadIdx.x, N = J.nx, P = J.P, m = J.m;
  const bool active = p < P;
  cx<R> *fld = fields + J.foff;
  const double k = J.k, dz = J.dz, dx = J.dx, xmin = J.xmin, zmin = J.zmin;
  const double rho = dz / (k * dx * dx);
  const double r = k * dz;
  const unsigned npix = J.absorb_npix;
  const double ainv = J.absorb_inv;
  const size_t soff = J.soff;
  const int nz = J.nz;
  const int len0 = N / P, rem = N % P;
  const int s = p * len0 + min(p, rem);
  const int len = active ? len0 + (p < rem ? 1 : 0) : 0;  // 2 <= len <= m <= MM for active lanes
  const int M = 2 * P;
  const int nw = J.mat.mode == MAT_CSG ? (J.mat.n_parts + 63) >> 6 : 0;
  const cx<double> zero{0.0, 0.0}, one{1.0, 0.0};
  // scratch: [field copy][ah][ch][dh], each coefficient array m rows x P lanes
  char *base = scratch + J.scoff;
  const size_t cst = align256(static_cast<size_t>(m) * P * 16);
  cx<double> *ca = reinterpret_cast<cx<double> *>(base + align256(static_cast<size_t>(N) * sizeof(cx<R>)));
  cx<double> *cc = reinterpret_cast<cx<double> *>(reinterpret_cast<char *>(ca) + cst);
  cx<double> *cd = reinterpret_cast<cx<double> *>(reinterpret_cast<char *>(cc) + cst);

  cx<double> u[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) u[j] = j < len ? cvt<double>(fld[s + j]) : zero;

  for (unsigned it = 0; it < nsteps; ++it) {
    const unsigned st = first + it;
    const double z = coord(zmin, st, dz);
    if (J.mat.mode == MAT_CSG) csg_cull(J.mat, mask, nw, xmin, coord(xmin, N - 1, dx), 0.0, 0.0, z - dz, z);
    cx<double> ulast = u[0], u2last = u[0];
#pragma unroll
    for (int j = 1; j < MM; ++j) {
      if (j == len - 1) ulast = u[j];
      if (j == len - 2) u2last = u[j];
    }
    if (active) {
      edge_lo[p] = u[0];
      edge_hi[p] = ulast;
    }
    lds_barrier();
    const cx<double> ul = (active && p > 0) ? edge_hi[p - 1] : zero;
    const cx<double> ur = (active && p < P - 1) ? edge_lo[p + 1] : zero;

    TriRow r0{}, r1{};
    if (active) {
      auto row = [&](int j, cx<double> ul_, cx<double> u_, cx<double> ur_, cx<double> uin, cx<double> &ai,
                     cx<double> &bi, cx<double> &ci, cx<double> &di) {
        cn_row_v(J, s + j, st, z, rho, r, mask, nw, ul_, u_, ur_, uin, ai, bi, ci, di);
      };
      cx<double> as, bs, cs, ds;
      row(0, ul, u[0], u[1], u[1], as, bs, cs, ds);
      cx<double> ah = zero, ch = zero, dh = zero;
      // forward composition of the backward recurrence over rows 1..len-2
      cx<double> Pa = zero, Pd = zero, Q = one;
#pragma unroll
      for (int j = 1; j < MM; ++j) {
        if (j < len) {
          const int i = s + j;
          const cx<double> un = (j + 1 < MM && j + 1 < len) ? u[j + 1 < MM ? j + 1 : 0] : ur;
          const cx<double> uin = (i == N - 1) ? u2last : u[j];
          cx<double> ai, bi, ci, di;
          row(j, u[j - 1], u[j], un, uin, ai, bi, ci, di);
          if (j == 1) {
            const cx<double> ib = crecip(bi);
            ah = ai * ib;
            ch = ci * ib;
            dh = di * ib;
          } else {
            const cx<double> iden = crecip(bi - ai * ch);
            ah = (zero - ai * ah) * iden;
            ch = ci * iden;
            dh = (di - ai * dh) * iden;
          }
          if (j <= len - 2) {
            ca[j * P + p] = ah;
            cc[j * P + p] = ch;
            cd[j * P + p] = dh;
            Pa = Pa + Q * ah;
            Pd = Pd + Q * dh;
            Q = zero - Q * ch;
          }
        }
      }
      // last row of the chunk: ah x_s + x_e + ch x_{e+1} = dh; first row through
      // at = Pa, dt = Pd, ct = -Q (base values 0, 0, -1 at the chunk's end)
      r1 = {ah, one, ch, dh};
      r0 = {as, bs - cs * Pa, cs * Q, ds - cs * Pd};
    }
    cx<double> xs, xe;
    cn_pcr(ra, rc, rd, 4 * CN_MAX_P, M, p, active, r0, r1, xs, xe);
    if (active) {
      // back substitution from the chunk's end; rows below m - 1 exist in the
      // scratch for every lane, so their loads need no per-lane guard
      cx<double> xn = xe;
#pragma unroll
      for (int j = MM - 1; j >= 1; --j) {
        if (j < m - 1) {
          const cx<double> a0 = ca[j * P + p], c0 = cc[j * P + p], d0 = cd[j * P + p];
          if (j < len - 1) {
            u[j] = d0 - a0 * xs - c0 * xn;
            xn = u[j];
          }
        }
        if (j == len - 1) u[j] = xe;
      }
      u[0] = xs;
      cx<R> *sc = (sol && soff != ~static_cast<size_t>(0) && static_cast<int>(st) < nz)
                      ? sol + soff + static_cast<size_t>(st) * N
                      : nullptr;
#pragma unroll
      for (int j = 0; j < MM; ++j) {
        if (j < len) {
          if (absorbing(s + j, N, npix)) u[j] = scale(u[j], absorb_factor(s + j, N, npix, ainv));
          if (sc) sc[s + j] = cvt<R>(u[j]);
        }
      }
    }
    lds_barrier();  // edge slots are rewritten next step
  }
#pragma unroll
  for (int j = 0; j < MM; ++j)
    if (j < len) fld[s + j] = cvt<R>(u[j]);
}

int cn_big_rows() { return CN_BIG_ROWS; }

template <class R>
hipError_t launch_cn_big(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, void *scratch, cx<R> *sol,
                         unsigned first, unsigned nsteps, hipStream_t s)
{
  hipLaunchKernelGGL((k_cn_big<R, CN_BIG_ROWS>), dim3(njobs), dim3(max_P), 0, s, jobs_dev, fields,
                     static_cast<char *>(scratch), sol, first, nsteps);
  return hipGetLastError();
}

size_t cn_job_scratch(int nx, int esz)
{
  return cn_field_scratch(nx, esz, cn_partition(nx));
}

template <class R>
hipError_t launch_cn_reg(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, cx<R> *sol,
                         unsigned first, unsigned nsteps, hipStream_t s)
{
  hipLaunchKernelGGL((k_cn_reg<R, CN_REG_ROWS>), dim3(njobs), dim3(max_P), 0, s, jobs_dev, fields, sol, first,
                     nsteps);
  return hipGetLastError();
}

template <class R>
hipError_t launch_cn_fixed(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, cx<R> *sol,
                           unsigned first, unsigned nsteps, hipStream_t s)
{
  if (max_P > CN_FIXED_P) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_cn_fixed<R, CN_REG_ROWS>), dim3(njobs), dim3(max_P), 0, s, jobs_dev, fields, sol, first,
                     nsteps);
  return hipGetLastError();
}

int cn_fixed_max_p() { return CN_FIXED_P; }

template <class R>
hipError_t launch_cn_jobs(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, void *scratch,
                          cx<R> *sol, unsigned first, unsigned nsteps, hipStream_t s)
{
  hipLaunchKernelGGL(k_cn_jobs<R>, dim3(njobs), dim3(max_P), 0, s, jobs_dev, fields,
                     static_cast<char *>(scratch), sol, first, nsteps);
  return hipGetLastError();
}

// ---------------------------------------------------------------- ADI
// One lane per line.  IMPX: line = row i (y), walking x (j); cell (i, j) at
// j*N + i; material at z_step.  Otherwise line = column i (x), walking y (j);
// cell (j, i) at i*N + j; material at z_step + dz/2.
template <class R, bool IMPX>
__global__ __launch_bounds__(256) void k_adi_line(TriArgs<R> a, const cx<R> *__restrict__ prev_all,
                                                  cx<R> *__restrict__ next_all,
                                                  cx<double> *__restrict__ work_all, size_t per_field)
{
  __shared__ uint64_t mask[TRI_MASK_WORDS];
  const int N = a.n;
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const size_t cells = static_cast<size_t>(N) * N;
  const cx<R> *prev = prev_all + b * cells;
  cx<R> *next = next_all + b * cells;
  cx<double> *wsup = reinterpret_cast<cx<double> *>(reinterpret_cast<char *>(work_all) + per_field * b);
  cx<double> *wrhs = wsup + cells;
  const double k = a.kvec[b];
  const double h = a.dz / 2.0;
  const unsigned st = a.first_step;
  const double z = IMPX ? coord(a.zmin, st, a.dz) : coord(a.zmin, st, a.dz) + h;
  const int nw = a.mat.mode == MAT_CSG ? (a.mat.n_parts + 63) >> 6 : 0;
  if (a.mat.mode == MAT_CSG) {
    const int l0 = blockIdx.x * blockDim.x, l1 = min(N - 1, l0 + static_cast<int>(blockDim.x) - 1);
    if (IMPX)
      csg_cull(a.mat, mask, nw, a.xmin, coord(a.xmin, N - 1, a.dx), coord(a.ymin, l0, a.dy),
               coord(a.ymin, l1, a.dy), z, z);
    else
      csg_cull(a.mat, mask, nw, coord(a.xmin, l0, a.dx), coord(a.xmin, l1, a.dx), a.ymin,
               coord(a.ymin, N - 1, a.dy), z, z);
  }
  if (line >= N) return;
  const double ddx = k * a.dx * a.dx, ddy = k * a.dy * a.dy;
  // implicit-direction spacing term and explicit-direction spacing term
  const double dimp = IMPX ? ddx : ddy;
  const double dexp = IMPX ? ddy : ddx;
  const cx<double> off = {0.0, -0.5 / dimp};
  // cell access: (pos along line j, neighbour line offset o) -> element index
  auto at = [&](int j, int lo) -> size_t {
    const int l = line + lo;
    return IMPX ? static_cast<size_t>(j) * N + l : static_cast<size_t>(l) * N + j;
  };
  const double *planes = nullptr;
  if (a.mat.mode == MAT_PLANES) {
    const int pi = a.mat.z_invariant ? 0 : static_cast<int>(2 * st + (IMPX ? 0 : 1)) - a.mat.first_plane;
    planes = a.mat.planes + static_cast<size_t>(pi) * a.mat.plane_stride;
  }
  // explicit-direction boundary: this line is the first/last line
  const bool edge_lo = line == 0, edge_hi = line == N - 1;

  cx<double> supp{0.0, 0.0}, rhsp{0.0, 0.0};
  for (int j = 0; j < N; ++j) {
    const size_t idx = at(j, 0);
    double d = 0.0, be = 0.0;
    if (a.mat.mode == MAT_CSG) {
      const double x = IMPX ? coord(a.xmin, j, a.dx) : coord(a.xmin, line, a.dx);
      const double y = IMPX ? coord(a.ymin, line, a.dy) : coord(a.ymin, j, a.dy);
      part_values(a.mat, csg_eval(a.mat, mask, nw, x, y, z), d, be);
    } else if (planes) {
      d = planes[2 * idx];
      be = planes[2 * idx + 1];
    }
    cx<double> diag;
    if (IMPX || (a.quirks & PXG_FIX_ADI_Y_DIAG)) {
      diag = {1.0 / h + be * k, 1.0 / dimp + d * k};
    } else {
      diag = cx<double>{1.0 / h, 0.0} + cdiv(cx<double>{0.0, 1.0}, cx<double>{ddy + be * k, d * k});
    }
    const cx<double> u = ld(prev, static_cast<int>(idx));
    cx<double> rhs = cx<double>{1.0 / h, -1.0 / dexp} * u;
    if (line > 0) {
      const cx<double> un = ld(prev, static_cast<int>(at(j, -1)));
      rhs = rhs + cx<double>{-0.5 * un.y / dexp, 0.5 * un.x / dexp};
    }
    if (line < N - 1) {
      const cx<double> un = ld(prev, static_cast<int>(at(j, +1)));
      rhs = rhs + cx<double>{-0.5 * un.y / dexp, 0.5 * un.x / dexp};
    }
    if (a.use_tbc) {
      cx<double> e;
      // implicit direction: diagonal of the line ends
      if (j == N - 1 && tbc_factor(u, ld(prev, static_cast<int>(at(N - 2, 0))), 1e-6, e))
        diag = diag - cx<double>{-0.5 * e.y / dimp, 0.5 * e.x / dimp};
      if (j == 0 && tbc_factor(u, ld(prev, static_cast<int>(at(1, 0))), 1e-6, e))
        diag = diag - cx<double>{-0.5 * e.y / dimp, 0.5 * e.x / dimp};
      // explicit direction: right-hand side of the first/last line
      if (edge_hi && tbc_factor(u, ld(prev, static_cast<int>(at(j, -1))), 1e-6, e)) {
        const cx<double> q = cx<double>{-0.5 * e.y, 0.5 * e.x} * u;
        rhs = rhs + cx<double>{q.x / dexp, q.y / dexp};
      }
      if (edge_lo && tbc_factor(u, ld(prev, static_cast<int>(at(j, +1))), 1e-6, e)) {
        const cx<double> q = cx<double>{-0.5 * e.y, 0.5 * e.x} * u;
        rhs = rhs + cx<double>{q.x / dexp, q.y / dexp};
      }
    }
    // Thomas forward sweep (sup = sub = off inside the line, 0 at its ends)
    const cx<double> sup = j < N - 1 ? off : cx<double>{0.0, 0.0};
    if (j == 0) {
      supp = cdiv(sup, diag);
      rhsp = cdiv(rhs, diag);
    } else {
      const cx<double> den = diag - off * supp;
      rhsp = cdiv(rhs - off * rhsp, den);
      supp = j < N - 1 ? cdiv(sup, den) : cx<double>{0.0, 0.0};
    }
    wsup[static_cast<size_t>(j) * N + line] = supp;
    wrhs[static_cast<size_t>(j) * N + line] = rhsp;
  }
  cx<double> x = rhsp;
  next[at(N - 1, 0)] = cvt<R>(x);
  for (int j = N - 2; j >= 0; --j) {
    x = wrhs[static_cast<size_t>(j) * N + line] - wsup[static_cast<size_t>(j) * N + line] * x;
    next[at(j, 0)] = cvt<R>(x);
  }
}

template <class R>
hipError_t launch_adi_half(const TriArgs<R> &a, bool implicit_x, hipStream_t s)
{
  const size_t cells = static_cast<size_t>(a.n) * a.n;
  const size_t fbytes = align256(cells * sizeof(cx<R>));
  const size_t per = fbytes + 2 * align256(cells * 16);
  // scratch per field: [field copy][sup'][rhs']
  char *s0 = static_cast<char *>(a.scratch);
  // Dominant systems: converged tile-Jacobi; otherwise one sequential sweep per
  // line.  FP64 fields take the sweeps that reach FP64 rounding; complex64
  // fields stop where the truncation is 1e-10 relative, nearly three orders of
  // magnitude below the complex64 rounding (6e-8) of the stored result.
  const double tol = sizeof(R) == sizeof(double) ? 1e-17 : 1e-10;
  const int iters = adi_jacobi_iters(a.dz, a.kmin, implicit_x ? a.dx : a.dy, tol);
  for (int b = 0; b < a.batch; ++b) {
    TriArgs<R> ab = a;
    ab.batch = 1;
    ab.field = a.field + b * cells;
    ab.kvec = a.kvec + b;
    ab.khost = a.khost + b;
    cx<R> *bb = reinterpret_cast<cx<R> *>(s0 + per * b);
    cx<double> *wb = reinterpret_cast<cx<double> *>(s0 + per * b + fbytes);
    const cx<R> *prev = implicit_x ? ab.field : bb;
    cx<R> *next = implicit_x ? bb : ab.field;
    if (iters > 0) {
      hipError_t e = launch_adi_jacobi<R>(ab, implicit_x, prev, next, iters, s);
      if (e != hipSuccess) return e;
      continue;
    }
    const dim3 g1((a.n + 255) / 256, 1);
    if (implicit_x)
      hipLaunchKernelGGL((k_adi_line<R, true>), g1, dim3(256), 0, s, ab, prev, next, wb, per);
    else
      hipLaunchKernelGGL((k_adi_line<R, false>), g1, dim3(256), 0, s, ab, prev, next, wb, per);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------- Thomas
// One lane per system, the reference's sweep: sup = sub, sup[0] /= diag[0],
// rhs[0] /= diag[0], then sup[i] /= (diag[i] - sub[i-1] sup[i-1]),
// rhs[i] = (rhs[i] - sub[i-1] rhs[i-1]) / (same), back substitution into diag.
__global__ void k_thomas(int layout, int nsys, int n, cx<double> *diag, const cx<double> *sub,
                         cx<double> *rhs, cx<double> *work)
{
  const int sy = blockIdx.x * blockDim.x + threadIdx.x;
  if (sy >= nsys) return;
  auto ix = [&](int i) -> size_t {
    return layout == 0 ? static_cast<size_t>(sy) * n + i : static_cast<size_t>(i) * nsys + sy;
  };
  auto sx = [&](int i) -> size_t {
    return layout == 0 ? static_cast<size_t>(sy) * (n - 1) + i : static_cast<size_t>(i) * nsys + sy;
  };
  if (n == 1) {
    rhs[ix(0)] = cdiv(rhs[ix(0)], diag[ix(0)]);
    diag[ix(0)] = rhs[ix(0)];
    return;
  }
  cx<double> supp = cdiv(sub[sx(0)], diag[ix(0)]);
  cx<double> rp = cdiv(rhs[ix(0)], diag[ix(0)]);
  work[ix(0)] = supp;
  rhs[ix(0)] = rp;
  for (int i = 1; i < n; ++i) {
    const cx<double> s1 = sub[sx(i - 1)];
    const cx<double> den = diag[ix(i)] - s1 * supp;
    if (i < n - 1) supp = cdiv(sub[sx(i)], den);
    rp = cdiv(rhs[ix(i)] - s1 * rp, den);
    work[ix(i)] = supp;
    rhs[ix(i)] = rp;
  }
  cx<double> x = rp;
  diag[ix(n - 1)] = x;
  for (int i = n - 2; i >= 0; --i) {
    x = rhs[ix(i)] - work[ix(i)] * x;
    diag[ix(i)] = x;
  }
}

hipError_t launch_thomas_batched(int layout, int nsys, int n, cx<double> *diag,
                                 const cx<double> *sub, cx<double> *rhs, cx<double> *work,
                                 hipStream_t s)
{
  hipLaunchKernelGGL(k_thomas, dim3((nsys + 127) / 128), dim3(128), 0, s, layout, nsys, n, diag,
                     sub, rhs, work);
  return hipGetLastError();
}

int cn_reg_rows() { return CN_REG_ROWS; }
template hipError_t launch_cn_fixed<float>(const CnJob *, int, int, cx<float> *, cx<float> *, unsigned,
                                           unsigned, hipStream_t);
template hipError_t launch_cn_fixed<double>(const CnJob *, int, int, cx<double> *, cx<double> *, unsigned,
                                            unsigned, hipStream_t);
template hipError_t launch_cn_reg<float>(const CnJob *, int, int, cx<float> *, cx<float> *, unsigned, unsigned,
                                         hipStream_t);
template hipError_t launch_cn_reg<double>(const CnJob *, int, int, cx<double> *, cx<double> *, unsigned,
                                          unsigned, hipStream_t);
template hipError_t launch_cn_big<float>(const CnJob *, int, int, cx<float> *, void *, cx<float> *, unsigned,
                                         unsigned, hipStream_t);
template hipError_t launch_cn_big<double>(const CnJob *, int, int, cx<double> *, void *, cx<double> *, unsigned,
                                          unsigned, hipStream_t);
template hipError_t launch_cn_jobs<float>(const CnJob *, int, int, cx<float> *, void *, cx<float> *,
                                          unsigned, unsigned, hipStream_t);
template hipError_t launch_cn_jobs<double>(const CnJob *, int, int, cx<double> *, void *, cx<double> *,
                                           unsigned, unsigned, hipStream_t);
template hipError_t launch_adi_half<float>(const TriArgs<float> &, bool, hipStream_t);
template hipError_t launch_adi_half<double>(const TriArgs<double> &, bool, hipStream_t);

}  // namespace pxg
