// Tridiagonal family: Crank-Nicolson march, ADI half-steps and the batched
// Thomas entry point.  Internal to libpaxgpu.
#pragma once

#include "../../include/paxgpu.h"
#include "pxg_internal.hpp"

namespace pxg {

template <class R>
struct TriArgs {
  cx<R> *field;          // batch of fields (CN: nx each; ADI: nx*nx each)
  int n;                 // nx
  int batch;
  const double *kvec;    // [batch]
  MatDev mat;
  double xmin, dx, ymin, dy, zmin, dz;
  int bc;                // CN boundary condition
  int use_tbc;           // ADI transparent boundary
  unsigned quirks;
  unsigned absorb_npix;  // CN edge absorber extension (0 = off)
  double absorb_inv;
  void *scratch;         // library-owned work space (tridiag_scratch_bytes)
  cx<R> *sol;            // CN: solution columns (nx x nz per field) or null
  size_t sol_fstride;
  int nz;
  unsigned first_step, nsteps;
  double kmin;           // smallest wavenumber of the batch (largest Jacobi contraction)
  const double *khost;   // [batch] wavenumbers in host memory (launch-time constants)
  int adi_tile;          // ADI: tile shape of the half-step launched (adi_tile_count() of each)
};

size_t tridiag_scratch_bytes(const pxg_problem &p, bool mixed);

// Transparent boundary factor of the reference, e = exp(i kd) with
// kd = -i log(r), r = u_out / u_in, and Re(kd) = arg(r) clamped at 0
// (src/crancNicholson.cpp:131-141, src/alternatingDirectionSolver.cpp:166-230).
// exp(i kd) = |r| e^{i arg r} is r itself when arg r >= 0, and |r| when
// arg r < 0 (the clamp removes the phase), so no logarithm or exponential is
// needed: arg r < 0 exactly when atan2(Im r, Re r) < 0, i.e. Im r < 0, or
// Im r = -0 with Re r < 0.  Returns false when |u_in| <= thresh (the reference
// then leaves the row unchanged).
// Every rounding is spelled out (one product rounded, then one fused
// multiply-add), so each kernel that inlines it gives the same bits.
__device__ __forceinline__ bool tbc_exp_ikd(cx<double> u_out, cx<double> u_in, double thresh, cx<double> &e)
{
#pragma clang fp contract(off)
  const double n2 = fma(u_in.x, u_in.x, u_in.y * u_in.y);
  if (!(sqrt(n2) > thresh)) return false;
  const double in2 = 1.0 / n2;
  const cx<double> r = {fma(u_out.x, u_in.x, u_out.y * u_in.y) * in2,
                        fma(u_out.y, u_in.x, -(u_out.x * u_in.y)) * in2};
  const bool clamp = r.y < 0.0 || (r.y == 0.0 && signbit(r.y) && r.x < 0.0);
  e = clamp ? cx<double>{sqrt(fma(r.x, r.x, r.y * r.y)), 0.0} : r;
  return true;
}


// One CN job: a field with its own grid, wavenumber and material, solved by
// one work-group (k_cn_jobs).  Offsets index the plan's field, solution and
// scratch buffers.
struct CnJob {
  int nx, P, m, bc;
  int nz;
  unsigned absorb_npix;
  size_t foff;   // elements into the field buffer
  size_t soff;   // elements into the solution buffer, ~0 when not stored
  size_t scoff;  // bytes into the scratch buffer
  double xmin, dx, zmin, dz, k;
  double absorb_inv;
  MatDev mat;
  // a ParaxialEquation other than the default: per plane 5 values per row,
  // F(x,z), G(x,z), H(x + dx/2, z), H(x - dx/2, z), J(x,z), plane p at
  // z = zmin + p dz (eq_inv: one plane for every z); null for the default
  const double *eq;
  int eq_first, eq_inv;
};
// Scratch bytes of one CN job of nx rows (256-byte aligned blocks).
size_t cn_job_scratch(int nx, int esz);
template <class R>
hipError_t launch_cn_jobs(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, void *scratch,
                          cx<R> *sol, unsigned first, unsigned nsteps, hipStream_t s);
// Register-resident march for jobs whose largest chunk has at most
// cn_reg_rows() rows (CnJob.m); needs no scratch.
template <class R>
hipError_t launch_cn_reg(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, cx<R> *sol,
                         unsigned first, unsigned nsteps, hipStream_t s);
int cn_reg_rows();
// Register-resident march with the matrix factored once (z-invariant material:
// planes or none), boundary terms by the Woodbury identity: reg-class jobs of
// at most cn_fixed_max_p() lanes.
template <class R>
hipError_t launch_cn_fixed(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, cx<R> *sol,
                           unsigned first, unsigned nsteps, hipStream_t s);
int cn_fixed_max_p();
// Register field, scratch coefficients: jobs whose chunks have more than
// cn_reg_rows() and at most cn_big_rows() rows.
template <class R>
hipError_t launch_cn_big(const CnJob *jobs_dev, int njobs, int max_P, cx<R> *fields, void *scratch, cx<R> *sol,
                         unsigned first, unsigned nsteps, hipStream_t s);
int cn_big_rows();
template <class R>
hipError_t launch_adi_half(const TriArgs<R> &a, bool implicit_x, hipStream_t s);
// Tile-Jacobi ADI half-step for diagonally dominant systems (pxg_adi.hip).
// Sweeps so that rho^K < tol (0 = not dominant enough for the tile path).
int adi_jacobi_iters(double dz, double k, double d, double tol);
// Tile shapes of each ADI half-step kernel (indices 0 .. adi_tile_count() - 1)
int adi_tile_count();
template <class R>
hipError_t launch_adi_jacobi(const TriArgs<R> &a, bool implicit_x, const cx<R> *prev, cx<R> *next,
                             int iters, hipStream_t s);
hipError_t launch_thomas_batched(int layout, int nsys, int n, cx<double> *diag,
                                 const cx<double> *sub, cx<double> *rhs, cx<double> *work,
                                 hipStream_t s);

// CN partition: P chunks of m rows (the last chunk may be shorter but keeps
// at least two rows); P is a power of two so the reduced system of 2P rows
// suits parallel cyclic reduction.
struct CnPartition {
  int P, m;
};
CnPartition cn_partition(int n);

}  // namespace pxg
