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
};

size_t tridiag_scratch_bytes(const pxg_problem &p, bool mixed);

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
template <class R>
hipError_t launch_adi_half(const TriArgs<R> &a, bool implicit_x, hipStream_t s);
// Tile-Jacobi ADI half-step for diagonally dominant systems (pxg_adi.hip).
// Sweeps so that rho^K < tol (0 = not dominant enough for the tile path).
int adi_jacobi_iters(double dz, double k, double d, double tol);
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
