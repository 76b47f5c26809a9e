// Phase-screen factors shared by the split-step kernels (pxg_fft3d.hip) and
// their Bluestein variants for other line lengths (pxg_bluestein.hip).
#pragma once

#include "pxg_fp64math.hpp"
#include "pxg_internal.hpp"

namespace pxg {

// Phase-screen factor exp(-k (beta + i delta) dz) / N^2 from FP64 material
// values, in FP64 (src/fftSolver3D.cpp:154).  Its callers are rare per-cell
// loops; exp_fp64/sincos_fp64 keep it inline without many registers.
__device__ __forceinline__ cd screen_from_values(double k, double dz, double delta, double beta,
                                                   double inv_norm)
{
  const double amp = exp_fp64((-k * beta) * dz);
  double s, c;
  sincos_fp64((-k * delta) * dz, s, c);
  return {c * amp * inv_norm, s * amp * inv_norm};
}

// FFTSolver2D's screen of cell i (src/fftSolver2D.cpp:36-51): the material at
// (x, 0, z) with z = z_step + dz/2, by CSG (parts culled into mask), curved
// guide or host-sampled plane.
__device__ __forceinline__ cd screen_2d(const MatDev &m, const uint64_t *mask, int nw, unsigned step,
                                     int i, double x, double z, double k, double dz, double inv_norm)
{
  double d = 0.0, be = 0.0;
  if (m.mode == MAT_CSG) {
    part_values(m, csg_eval(m, mask, nw, x, 0.0, z), d, be);
  } else if (m.mode == MAT_CURVED) {
    curved_material(m.wg, x, z, d, be);
  } else if (m.mode == MAT_PLANES) {
    const int p = m.z_invariant ? 0 : static_cast<int>(step) - m.first_plane;
    const double *pl = m.planes + static_cast<size_t>(p) * m.plane_stride;
    d = pl[2 * i];
    be = pl[2 * i + 1];
  }
  return screen_from_values(k, dz, d, be, inv_norm);
}

}  // namespace pxg
