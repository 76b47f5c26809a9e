// Per-step reductions over a stored 2D solution: the trapezoid sums of
// post::Transmittivity::trapezoidalIntegrateIntensityZ (src/transmittivity.cpp:30-42),
// which CurvedWaveGuideFD::solve evaluates after every z-step
// (src/curvedWaveGuide2D.cpp:207-219).  The march keeps every column of the
// solution on the device, so all steps reduce in one launch after it, one
// work-group per step, instead of a reduction inside each dependent step.
#include "pxg_internal.hpp"

namespace pxg {

__device__ __forceinline__ double intensity_at(const void *p, int c64, size_t i)
{
  const cd u = c64 ? cvt<double>(static_cast<const cf *>(p)[i]) : static_cast<const cd *>(p)[i];
  const double h = hypot(u.x, u.y);  // pow(abs(u), 2) of the reference
  return h * h;
}

// sums[i] = |u(s)|^2 + |u(e)|^2 + 2 sum_{s<j<e} |u(j)|^2 of column steps[i]
__global__ __launch_bounds__(256) void k_trapz_intensity(const void *sol, int c64, int nx,
                                                         const int *steps, const int *ixs,
                                                         const int *ixe, double *sums)
{
  __shared__ double part[256];
  const int i = blockIdx.x, t = threadIdx.x;
  const size_t base = static_cast<size_t>(steps[i]) * nx;
  const int s = ixs[i], e = ixe[i];
  double acc = 0.0;
  for (int j = s + 1 + t; j < e; j += 256) acc += 2.0 * intensity_at(sol, c64, base + j);
  part[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) part[t] += part[t + w];
    __syncthreads();
  }
  if (t == 0) {
    sums[i] = intensity_at(sol, c64, base + s) + intensity_at(sol, c64, base + e) + part[0];
  }
}

hipError_t launch_trapz_intensity(const void *sol, int c64, int nx, int n, const int *steps,
                                  const int *ixs, const int *ixe, double *sums, hipStream_t s)
{
  hipLaunchKernelGGL(k_trapz_intensity, dim3(n), dim3(256), 0, s, sol, c64, nx, steps, ixs, ixe, sums);
  return hipGetLastError();
}

}  // namespace pxg
