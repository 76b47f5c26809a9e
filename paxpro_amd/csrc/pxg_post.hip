// Post-processing on the device.
//
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

// ---------------------------------------------------------------------------
// Field quantities of the post-processing modules (src/postProcessMod.cpp:7-151):
// Intensity |u| (arma::abs), ExitIntensity |u|^2 (pow(abs(u), 2)), Phase and
// ExitPhase arg u (atan2(Im, Re)), ExitField Re u.  One cell per lane,
// grid-stride; the launch reads s bytes and writes 8 bytes per cell.
__device__ __forceinline__ cd cell_at(const void *p, int c64, size_t i)
{
  return c64 ? cvt<double>(static_cast<const cf *>(p)[i]) : static_cast<const cd *>(p)[i];
}

__global__ __launch_bounds__(256) void k_field_quantity(const void *src, int c64, size_t n, int q,
                                                        double *out)
{
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    const cd u = cell_at(src, c64, i);
    double r;
    if (q == 0) {
      r = hypot(u.x, u.y);
    } else if (q == 1) {
      const double h = hypot(u.x, u.y);
      r = h * h;
    } else if (q == 2) {
      r = atan2(u.y, u.x);
    } else {
      r = u.x;
    }
    out[i] = r;
  }
}

// Per work-group maximum and minimum of |u| (Armadillo's max() / min() of a
// complex cube compare absolute values): part[blk] = max, part[G + blk] = min.
__global__ __launch_bounds__(256) void k_abs_extrema(const void *src, int c64, size_t n, double *part)
{
  __shared__ double smax[256], smin[256];
  const int t = threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  double mx = -HUGE_VAL, mn = HUGE_VAL;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + t; i < n; i += stride) {
    const cd u = cell_at(src, c64, i);
    const double h = hypot(u.x, u.y);
    mx = fmax(mx, h);
    mn = fmin(mn, h);
  }
  smax[t] = mx;
  smin[t] = mn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      smax[t] = fmax(smax[t], smax[t + w]);
      smin[t] = fmin(smin[t], smin[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x] = smax[0];
    part[gridDim.x + blockIdx.x] = smin[0];
  }
}

// IntensityUint8 (src/postProcessMod.cpp:17-30): 255 |u| / max|u|.
// LogIntensityUint8 (src/postProcessMod.cpp:32-76): log|u| mapped from
// [bot, top] to [0, 255], clamped outside; top = log max|u| (or the set maximum
// when that is smaller), bot = log min|u| (or the set minimum when larger).
// Values are converted to unsigned char by truncation, as the reference's
// assignment does; a value that is not a number (a field of zeros, top = bot)
// is stored as 0, where the reference's conversion is undefined.
// Every work-group first reduces the G partial extrema of k_abs_extrema.
__global__ __launch_bounds__(256) void k_field_uint8(const void *src, int c64, size_t n,
                                                     const double *part, int G, int log_scale,
                                                     int lo_set, double lo, int hi_set, double hi,
                                                     unsigned char *out)
{
  __shared__ double smax[256], smin[256];
  const int t = threadIdx.x;
  double mx = -HUGE_VAL, mn = HUGE_VAL;
  for (int j = t; j < G; j += 256) {
    mx = fmax(mx, part[j]);
    mn = fmin(mn, part[G + j]);
  }
  smax[t] = mx;
  smin[t] = mn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      smax[t] = fmax(smax[t], smax[t + w]);
      smin[t] = fmin(smin[t], smin[t + w]);
    }
    __syncthreads();
  }
  const double maxabs = smax[0], minabs = smin[0];
  double top = 0.0, bot = 0.0;
  if (log_scale) {
    const double lmax = log(maxabs), lmin = log(minabs);
    top = hi_set ? (lmax > hi ? hi : lmax) : lmax;
    bot = lo_set ? (lmin < lo ? lo : lmin) : lmin;
  }
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + t; i < n; i += stride) {
    const cd u = cell_at(src, c64, i);
    const double h = hypot(u.x, u.y);
    double v;
    if (!log_scale) {
      v = 255.0 * h / maxabs;
    } else {
      const double l = log(h);
      if (l > top)
        v = 255.0;
      else if (l < bot)
        v = 0.0;
      else
        v = 255.0 * (l - bot) / (top - bot);
    }
    out[i] = (v >= 0.0 && v < 256.0) ? static_cast<unsigned char>(v) : static_cast<unsigned char>(0);
  }
}

namespace {
int grid_for(size_t n)
{
  const size_t g = (n + 255) / 256;
  return static_cast<int>(g < 4096 ? (g > 0 ? g : 1) : 4096);
}
}  // namespace

int extrema_groups(size_t n)
{
  const int g = grid_for(n);
  return g < 1024 ? g : 1024;
}

hipError_t launch_field_quantity(const void *src, int c64, size_t n, int q, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_field_quantity, dim3(grid_for(n)), dim3(256), 0, s, src, c64, n, q, out);
  return hipGetLastError();
}

hipError_t launch_field_uint8(const void *src, int c64, size_t n, double *part, int log_scale,
                              int lo_set, double lo, int hi_set, double hi, unsigned char *out,
                              hipStream_t s)
{
  const int G = extrema_groups(n);
  hipLaunchKernelGGL(k_abs_extrema, dim3(G), dim3(256), 0, s, src, c64, n, part);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_field_uint8, dim3(grid_for(n)), dim3(256), 0, s, src, c64, n,
                     static_cast<const double *>(part), G, log_scale, lo_set, lo, hi_set, hi, out);
  return hipGetLastError();
}

}  // namespace pxg
