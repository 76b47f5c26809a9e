// Split-step passes for line lengths without a power-of-two kernel: FFTW plans
// any length (src/fftSolver3D.cpp:61-62, src/fftSolver2D.cpp:68-69), and the
// reference's truncating node count makes lengths such as 511 ordinary
// (src/paraxialSimulation.cpp:107-120).
//
// An N-point DFT with sign S, X_k = sum_j x_j e^{S 2 pi i jk/N}, is computed as
// Bluestein's chirp convolution (Bluestein 1970; Rabiner, Schafer & Rader 1969):
// with jk = (j^2 + k^2 - (k - j)^2)/2 and c_j = e^{-i pi j^2/N},
//   S = -1:  X_k = c_k   sum_j (x_j c_j)      conj(c)_{k-j}
//   S = +1:  X_k = c*_k  sum_j (x_j conj(c_j)) c_{k-j}
// a linear convolution of length 2N - 1, done cyclically on M >= 2N - 1 points
// (a power of two) with the LDS Stockham FFT of pxg_fft.hpp: FFT_M, times the
// transformed kernel (an FP64 table with the 1/M folded in), FFT_M^{-1}.  The
// chirp angles pi j^2/N are reduced exactly (j^2 mod 2N) on the host.  A line is
// one work-group of M/8 lanes holding the M points in canonical order
// (positions t + m M/8); positions >= N are zero on input and discarded on
// output.  Each pass performs the operations of its power-of-two counterpart
// (pxg_fft3d.hip) in the same order, with every N-point DFT replaced by this one:
//   pass Y:  IFFT_y -> phase screen (material map, planes) -> absorber -> FFT_y
//   pass X:  FFT_x -> x Kx Ky -> IFFT_x
//   FFT2D:   FFT -> x Kx -> IFFT -> screen at z + dz/2 (+ solution column)
// Four M-point transforms per N-point pair, with M between 2N and 4N: several
// times the work of a power-of-two line.  These kernels favour simplicity: one
// line per work-group, strided x-lines read as they lie.
#include "pxg_fft.hpp"
#include "pxg_fp64math.hpp"
#include "pxg_internal.hpp"
#include "pxg_screen.hpp"

namespace pxg {

int blue_length(int n)
{
  int m = 16;
  while (m < 2 * n - 1) m *= 2;
  return m;
}

// N-point DFT of sign S of the canonical registers of an M-point line whose
// positions >= n are zero; on return positions < n hold the DFT and positions
// >= n are zero.  Contains LDS synchronisation: every lane of the work-group
// calls it.
template <int M, int S, bool WL>
__device__ __forceinline__ void blue_dft(cd (&v)[8], cd *lds, int t, const TwBase<M, double> &twb, int n,
                                         const cd *__restrict__ chirp, const cd *__restrict__ kern)
{
  constexpr int T = M / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * T;
    if (j < n) v[m] = v[m] * (S < 0 ? chirp[j] : conj(chirp[j]));
  }
  fft_line<M, -1, double, WL>(v, lds, t, twb);
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = v[m] * kern[t + m * T];
  fft_line<M, +1, double, WL>(v, lds, t, twb);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * T;
    v[m] = j < n ? v[m] * (S < 0 ? chirp[j] : conj(chirp[j])) : cd{0.0, 0.0};
  }
}

template <int M, class R>
__device__ __forceinline__ void load_line(cd (&v)[8], const cx<R> *f, size_t stride, int n, int t)
{
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * (M / 8);
    v[m] = j < n ? cvt<double>(f[static_cast<size_t>(j) * stride]) : cd{0.0, 0.0};
  }
}

template <int M, class R>
__device__ __forceinline__ void store_line(const cd (&v)[8], cx<R> *f, size_t stride, int n, int t)
{
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * (M / 8);
    if (j < n) f[static_cast<size_t>(j) * stride] = cvt<R>(v[m]);
  }
}

// Pass Y: column c (y-line, contiguous) of field b.
template <int M, class R>
__global__ __launch_bounds__(M / 8) void k_blue_pass_y(FftArgs<R> a, int n)
{
  constexpr int T = M / 8;
  constexpr bool WL = T <= 64;
  __shared__ cd lds[LdsLine<M>::STRIDE];
  const int t = threadIdx.x, c = blockIdx.x, b = blockIdx.y;
  cx<R> *f = a.field + b * a.fstride + static_cast<size_t>(c) * n;
  const TwBase<M, double> twb = tw_base<M>(a.tw, t);
  cd v[8];
  load_line<M>(v, f, 1, n, t);
  if (a.mode & Y_INV) {
    blue_dft<M, +1, WL>(v, lds, t, twb, n, a.chirp, a.bi);
    // the screen, one cell at a time through the lane's own LDS slots (the
    // out-of-line FP64 exp/sincos of border cells and planes stay off the
    // registers of the transform)
#pragma unroll
    for (int m = 0; m < 8; ++m) lds[lds_pos(t + m * T)] = v[m];
    const double k = a.kvec[b];
#pragma unroll 1
    for (int m = 0; m < 8; ++m) {
      const int r = t + m * T;
      if (r >= n) continue;
      cd fac{a.inv_norm, 0.0};
      if (a.mat.mode == MAT_CSG) {
        const size_t cell = static_cast<size_t>(c) * n + r;
        const unsigned code = a.mcode[cell];
        if (code & MAP_CROSSING) {
          const cd db = a.mxdb[cell];
          fac = screen_from_values(k, a.dz, db.x, db.y, a.inv_norm);
        } else {
          fac = static_cast<const cd *>(a.mat.part_fac)[b * (a.mat.n_parts + 1) + code];
        }
      } else if (a.mat.mode == MAT_PLANES) {
        const int p = a.mat.z_invariant ? 0 : static_cast<int>(a.step) - a.mat.first_plane;
        const double *pl = a.mat.planes + static_cast<size_t>(p) * a.mat.plane_stride +
                           2 * (static_cast<size_t>(c) * n + r);
        fac = screen_from_values(k, a.dz, pl[0], pl[1], a.inv_norm);
      }
      cd &slot = lds[lds_pos(r)];
      slot = slot * fac;
      if (a.absy) slot = scale(scale(slot, a.absy[r]), a.absx[c]);
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = lds[lds_pos(t + m * T)];
  }
  if (a.mode & Y_FWD) blue_dft<M, -1, WL>(v, lds, t, twb, n, a.chirp, a.bf);
  store_line<M>(v, f, 1, n, t);
}

// Pass X: row r (x-line, stride n) of field b.
template <int M, class R>
__global__ __launch_bounds__(M / 8) void k_blue_pass_x(FftArgs<R> a, int n)
{
  constexpr int T = M / 8;
  constexpr bool WL = T <= 64;
  __shared__ cd lds[LdsLine<M>::STRIDE];
  const int t = threadIdx.x, r = blockIdx.x, b = blockIdx.y;
  cx<R> *f = a.field + b * a.fstride + r;
  const TwBase<M, double> twb = tw_base<M>(a.tw, t);
  cd v[8];
  load_line<M>(v, f, n, n, t);
  blue_dft<M, -1, WL>(v, lds, t, twb, n, a.chirp, a.bf);
  const cd kyr = a.ky[b * n + r];
  const cd *kx = a.kx + b * n;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int c = t + m * T;
    if (c < n) v[m] = v[m] * (kx[c] * kyr);
  }
  blue_dft<M, +1, WL>(v, lds, t, twb, n, a.chirp, a.bi);
  store_line<M>(v, f, n, n, t);
}

// FFTSolver2D step: field b is one line.
template <int M, class R>
__global__ __launch_bounds__(M / 8) void k_blue_fft2d(FftArgs<R> a, int n)
{
  constexpr int T = M / 8;
  constexpr bool WL = T <= 64;
  __shared__ cd lds[LdsLine<M>::STRIDE];
  __shared__ uint64_t mask[16];
  const int t = threadIdx.x, b = blockIdx.x;
  cx<R> *f = a.field + b * a.fstride;
  const TwBase<M, double> twb = tw_base<M>(a.tw, t);
  cd v[8];
  load_line<M>(v, f, 1, n, t);
  blue_dft<M, -1, WL>(v, lds, t, twb, n, a.chirp, a.bf);
  const cd *kx = a.kx + b * n;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int i = t + m * T;
    if (i < n) v[m] = v[m] * kx[i];
  }
  blue_dft<M, +1, WL>(v, lds, t, twb, n, a.chirp, a.bi);
  const double k = a.kvec[b];
  const double z = coord(a.zmin, a.step, a.dz) + 0.5 * a.dz;
  const int nw = a.mat.mode == MAT_CSG ? (a.mat.n_parts + 63) >> 6 : 0;
  if (a.mat.mode == MAT_CSG) csg_cull(a.mat, mask, nw, a.xmin, coord(a.xmin, n - 1, a.dx), 0.0, 0.0, z, z);
#pragma unroll 1
  for (int m = 0; m < 8; ++m) {
    const int i = t + m * T;
    if (i < n) lds[lds_pos(i)] = screen_2d(a.mat, mask, nw, a.step, i, coord(a.xmin, i, a.dx), z, k, a.dz, a.inv_norm);
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int i = t + m * T;
    if (i < n) v[m] = v[m] * lds[lds_pos(i)];
  }
  store_line<M>(v, f, 1, n, t);
  if (a.sol) store_line<M>(v, a.sol + b * a.sol_fstride + static_cast<size_t>(a.step) * n, 1, n, t);
}

#define PXG_BLUE_SWITCH(KER, GRID)                                             \
  switch (blue_length(n)) {                                                    \
    case 16: hipLaunchKernelGGL((KER<16, R>), GRID, dim3(2), 0, s, a, n); break;     \
    case 32: hipLaunchKernelGGL((KER<32, R>), GRID, dim3(4), 0, s, a, n); break;     \
    case 64: hipLaunchKernelGGL((KER<64, R>), GRID, dim3(8), 0, s, a, n); break;     \
    case 128: hipLaunchKernelGGL((KER<128, R>), GRID, dim3(16), 0, s, a, n); break;  \
    case 256: hipLaunchKernelGGL((KER<256, R>), GRID, dim3(32), 0, s, a, n); break;  \
    case 512: hipLaunchKernelGGL((KER<512, R>), GRID, dim3(64), 0, s, a, n); break;  \
    case 1024: hipLaunchKernelGGL((KER<1024, R>), GRID, dim3(128), 0, s, a, n); break; \
    case 2048: hipLaunchKernelGGL((KER<2048, R>), GRID, dim3(256), 0, s, a, n); break; \
    case 4096: hipLaunchKernelGGL((KER<4096, R>), GRID, dim3(512), 0, s, a, n); break; \
    case 8192: hipLaunchKernelGGL((KER<8192, R>), GRID, dim3(1024), 0, s, a, n); break; \
    default: return hipErrorInvalidValue;                                      \
  }                                                                            \
  return hipGetLastError();

template <class R>
hipError_t launch_blue_pass_y(int n, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (n < 2 || n > 4096 || a.blue_m != blue_length(n)) return hipErrorInvalidValue;
  PXG_BLUE_SWITCH(k_blue_pass_y, dim3(n, batch))
}

template <class R>
hipError_t launch_blue_pass_x(int n, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (n < 2 || n > 4096 || a.blue_m != blue_length(n)) return hipErrorInvalidValue;
  PXG_BLUE_SWITCH(k_blue_pass_x, dim3(n, batch))
}

template <class R>
hipError_t launch_blue_fft2d(int n, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (n < 2 || n > 4096 || a.blue_m != blue_length(n)) return hipErrorInvalidValue;
  PXG_BLUE_SWITCH(k_blue_fft2d, dim3(batch))
}

template hipError_t launch_blue_pass_y<float>(int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_blue_pass_y<double>(int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_blue_pass_x<float>(int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_blue_pass_x<double>(int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_blue_fft2d<float>(int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_blue_fft2d<double>(int, const FftArgs<double> &, int, hipStream_t);

}  // namespace pxg
