// Far-field intensity of a 3D exit field: post::FarField::result
// (src/farFieldPost.cpp:141-230) as GenericScattering uses it
// (src/genericScattering.cpp:148-213,235-238).
//
//   D = field - reference (the vacuum run), n x n, column-major (rows = y)
//   column pass: every column zero-padded at its end to the signal length M,
//                forward FFT (unnormalised), fftshift, rows of the y window kept
//   row pass:    every row of that, conjugated (Armadillo's t() of a complex
//                row, src/farFieldPost.cpp:227), zero-padded to M, forward FFT,
//                fftshift, columns of the x window kept, |.|^2 / M
//   resize:      block means down to the export size
//                (PostProcessingModule::resizeMatrix, src/postProcessing.cpp:29-49)
//
// A line holds at most n non-zero values, so its M-point DFT is computed as
// D = M/NP transforms of NP = nextpow2(n) points: with f = D g + h,
//   X[f] = sum_j x_j w_M^{h j} w_NP^{g j},
// i.e. residue h pre-multiplies the line by w_M^{h j} and one NP-point FFT gives
// every f = h (mod D).  Each (line, h) is one work-group running the LDS Stockham
// FFT of the split-step passes; it writes only the outputs inside the window.
// Pad lengths far beyond the field (the reference's examples use 1024 and 32768
// for 512- and 1024-point fields) then need no transform longer than the field's.
// Everything after the field load is FP64.
//
// Constant padding (post::FarField::constantPadding, src/farFieldPost.cpp:314-321)
// fills entries n .. M-1 of every line with its entry 0, x_0.  Its spectrum is
// the zero-padded one plus x_0 G(f), G(f) = sum_{j=n}^{M-1} w_M^{f j}, a geometric
// sum: M - n at f = 0, else (w_M^{f n} - 1) / (1 - w_M^f).
//
// Linear-fit padding (post::FarField::linearPadding, src/farFieldPost.cpp:265-312)
// fits y = c0 + c1 x to the real and to the imaginary parts of entries 0..9 of a
// line (gsl_fit_linear: running means) twice: against x = linspace(0, x_10 - x_0)
// for the L = (M - n)/2 entries M-L .. M-1, valued a1 + b1 dxf (i - M + 1), and
// against x = linspace(x_{n-10}, x_n) for entries n .. n+L-1, valued
// a2 + b2 (x_n + (i - n) dxf); dxf is the first grid's step, x_i the transverse grid.
// (The second fit reads entries 0..9 again, as the reference does.)  Each segment
// is P + Q k over k = 0..L-1, so its spectrum is a phase times P S0 + Q S1 with
// S0 = sum_k w^{fk} = e^{i c t} D(t) and S1 = sum_k k w^{fk} = e^{i c t} (c D(t) - i D'(t)),
// t = -2 pi f / M, c = (L-1)/2 and D the Dirichlet kernel sin(L t/2) / sin(t/2).
//
// The reference forms the first segment's abscissa as dxf * (i - n_elem + 1) with
// an int i and Armadillo's unsigned 64-bit n_elem (src/farFieldPost.cpp:288), so the
// difference wraps: for j = M - 1 - i >= 1 it is the double nearest to 2^64 - j,
// i.e. 2^64 - 2048 q(j) with q(j) = j / 2048 rounded half to even (doubles in
// [2^63, 2^64) are 2048 apart), and 0 at i = M - 1.  That segment is therefore
// C = a + b dxf 2^64 on k = 0..L-2 and a at k = L-1, less b dxf 2048 q(j) on the
// runs of constant q, each a geometric sum (none while L <= 1024).  This is the
// default; PXG_FIX_LINEAR_PAD selects the signed ramp P + Q k above.
#include "pxg_fft.hpp"
#include "pxg_fp64math.hpp"
#include "pxg_internal.hpp"

namespace pxg {

__device__ __forceinline__ cd ff_load(const void *p, int c64, size_t i)
{
  return c64 ? cvt<double>(static_cast<const cf *>(p)[i]) : static_cast<const cd *>(p)[i];
}

// x_0 G(f) of constant padding, f < M
__device__ __forceinline__ cd ff_pad_term(const FfArgs &a, cd x0, long f)
{
  cd g;
  if (f == 0) {
    g = cd{static_cast<double>(a.M - a.n), 0.0};
  } else {
    const cd wfn = a.tw_m[(static_cast<size_t>(f) * a.n) % static_cast<size_t>(a.M)];
    const cd wf = a.tw_m[f];
    const cd num{wfn.x - 1.0, wfn.y};
    const cd den{1.0 - wf.x, -wf.y};
    const double inv = 1.0 / (den.x * den.x + den.y * den.y);
    g = cd{(num.x * den.x + num.y * den.y) * inv, (num.y * den.x - num.x * den.y) * inv};
  }
  return x0 * g;
}

struct LinFit {
  cd a, b;  // intercept and slope (real and imaginary parts fitted separately)
};

// gsl_fit_linear's running means over the 10 points y of one grid
__device__ __forceinline__ LinFit lin_fit(const double *dxk, double mx, double mdx2, const cd *y)
{
  double myr = 0.0, myi = 0.0;
  for (int k = 0; k < 10; ++k) {
    myr += (y[k].x - myr) / (k + 1.0);
    myi += (y[k].y - myi) / (k + 1.0);
  }
  double mxyr = 0.0, mxyi = 0.0;
  for (int k = 0; k < 10; ++k) {
    mxyr += (dxk[k] * (y[k].x - myr) - mxyr) / (k + 1.0);
    mxyi += (dxk[k] * (y[k].y - myi) - mxyi) / (k + 1.0);
  }
  const double br = mxyr / mdx2, bi = mxyi / mdx2;
  return {cd{myr - mx * br, myi - mx * bi}, cd{br, bi}};
}

// sin and cos of pi m / M for an integer m, reduced modulo 2M first
__device__ __forceinline__ void sincos_pi_frac(long long m, int M, double &s, double &c)
{
  const long long r = m % (2LL * M);
  sincos_fp64(3.141592653589793 * static_cast<double>(r) / M, s, c);
}

// sum_{k=0}^{m-1} w_M^{f k}, f in [0, M): m at f = 0, else (1 - w^{fm}) / (1 - w^f)
__device__ __forceinline__ cd ff_geometric(long f, long long m, int M)
{
  if (f == 0) return cd{static_cast<double>(m), 0.0};
  double s1, c1, sm, cm;
  sincos_pi_frac(2LL * f, M, s1, c1);                        // w^f = c1 - i s1
  sincos_pi_frac(2LL * f * m, M, sm, cm);                    // w^{fm}
  const cd num{1.0 - cm, sm};
  const cd den{1.0 - c1, s1};
  const double inv = 1.0 / (den.x * den.x + den.y * den.y);
  return cd{(num.x * den.x + num.y * den.y) * inv, (num.y * den.x - num.x * den.y) * inv};
}

// w_M^{x} for an integer x
__device__ __forceinline__ cd ff_w(long long x, int M)
{
  double s, c;
  sincos_pi_frac(2LL * x, M, s, c);
  return cd{c, -s};
}

// sum_{k=0}^{L-2} q(L-1-k) w^{fk} of the wrapped first segment: over the runs
// q >= 1 of j = L-1-k (q even: j in [2048q - 1024, 2048q + 1024], odd: the open
// range), each clipped to j <= L-1
__device__ __forceinline__ cd ff_wrap_staircase(long f, int L, int M)
{
  cd t{0.0, 0.0};
  for (long long q = 1;; ++q) {
    const long long jlo = 2048 * q - 1024 + (q & 1), jhi = 2048 * q + 1024 - (q & 1);
    if (jlo > L - 1) break;
    const long long k0 = (L - 1) - std::min<long long>(jhi, L - 1), k1 = (L - 1) - jlo;
    t = t + cd{static_cast<double>(q), 0.0} * ff_w(static_cast<long long>(f) * k0, M) *
                ff_geometric(f, k1 - k0 + 1, M);
  }
  return t;
}

// spectrum at f of the two linear-fit segments of a line
__device__ __forceinline__ cd ff_linear_term(const FfArgs &a, const LinFit &B, const LinFit &E, long f)
{
  const int L = a.lin_L;
  cd S0, S1;
  if (f == 0) {
    S0 = cd{static_cast<double>(L), 0.0};
    S1 = cd{0.5 * L * (L - 1.0), 0.0};
  } else {
    double sh, ch, sl, cl, se, ce;
    sincos_pi_frac(f, a.M, sh, ch);                                  // sin(t/2) = -sh
    sincos_pi_frac(static_cast<long long>(f) * L, a.M, sl, cl);      // sin(L t/2) = -sl
    sincos_pi_frac(static_cast<long long>(f) * (L - 1), a.M, se, ce);  // e^{i c t} = ce - i se
    const double D = sl / sh;
    const double Dp = (0.5 * L * cl * (-sh) + 0.5 * sl * ch) / (sh * sh);
    const cd e{ce, -se};
    const double chalf = 0.5 * (L - 1.0);
    S0 = e * cd{D, 0.0};
    S1 = e * cd{chalf * D, -Dp};
  }
  const cd PE = E.a + E.b * cd{a.lin_x2end, 0.0};
  const cd QE = E.b * cd{a.lin_dxf, 0.0};
  const cd phB = a.tw_m[(static_cast<size_t>(f) * static_cast<size_t>(a.M - L)) % static_cast<size_t>(a.M)];
  const cd phE = a.tw_m[(static_cast<size_t>(f) * a.n) % static_cast<size_t>(a.M)];
  cd segB;
  if (a.lin_wrap) {
    const cd C = B.a + B.b * cd{a.lin_dxf * 18446744073709551616.0, 0.0};   // 2^64
    const cd wl = ff_w(static_cast<long long>(f) * (L - 1), a.M);
    segB = C * (S0 - wl) + B.a * wl;
    if (L > 1025) segB = segB - B.b * cd{a.lin_dxf * 2048.0, 0.0} * ff_wrap_staircase(f, L, a.M);
  } else {
    const cd PB = B.a + B.b * cd{a.lin_dxf * (1.0 - L), 0.0};
    const cd QB = B.b * cd{a.lin_dxf, 0.0};
    segB = PB * S0 + QB * S1;
  }
  return phB * segB + phE * (PE * S0 + QE * S1);
}

template <int NP>
__global__ __launch_bounds__(NP / 8) void k_ff_cols(FfArgs a)
{
  constexpr int T = NP / 8;
  constexpr bool WL = T <= 64;
  __shared__ cd lds[LdsLine<NP>::SIZE];
  const int t = threadIdx.x, c = blockIdx.x, h = blockIdx.y;
  const TwBase<NP, double> twb = tw_base<NP>(a.tw_np, t);
  cd v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * T;
    if (j < a.n) {
      const size_t i = static_cast<size_t>(c) * a.n + j;
      cd u = ff_load(a.field, a.field_c64, i);
      if (a.ref) u = u - ff_load(a.ref, a.ref_c64, i);
      if (a.D > 1) u = u * a.tw_m[(static_cast<size_t>(h) * j) & (a.M - 1)];
      v[m] = u;
    } else {
      v[m] = cd{0.0, 0.0};
    }
  }
  cd x0{0.0, 0.0};
  if (a.pad_const) {
    const size_t i0 = static_cast<size_t>(c) * a.n;
    x0 = ff_load(a.field, a.field_c64, i0);
    if (a.ref) x0 = x0 - ff_load(a.ref, a.ref_c64, i0);
  }
  LinFit fb{}, fe{};
  if (a.pad_linear) {
    cd y[10];
    for (int k = 0; k < 10; ++k) {
      const size_t i = static_cast<size_t>(c) * a.n + k;
      y[k] = ff_load(a.field, a.field_c64, i);
      if (a.ref) y[k] = y[k] - ff_load(a.ref, a.ref_c64, i);
    }
    fb = lin_fit(a.fit1_dx, a.fit1_mx, a.fit1_mdx2, y);
    fe = lin_fit(a.fit2_dx, a.fit2_mx, a.fit2_mdx2, y);
  }
  fft_line<NP, -1, double, WL, false>(v, lds, t, twb);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long f = static_cast<long>(a.D) * (t + m * T) + h;
    const int r = static_cast<int>((f + a.M / 2) & (a.M - 1)) - a.ylo;  // fftshift, window
    if (static_cast<unsigned>(r) < static_cast<unsigned>(a.ny))
      a.tmp[static_cast<size_t>(c) * a.ny + r] =
          a.pad_const ? v[m] + ff_pad_term(a, x0, f) : (a.pad_linear ? v[m] + ff_linear_term(a, fb, fe, f) : v[m]);
  }
}

template <int NP>
__global__ __launch_bounds__(NP / 8) void k_ff_rows(FfArgs a)
{
  constexpr int T = NP / 8;
  constexpr bool WL = T <= 64;
  __shared__ cd lds[LdsLine<NP>::SIZE];
  const int t = threadIdx.x, r = blockIdx.x, h = blockIdx.y;
  const TwBase<NP, double> twb = tw_base<NP>(a.tw_np, t);
  cd v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * T;
    if (j < a.n) {
      cd u = conj(a.tmp[static_cast<size_t>(j) * a.ny + r]);
      if (a.D > 1) u = u * a.tw_m[(static_cast<size_t>(h) * j) & (a.M - 1)];
      v[m] = u;
    } else {
      v[m] = cd{0.0, 0.0};
    }
  }
  const cd x0 = a.pad_const ? conj(a.tmp[r]) : cd{0.0, 0.0};  // entry 0 of the row
  LinFit fb{}, fe{};
  if (a.pad_linear) {
    cd y[10];
    for (int k = 0; k < 10; ++k) y[k] = conj(a.tmp[static_cast<size_t>(k) * a.ny + r]);
    fb = lin_fit(a.fit1_dx, a.fit1_mx, a.fit1_mdx2, y);
    fe = lin_fit(a.fit2_dx, a.fit2_mx, a.fit2_mdx2, y);
  }
  fft_line<NP, -1, double, WL, false>(v, lds, t, twb);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long f = static_cast<long>(a.D) * (t + m * T) + h;
    const int col = static_cast<int>((f + a.M / 2) & (a.M - 1)) - a.xlo;
    if (static_cast<unsigned>(col) < static_cast<unsigned>(a.nx)) {
      if (a.pad_const) v[m] = v[m] + ff_pad_term(a, x0, f);
      if (a.pad_linear) v[m] = v[m] + ff_linear_term(a, fb, fe, f);
      const double mag = hypot(v[m].x, v[m].y);
      a.out[static_cast<size_t>(col) * a.ny + r] = mag * mag * a.inv_len;
    }
  }
}

// resizeMatrix: output (j, i) is the mean of the column means of the block of
// rows [j dY] .. [(j+1) dY] and columns [i dX] .. [(i+1) dX] (inclusive, indices
// truncated), dY = rows/(er+1), dX = cols/(ec+1).
__global__ __launch_bounds__(256) void k_ff_resize(const double *in, int rows, int cols,
                                                   double *out, int er, int ec)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= er * ec) return;
  const int j = idx % er, i = idx / er;
  const double dX = cols / (ec + 1.0), dY = rows / (er + 1.0);
  const int r0 = static_cast<int>(j * dY), r1 = static_cast<int>((j + 1) * dY);
  const int c0 = static_cast<int>(i * dX), c1 = static_cast<int>((i + 1) * dX);
  double s = 0.0;
  for (int c = c0; c <= c1; ++c) {
    double cs = 0.0;
    for (int r = r0; r <= r1; ++r) cs += in[static_cast<size_t>(c) * rows + r];
    s += cs / (r1 - r0 + 1);
  }
  out[idx] = s / (c1 - c0 + 1);
}

// ---------------------------------------------------------------------------
// Signal lengths M that are not a power of two (FFTW takes any length,
// src/farFieldPost.cpp:53): every output inside the window is its DFT sum,
// X[f] = sum_{j<n} x_j w_M^{f j} (w_M from the M-entry table, the index f j mod M
// advanced by f per term), plus the padded segments' closed-form terms.  The
// window gives the frequencies: the reference's fftshift puts X[f] at position
// f + h for f < h, f - h for h <= f < 2h and leaves position 2h (odd M) in place,
// h = M/2 (src/farFieldPost.cpp:60-71).  One work-group per line, its values in
// LDS, 256 lanes over the window's outputs.
__device__ __forceinline__ long ff_unshift(long p, long M)
{
  const long h = M / 2;
  return p < h ? p + h : (p < 2 * h ? p - h : p);
}

__device__ __forceinline__ cd ff_direct_sum(const cd *xs, int n, long f, const cd *tw, long M)
{
  cd s{0.0, 0.0};
  long idx = 0;
  for (int j = 0; j < n; ++j) {
    s = s + xs[j] * tw[idx];
    idx += f;
    if (idx >= M) idx -= M;
  }
  return s;
}

__global__ __launch_bounds__(256) void k_ff_cols_direct(FfArgs a)
{
  __shared__ cd xs[4096];
  const int c = blockIdx.x;
  for (int j = threadIdx.x; j < a.n; j += blockDim.x) {
    const size_t i = static_cast<size_t>(c) * a.n + j;
    cd u = ff_load(a.field, a.field_c64, i);
    if (a.ref) u = u - ff_load(a.ref, a.ref_c64, i);
    xs[j] = u;
  }
  __syncthreads();
  LinFit fb{}, fe{};
  if (a.pad_linear) {
    fb = lin_fit(a.fit1_dx, a.fit1_mx, a.fit1_mdx2, xs);
    fe = lin_fit(a.fit2_dx, a.fit2_mx, a.fit2_mdx2, xs);
  }
  for (int r = threadIdx.x; r < a.ny; r += blockDim.x) {
    const long f = ff_unshift(a.ylo + r, a.M);
    cd v = ff_direct_sum(xs, a.n, f, a.tw_m, a.M);
    if (a.pad_const) v = v + ff_pad_term(a, xs[0], f);
    if (a.pad_linear) v = v + ff_linear_term(a, fb, fe, f);
    a.tmp[static_cast<size_t>(c) * a.ny + r] = v;
  }
}

__global__ __launch_bounds__(256) void k_ff_rows_direct(FfArgs a)
{
  __shared__ cd xs[4096];
  const int r = blockIdx.x;
  for (int j = threadIdx.x; j < a.n; j += blockDim.x) xs[j] = conj(a.tmp[static_cast<size_t>(j) * a.ny + r]);
  __syncthreads();
  LinFit fb{}, fe{};
  if (a.pad_linear) {
    fb = lin_fit(a.fit1_dx, a.fit1_mx, a.fit1_mdx2, xs);
    fe = lin_fit(a.fit2_dx, a.fit2_mx, a.fit2_mdx2, xs);
  }
  for (int col = threadIdx.x; col < a.nx; col += blockDim.x) {
    const long f = ff_unshift(a.xlo + col, a.M);
    cd v = ff_direct_sum(xs, a.n, f, a.tw_m, a.M);
    if (a.pad_const) v = v + ff_pad_term(a, xs[0], f);
    if (a.pad_linear) v = v + ff_linear_term(a, fb, fe, f);
    const double mag = hypot(v.x, v.y);
    a.out[static_cast<size_t>(col) * a.ny + r] = mag * mag * a.inv_len;
  }
}

hipError_t launch_far_field_direct(const FfArgs &a, hipStream_t s)
{
  if (a.n > 4096) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ff_cols_direct, dim3(a.n), dim3(256), 0, s, a);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_ff_rows_direct, dim3(a.ny), dim3(256), 0, s, a);
  return hipGetLastError();
}

#define PXG_FF_CASE(NN)                                                                   \
  case NN:                                                                                \
    hipLaunchKernelGGL(k_ff_cols<NN>, dim3(a.n, a.D), dim3(NN / 8), 0, s, a);             \
    if ((err = hipGetLastError()) != hipSuccess) return err;                              \
    hipLaunchKernelGGL(k_ff_rows<NN>, dim3(a.ny, a.D), dim3(NN / 8), 0, s, a);            \
    return hipGetLastError();

hipError_t launch_far_field(const FfArgs &a, int np, hipStream_t s)
{
  hipError_t err = hipSuccess;
  switch (np) {
    PXG_FF_CASE(8)
    PXG_FF_CASE(16)
    PXG_FF_CASE(32)
    PXG_FF_CASE(64)
    PXG_FF_CASE(128)
    PXG_FF_CASE(256)
    PXG_FF_CASE(512)
    PXG_FF_CASE(1024)
    PXG_FF_CASE(2048)
    PXG_FF_CASE(4096)
  default:
    return hipErrorInvalidValue;
  }
}

hipError_t launch_ff_resize(const double *in, int rows, int cols, double *out, int er, int ec,
                            hipStream_t s)
{
  const int cells = er * ec;
  hipLaunchKernelGGL(k_ff_resize, dim3((cells + 255) / 256), dim3(256), 0, s, in, rows, cols, out,
                     er, ec);
  return hipGetLastError();
}

}  // namespace pxg
