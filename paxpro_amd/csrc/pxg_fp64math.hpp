// FP64 exp and sincos for the phase screens, written to be inlined into the
// FFT passes: an out-of-line library call makes the calling kernel keep its
// live values in callee-saved registers, which cost the passes a wave per SIMD.
//
// Method (Cody and Waite, "Software Manual for the Elementary Functions",
// 1980): x = n c + r with n = rint(x / c), c = ln 2 or pi/2, the reduction
// x - n c formed by fused multiply-adds over the double-precision parts
// c1 + c2 (+ c3) of c (each part the rounding of what the previous parts
// leave; the first is the double nearest c).  r is then within about
// 1e-16 of the true remainder for |n| up to about 2^50, because each fused step
// rounds once, at the magnitude of r.  e^r, sin r and cos r are Taylor
// polynomials in Horner form with coefficients 1/k!: factorials up to 18! are
// exact doubles and the quotients round once at compile time.  Truncation:
// e^r to r^13 on |r| <= ln2/2 leaves < 5e-18, sin to r^17 and cos to r^18 on
// |r| <= pi/4 leave < 1e-19.  Every multiply feeding an add is an explicit
// fma, so the host and device builds round identically.
//
// Range: exp_fp64 is 0 below -746 and +inf above 710 (NaN stays NaN);
// sincos_fp64 is accurate to about 1e-16 absolute for |x| < 1e15 (the phase
// of a screen is k delta dz, some radians) and NaN for infinite x.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define PXG_HD __host__ __device__ __forceinline__
#else
#define PXG_HD inline
#endif

namespace pxg {

// A coefficient materialised in scalar registers where it is used.  Without
// this the compiler hoists every coefficient of the polynomials into vector
// registers for the whole kernel (v_fmac needs its addend in the destination),
// some 40 VGPRs that lower the occupancy of the kernels these are inlined into.
PXG_HD double kc(double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+s"(c));
#endif
  return c;
}

PXG_HD double exp_fp64(double x)
{
  if (x < -746.0) return 0.0;
  if (x > 710.0) return HUGE_VAL;
  const double n = std::rint(x * kc(1.4426950408889634));        // x / ln 2
  double r = std::fma(-n, kc(0.6931471805599453), x);            // ln 2, part 1
  r = std::fma(-n, kc(2.3190468138462996e-17), r);               // part 2
  double p = kc(1.0 / 6227020800.0);  // 1/13!
  p = std::fma(p, r, kc(1.0 / 479001600.0));
  p = std::fma(p, r, kc(1.0 / 39916800.0));
  p = std::fma(p, r, kc(1.0 / 3628800.0));
  p = std::fma(p, r, kc(1.0 / 362880.0));
  p = std::fma(p, r, kc(1.0 / 40320.0));
  p = std::fma(p, r, kc(1.0 / 5040.0));
  p = std::fma(p, r, kc(1.0 / 720.0));
  p = std::fma(p, r, kc(1.0 / 120.0));
  p = std::fma(p, r, kc(1.0 / 24.0));
  p = std::fma(p, r, kc(1.0 / 6.0));
  p = std::fma(p, r, 0.5);
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, static_cast<int>(n));
}

PXG_HD void sincos_fp64(double x, double &s, double &c)
{
  if (!std::isfinite(x)) {
    s = c = x - x;  // NaN
    return;
  }
  const double q = std::rint(x * kc(0.6366197723675814));       // x / (pi/2)
  double r = std::fma(-q, kc(1.5707963267948966), x);            // pi/2, part 1
  r = std::fma(-q, kc(6.123233995736766e-17), r);                // part 2
  r = std::fma(-q, kc(-1.4973849048591698e-33), r);              // part 3
  const double r2 = r * r;
  // sin r = r + r^3 P(r^2)
  double ps = kc(1.0 / 355687428096000.0);  // 1/17!
  ps = std::fma(ps, r2, kc(-1.0 / 1307674368000.0));
  ps = std::fma(ps, r2, kc(1.0 / 6227020800.0));
  ps = std::fma(ps, r2, kc(-1.0 / 39916800.0));
  ps = std::fma(ps, r2, kc(1.0 / 362880.0));
  ps = std::fma(ps, r2, kc(-1.0 / 5040.0));
  ps = std::fma(ps, r2, kc(1.0 / 120.0));
  ps = std::fma(ps, r2, kc(-1.0 / 6.0));
  const double sr = std::fma(r * r2, ps, r);
  // cos r = 1 + r^2 Q(r^2)
  double pc = kc(-1.0 / 6402373705728000.0);  // -1/18!
  pc = std::fma(pc, r2, kc(1.0 / 20922789888000.0));
  pc = std::fma(pc, r2, kc(-1.0 / 87178291200.0));
  pc = std::fma(pc, r2, kc(1.0 / 479001600.0));
  pc = std::fma(pc, r2, kc(-1.0 / 3628800.0));
  pc = std::fma(pc, r2, kc(1.0 / 40320.0));
  pc = std::fma(pc, r2, kc(-1.0 / 720.0));
  pc = std::fma(pc, r2, kc(1.0 / 24.0));
  pc = std::fma(pc, r2, -0.5);
  const double cr = std::fma(r2, pc, 1.0);
  // quadrant q mod 4 (exact for |q| < 2^53)
  const int quad = static_cast<int>(q - 4.0 * std::floor(q * 0.25));
  const double a = (quad & 1) ? cr : sr, b = (quad & 1) ? sr : cr;
  s = (quad & 2) ? -a : a;
  c = ((quad + 1) & 2) ? -b : b;
}

}  // namespace pxg
