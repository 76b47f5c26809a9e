// Shared device/host helpers of libpaxgpu: complex types, arithmetic, and the
// status/error plumbing behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace pxg {

// Interleaved complex of real type R; 8-byte (float) or 16-byte (double) aligned
// so a lane moves one complex per load.
template <class R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};
using cf = cx<float>;
using cd = cx<double>;

template <class R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b)
{
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R>
__host__ __device__ __forceinline__ cx<R> scale(cx<R> a, R s) { return {a.x * s, a.y * s}; }
template <class R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return {a.x, -a.y}; }
// multiply by S*i, S = +1 or -1
template <int S, class R>
__host__ __device__ __forceinline__ cx<R> mul_si(cx<R> a) { return S > 0 ? cx<R>{-a.y, a.x} : cx<R>{a.y, -a.x}; }
// complex division (a / b), scaled to avoid overflow (Smith's algorithm)
template <class R>
__host__ __device__ __forceinline__ cx<R> cdiv(cx<R> a, cx<R> b)
{
  R ax = b.x < 0 ? -b.x : b.x, ay = b.y < 0 ? -b.y : b.y;
  if (ax >= ay) {
    R r = b.y / b.x, den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  R r = b.x / b.y, den = b.x * r + b.y;
  return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
template <class R>
__host__ __device__ __forceinline__ R norm2(cx<R> a) { return a.x * a.x + a.y * a.y; }

template <class D, class S>
__host__ __device__ __forceinline__ cx<D> cvt(cx<S> a) { return {static_cast<D>(a.x), static_cast<D>(a.y)}; }

// exp(i*phi) for a phase argument held in FP64: reduced to [-pi, pi] in FP64
// before the sine and cosine are taken, then rounded to R.
template <class R>
__device__ __forceinline__ cx<R> expi_fp64(double phi)
{
  const double two_pi = 6.283185307179586476925286766559;
  double r = phi - two_pi * rint(phi / two_pi);
  double s, c;
  sincos(r, &s, &c);
  return {static_cast<R>(c), static_cast<R>(s)};
}

// 1/d for a finite, normal, nonzero d: the hardware reciprocal estimate refined
// by two Newton-Raphson steps r <- r + r (1 - d r), each of which squares the
// relative error; the result is within an ulp or two of the rounded quotient.
// Used where a quotient feeds FP64 arithmetic whose tolerance is far above an
// ulp (tridiagonal eliminations), in place of the IEEE division sequence.
__device__ __forceinline__ double recip(double d)
{
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// 1/b = conj(b) / |b|^2 for |b|^2 in the normal range.
__device__ __forceinline__ cx<double> crecip(cx<double> b)
{
  const double r = recip(b.x * b.x + b.y * b.y);
  return {b.x * r, -b.y * r};
}

// Grid coordinate min + i*step, evaluated without FMA contraction like the
// reference's getX/getY/getZ (src/paraxialSimulation.cpp:250-263).
__host__ __device__ __forceinline__ double coord(double min, unsigned i, double step)
{
#pragma clang fp contract(off)
  return min + static_cast<double>(i) * step;
}

// Work-group barrier that orders LDS accesses only: the fences are restricted
// to the local address space, so the barrier waits for the wavefront's LDS
// operations (lgkmcnt) but not for its outstanding global loads and stores,
// which __syncthreads would drain (vmcnt).  For phases that exchange data
// through LDS alone.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS ordering between the phases of a computation.  When the LDS slots a
// wavefront reads were written by that wavefront alone, completing the
// wavefront's own LDS accesses (waitcnt, emitted by the work-group-scope
// fences) orders them: no s_barrier.  Otherwise a work-group barrier that
// orders LDS only.
template <bool WAVE_LOCAL>
__device__ __forceinline__ void lds_sync()
{
  if constexpr (WAVE_LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    lds_barrier();
  }
}

}  // namespace pxg
