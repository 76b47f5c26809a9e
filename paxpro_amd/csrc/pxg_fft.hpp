// LDS-staged Stockham FFT for power-of-two line lengths on gfx950.
//
// A line of N points is held by T = N/8 lanes, 8 complex values per lane in
// "canonical" register order: lane t holds positions t + m*N/8, m = 0..7.
// Stages follow the self-sorting Stockham formulation (Govindaraju et al.,
// SC08): with Ns the product of the radices already applied, butterfly j reads
// in[j + r*N/R], multiplies by the twiddle exp(s*2*pi*i*r*(j mod Ns)/(Ns*R)),
// applies a radix-R DFT and writes out[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].
// Radix-8 stages are used throughout, preceded by one radix-2 or radix-4
// stage when log2(N) is not a multiple of 3.  The first stage reads from
// registers and the last stage writes to registers in canonical order, so a
// kernel can load a line from HBM straight into registers, chain transforms and
// point-wise operations in registers, and store straight back: only the inner
// stage exchanges go through LDS.  Twiddles come from an FP64-accurate table
// tw[m] = exp(-2*pi*i*m/N) rounded once to the working precision.
#pragma once

#include "pxg_common.hpp"

namespace pxg {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

template <int N>
struct FftShape {
  static_assert(N >= 8 && (N & (N - 1)) == 0, "power-of-two N >= 8");
  static constexpr int LOG = ilog2(N);
  static constexpr int REM = LOG % 3;           // 0, 1 or 2
  static constexpr int FIRST = 1 << (REM ? REM : 3);  // radix of the first stage
  static constexpr int N8 = N / 8;              // lanes per line
};

template <int S, class R>
__device__ __forceinline__ void dft2(cx<R> &a, cx<R> &b)
{
  cx<R> t = a;
  a = t + b;
  b = t - b;
}

template <int S, class R>
__device__ __forceinline__ void dft4(cx<R> &u0, cx<R> &u1, cx<R> &u2, cx<R> &u3)
{
  cx<R> a0 = u0 + u2, a1 = u0 - u2, a2 = u1 + u3, a3 = mul_si<S>(u1 - u3);
  u0 = a0 + a2;
  u1 = a1 + a3;
  u2 = a0 - a2;
  u3 = a1 - a3;
}

template <int S, class R>
__device__ __forceinline__ void dft8(cx<R> (&u)[8])
{
  const R c = static_cast<R>(0.70710678118654752440084436210485);
  cx<R> e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
  cx<R> o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
  dft4<S>(e0, e1, e2, e3);
  dft4<S>(o0, o1, o2, o3);
  // w8^1 = (1 + S i)/sqrt2, w8^2 = S i, w8^3 = (-1 + S i)/sqrt2
  cx<R> t1 = {(o1.x - S * o1.y) * c, (o1.y + S * o1.x) * c};
  cx<R> t2 = mul_si<S>(o2);
  cx<R> t3 = {(-o3.x - S * o3.y) * c, (-o3.y + S * o3.x) * c};
  u[0] = e0 + o0;
  u[4] = e0 - o0;
  u[1] = e1 + t1;
  u[5] = e1 - t1;
  u[2] = e2 + t2;
  u[6] = e2 - t2;
  u[3] = e3 + t3;
  u[7] = e3 - t3;
}

// Radix-R butterflies of one stage on the canonical registers of lane t.
template <int N, int R, int NS, int S, class Rt>
__device__ __forceinline__ void stage_compute(cx<Rt> (&v)[8], int t, const cx<Rt> *__restrict__ tw)
{
  constexpr int Q = 8 / R;
  constexpr int N8 = N / 8;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = t + q * N8;
    cx<Rt> u[8];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
    if constexpr (NS > 1) {
      // One table load per butterfly: w = e^{-2 pi i jm/(NS R)}; its powers
      // w^2..w^{R-1} are products at most three deep (a few ulp in FP64).
      // Loading all R-1 powers cost R-1 scattered reads per lane and stage,
      // which dominated the pass time.
      const int jm = j & (NS - 1);
      cx<Rt> w[8];
      w[1] = tw[jm * (N / (NS * R))];
      if constexpr (S > 0) w[1] = conj(w[1]);
      if constexpr (R >= 4) {
        w[2] = w[1] * w[1];
        w[3] = w[2] * w[1];
      }
      if constexpr (R == 8) {
        w[4] = w[2] * w[2];
        w[5] = w[4] * w[1];
        w[6] = w[3] * w[3];
        w[7] = w[4] * w[3];
      }
#pragma unroll
      for (int r = 1; r < R; ++r) u[r] = u[r] * w[r];
    }
    if constexpr (R == 8) {
      dft8<S>(u);
    } else if constexpr (R == 4) {
      dft4<S>(u[0], u[1], u[2], u[3]);
    } else {
      dft2<S>(u[0], u[1]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[q + r * Q] = u[r];
  }
}

// LDS position of logical line index i: one pad slot every 8 entries keeps the
// radix-8 scatter of the early stages free of bank conflicts.
__device__ __forceinline__ constexpr int lds_pos(int i) { return i + (i >> 3); }
template <int N>
struct LdsLine {
  static constexpr int SIZE = N + N / 8;      // padded line length
  static constexpr int STRIDE = SIZE + 1;     // odd stride between lines
};

template <int N, int R, int NS, class Rt>
__device__ __forceinline__ void stage_store(const cx<Rt> (&v)[8], cx<Rt> *lds, int t)
{
  constexpr int Q = 8 / R;
  constexpr int N8 = N / 8;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = t + q * N8;
    const int idx = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
    for (int r = 0; r < R; ++r) lds[lds_pos(idx + r * NS)] = v[q + r * Q];
  }
}

template <int N, class Rt>
__device__ __forceinline__ void load_canon(cx<Rt> (&v)[8], const cx<Rt> *lds, int t)
{
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = lds[lds_pos(t + m * (N / 8))];
}

// LDS ordering between the stages.  When every lane of a line sits in one
// wavefront and the line's LDS slice belongs to that wavefront alone (N <= 512
// with lanes of a line adjacent), completing the wavefront's own LDS accesses
// (waitcnt, emitted by the work-group-scope fences) orders them: no s_barrier.
template <bool WAVE_LOCAL>
__device__ __forceinline__ void lds_sync()
{
  if constexpr (WAVE_LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

template <int N, int S, int NS, class Rt, bool WL>
struct StockhamStages {
  static __device__ __forceinline__ void run(cx<Rt> (&v)[8], cx<Rt> *lds, int t,
                                             const cx<Rt> *__restrict__ tw)
  {
    constexpr int R = (NS == 1) ? FftShape<N>::FIRST : 8;
    stage_compute<N, R, NS, S>(v, t, tw);
    if constexpr (NS * R < N) {
      stage_store<N, R, NS>(v, lds, t);
      lds_sync<WL>();
      load_canon<N>(v, lds, t);
      // Ordering before the next stage's stores.  After the last exchange each
      // lane has read only its own canonical slots, so nothing is needed
      // there: a following transform starts with its own sync.
      if constexpr (NS * R * 8 < N) lds_sync<WL>();
      StockhamStages<N, S, NS * R, Rt, WL>::run(v, lds, t, tw);
    }
  }
};

// Full length-N transform of the canonical registers (S = -1 forward, +1
// backward, unnormalised).  `lds` is this line's LdsLine<N>::SIZE slice; every
// lane of the work-group must call it (it contains work-group barriers unless
// WL, see lds_sync).  Between transforms a lane may touch only its own
// canonical slots (lds_pos(t + m N/8)).  The leading sync protects LDS still
// being read by a preceding transform; LEAD = false skips it where the line's
// LDS has not been used before in the kernel.
template <int N, int S, class Rt, bool WL = false, bool LEAD = true>
__device__ __forceinline__ void fft_line(cx<Rt> (&v)[8], cx<Rt> *lds, int t,
                                         const cx<Rt> *__restrict__ tw)
{
  if constexpr (LEAD) lds_sync<WL>();
  StockhamStages<N, S, 1, Rt, WL>::run(v, lds, t, tw);
}

}  // namespace pxg
