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
// tw[m] = exp(-2*pi*i*m/N) rounded once to the working precision: one base per
// stage and lane, its powers formed by products.
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

// Stages after the first carry twiddles; they are radix 8, so each lane has
// one butterfly (j = t) and needs one base twiddle w = tw[(t mod Ns) N/(8 Ns)]
// per stage.  The bases depend only on the lane: a kernel loads them once
// (tw_base) and passes them down, so no table read sits behind a barrier.
template <int N>
constexpr int tw_stages()
{
  int n = N / FftShape<N>::FIRST, s = 0;
  while (n > 1) {
    n /= 8;
    ++s;
  }
  return s;
}

// Twiddle slot of the stage whose input spans NS = FIRST * 8^slot.
template <int N, int NS>
constexpr int tw_slot()
{
  int ns = FftShape<N>::FIRST, s = 0;
  while (ns < NS) {
    ns *= 8;
    ++s;
  }
  return s;
}

template <int N, class Rt>
struct TwBase {
  static constexpr bool POW = false;
  cx<Rt> w[tw_stages<N>() > 0 ? tw_stages<N>() : 1];
};

// The seven powers of the last twiddle stage's base, formed once (the same
// products as the per-stage computation), with the bases of the other stages:
// for kernels that run several transforms per lane and have registers to spare.
// Inverse transforms multiply by the conjugates (the per-stage path conjugates
// the base first; the two agree to rounding, as contraction may pair the
// operands differently).
template <int N, class Rt>
struct TwPow {
  static constexpr bool POW = true;
  static constexpr int KEEP = tw_stages<N>() - 1;  // slots from here hold powers
  cx<Rt> b[tw_stages<N>() > 0 ? tw_stages<N>() : 1];     // bases
  cx<Rt> w[tw_stages<N>() > 0 ? tw_stages<N>() : 1][7];  // w[s][r - 1] = base^r
};

template <int N, class Rt>
__device__ __forceinline__ TwBase<N, Rt> tw_base(const cx<Rt> *__restrict__ tw, int t)
{
  TwBase<N, Rt> b;
#pragma unroll
  for (int s = 0, ns = FftShape<N>::FIRST; s < tw_stages<N>(); ++s, ns *= 8)
    b.w[s] = tw[(t & (ns - 1)) * (N / (ns * 8))];
  return b;
}

template <int N, class Rt>
__device__ __forceinline__ TwPow<N, Rt> tw_pow(const TwBase<N, Rt> &b)
{
  TwPow<N, Rt> p;
#pragma unroll
  for (int s = 0; s < tw_stages<N>(); ++s) p.b[s] = b.w[s];
#pragma unroll
  for (int s = TwPow<N, Rt>::KEEP; s < tw_stages<N>(); ++s) {
    const cx<Rt> w1 = b.w[s];
    const cx<Rt> w2 = w1 * w1, w3 = w2 * w1, w4 = w2 * w2;
    p.w[s][0] = w1;
    p.w[s][1] = w2;
    p.w[s][2] = w3;
    p.w[s][3] = w4;
    p.w[s][4] = w4 * w1;
    p.w[s][5] = w3 * w3;
    p.w[s][6] = w4 * w3;
  }
  return p;
}

// Radix-R butterflies of one stage on the canonical registers of lane t.
template <int N, int R, int NS, int S, class Rt, class TW>
__device__ __forceinline__ void stage_compute(cx<Rt> (&v)[8], int t, const TW &twb)
{
  constexpr int Q = 8 / R;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cx<Rt> u[8];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
    if constexpr (NS > 1 && TW::POW && tw_slot<N, NS>() >= TwPow<N, Rt>::KEEP) {
      static_assert(R == 8, "twiddle stages are radix 8");
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cx<Rt> w = twb.w[tw_slot<N, NS>()][r - 1];
        u[r] = u[r] * (S > 0 ? conj(w) : w);
      }
    } else if constexpr (NS > 1) {
      // The base w = e^{-2 pi i (j mod NS)/(NS R)} of butterfly j = t + q N/8
      // (held in twb); its powers w^2..w^{R-1}
      // are products at most three deep (a few ulp in FP64).  Loading all R-1
      // powers cost R-1 scattered reads per lane and stage, which dominated
      // the pass time.
      static_assert(R == 8, "twiddle stages are radix 8");
      cx<Rt> w[8];
      if constexpr (TW::POW)
        w[1] = twb.b[tw_slot<N, NS>()];
      else
        w[1] = twb.w[tw_slot<N, NS>()];
      // Opaque copy: keeps the powers local to this stage.  Without it the
      // compiler forms them once per kernel and keeps up to 18 complex values
      // live across both transforms (spills at 1024 lanes).
      asm volatile("" : "+v"(w[1].x), "+v"(w[1].y));
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

// WL: every lane of a line sits in one wavefront and the line's LDS slice
// belongs to that wavefront alone (N <= 512 with lanes of a line adjacent), so
// lds_sync<true> orders the stages.
template <int N, int S, int NS, class Rt, bool WL>
struct StockhamStages {
  template <class TW>
  static __device__ __forceinline__ void run(cx<Rt> (&v)[8], cx<Rt> *lds, int t, const TW &twb)
  {
    constexpr int R = (NS == 1) ? FftShape<N>::FIRST : 8;
    stage_compute<N, R, NS, S>(v, t, twb);
    if constexpr (NS * R < N) {
      stage_store<N, R, NS>(v, lds, t);
      lds_sync<WL>();
      load_canon<N>(v, lds, t);
      // Ordering before the next stage's stores.  After the last exchange each
      // lane has read only its own canonical slots, so nothing is needed
      // there: a following transform starts with its own sync.
      if constexpr (NS * R * 8 < N) lds_sync<WL>();
      StockhamStages<N, S, NS * R, Rt, WL>::run(v, lds, t, twb);
    }
  }
};

// Transforms whose first or last exchange also moves each line to another set
// of lanes.  Mapping A is any lane layout (e.g. lines fastest, for coalesced
// strided global access); mapping B keeps a line's N/8 lanes inside one
// wavefront (N <= 512), so that its other exchanges need no work-group
// barrier.  `la`/`lb` are the line's LDS slices under each mapping.
//
// Forward, A -> B: first stage in A, remapping exchange (work-group barrier),
// remaining stages wave-local in B.  Registers end canonical in B.
template <int N, int S>
__device__ __forceinline__ void fft_line_remap_in(cd (&v)[8], cd *la, cd *lb, int ta, int tb,
                                                  const TwBase<N, double> &twb_b)
{
  static_assert(N / 8 <= 64, "mapping B needs a line inside one wavefront");
  constexpr int F = FftShape<N>::FIRST;
  const TwBase<N, double> none{};
  stage_compute<N, F, 1, S>(v, ta, none);
  stage_store<N, F, 1>(v, la, ta);
  lds_barrier();
  load_canon<N>(v, lb, tb);
  if constexpr (F * 8 < N) lds_sync<true>();
  StockhamStages<N, S, F, double, true>::run(v, lb, tb, twb_b);
}

// Inverse, B -> A: stages wave-local in B up to the last exchange, which
// remaps to A (work-group barrier); the last stage runs in A.  The caller's
// previous use of lb must be wave-local (a leading wave-local sync is done).
template <int N, int S, int NS>
struct RemapOutStages {
  static __device__ __forceinline__ void run(cd (&v)[8], cd *la, cd *lb, int ta, int tb,
                                             const TwBase<N, double> &twb_a, const TwBase<N, double> &twb_b)
  {
    constexpr int R = (NS == 1) ? FftShape<N>::FIRST : 8;
    static_assert(NS * R < N, "at least one stage follows");
    stage_compute<N, R, NS, S>(v, tb, twb_b);
    stage_store<N, R, NS>(v, lb, tb);
    if constexpr (NS * R * 8 < N) {
      lds_sync<true>();
      load_canon<N>(v, lb, tb);
      lds_sync<true>();
      RemapOutStages<N, S, NS * R>::run(v, la, lb, ta, tb, twb_a, twb_b);
    } else {
      lds_barrier();
      load_canon<N>(v, la, ta);
      stage_compute<N, 8, NS * R, S>(v, ta, twb_a);
    }
  }
};

template <int N, int S>
__device__ __forceinline__ void fft_line_remap_out(cd (&v)[8], cd *la, cd *lb, int ta, int tb,
                                                   const TwBase<N, double> &twb_a,
                                                   const TwBase<N, double> &twb_b)
{
  static_assert(N / 8 <= 64, "mapping B needs a line inside one wavefront");
  lds_sync<true>();
  RemapOutStages<N, S, 1>::run(v, la, lb, ta, tb, twb_a, twb_b);
}

// Full length-N transform of the canonical registers (S = -1 forward, +1
// backward, unnormalised).  `lds` is this line's LdsLine<N>::SIZE slice; every
// lane of the work-group must call it (it contains work-group barriers unless
// WL, see lds_sync).  Between transforms a lane may touch only its own
// canonical slots (lds_pos(t + m N/8)).  The leading sync protects LDS still
// being read by a preceding transform; LEAD = false skips it where the line's
// LDS has not been used before in the kernel.
template <int N, int S, class Rt, bool WL = false, bool LEAD = true, class TW>
__device__ __forceinline__ void fft_line(cx<Rt> (&v)[8], cx<Rt> *lds, int t, const TW &twb)
{
  if constexpr (LEAD) lds_sync<WL>();
  StockhamStages<N, S, 1, Rt, WL>::run(v, lds, t, twb);
}

}  // namespace pxg
