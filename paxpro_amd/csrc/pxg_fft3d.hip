// Split-step Fourier kernels: FFTSolver3D (two fused passes per z-step) and
// FFTSolver2D (one fused pass per z-step), plus the Solver3D cube storage.
//
// The 3D step of the reference (src/fftSolver3D.cpp:52-156,327-374) is
//   u <- A . P_z . IFFT2[ K . FFT2[u] ] / N^2
// with K(kx, ky) = exp(-i dz/(2k) (kx^2 + ky^2)) = Kx(kx) Ky(ky) separable, P_z
// the per-cell phase screen exp(-k (beta + i delta) dz) and A the separable edge
// absorber.  The field is stored with y contiguous (Armadillo column-major), so
// FFT2 = FFT_y (contiguous lines) then FFT_x (strided lines).  Because Ky
// commutes with the x transforms, one z-step becomes two passes over HBM:
//   pass X (strided x-lines):     FFT_x -> x Kx Ky -> IFFT_x
//   pass Y (contiguous y-lines):  IFFT_y -> x P_z/N^2 -> x A -> FFT_y (next step)
// A run of n steps is one forward FFT_y pass, then n (X, Y) pairs, the last Y
// pass leaving the field in real space.  Each pass reads and writes the field
// once: 4 x sizeof(complex) bytes per cell and step.
#include <type_traits>

#include "pxg_fft.hpp"
#include "pxg_fp64math.hpp"
#include "pxg_internal.hpp"
#include "pxg_screen.hpp"

namespace pxg {

constexpr int MAX_MASK_WORDS = 16;  // CSG parts per material: up to 1024

// Border-crossing integral of FFTSolver3D::refractionIntegral
// (src/fftSolver3D.cpp:253-273): f(z0) + f(z1) + 2 sum_{i=1}^{n-2} f(z0 + i dz/n),
// divided by 2n; the reference skips sample n-1 (PXG_FIX_REFR_SAMPLES restores it).
// The points (z0, z1, then the interior samples, summed in that order) are
// classified CK at a time in one walk over the culled parts each.  The map
// uses it only when a cell has more samples than a wavefront has lanes
// (map_border_cells spreads the samples over the lanes otherwise).
template <int CK>
__device__ __forceinline__ void crossing_integral(const MatDev &m, const uint64_t *mask, int nw,
                                                  unsigned n_refr, unsigned quirks, double x,
                                                  double y, double za, double zb, double &delta,
                                                  double &beta)
{
  const double h = (zb - za) / n_refr;
  const unsigned last = (quirks & 2u) ? n_refr : n_refr - 1;
  const unsigned npts = last + 1;  // z0, z1 and samples 1 .. last-1
  delta = 0.0;
  beta = 0.0;
  for (unsigned j0 = 0; j0 < npts; j0 += CK) {
    double zs[CK];
#pragma unroll
    for (int q = 0; q < CK; ++q) {
      const unsigned j = j0 + q;
      zs[q] = j == 0 ? za : (j == 1 ? zb : za + (j - 1) * h);
    }
    int es[CK];
    csg_eval_column<CK>(m, mask, nw, x, y, zs, es);
#pragma unroll
    for (int q = 0; q < CK; ++q) {
      const unsigned j = j0 + q;
      if (j >= npts) break;
      double dt, bt;
      part_values(m, es[q], dt, bt);
      const double wgt = j < 2 ? 1.0 : 2.0;
      delta += wgt * dt;
      beta += wgt * bt;
    }
  }
  delta /= (2.0 * n_refr);
  beta /= (2.0 * n_refr);
}

template <class R>
__device__ __forceinline__ void load8(cd (&v)[8], const cx<R> *f, size_t stride)
{
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = cvt<double>(f[m * stride]);
}

template <class R>
__device__ __forceinline__ void store8(const cd (&v)[8], cx<R> *f, size_t stride)
{
#pragma unroll
  for (int m = 0; m < 8; ++m) f[m * stride] = cvt<R>(v[m]);
}

// ---------------------------------------------------------------------------
// Pass Y: contiguous y-lines (matrix columns); lane t of line `line` holds rows
// t + m*N/8.  L lines (columns) per work-group, field b of the batch; `lds`
// holds L * LdsLine<N>::STRIDE values.
template <int N, class R, int L>
__device__ __forceinline__ void pass_y_body(const FftArgs<R> &a, int b, cd *lds)
{
  constexpr int T = N / 8;
  constexpr bool WL = (N / 8) <= 64;
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int c = blockIdx.x * L + line;
  cx<R> *f = a.field + b * a.fstride + static_cast<size_t>(c) * N + t;
  cd *ll = lds + line * LdsLine<N>::STRIDE;
  const bool csg = (a.mode & Y_INV) && a.mat.mode == MAT_CSG;
  // Up to 512 points the powers of the last twiddle stage are formed once for both
  // transforms (the registers are free: LDS bounds the occupancy); longer lines
  // form them per stage and transform
  using TwHold = std::conditional_t<(N <= 512), TwPow<N, double>, TwBase<N, double>>;
  const TwHold twb = [&] {
    if constexpr (N <= 512)
      return tw_pow<N>(tw_base<N>(a.tw, t));
    else
      return tw_base<N>(a.tw, t);
  }();

  // Field loads and, for a CSG material, the cells' material-map codes are
  // issued together before any arithmetic.
  cx<R> raw[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) raw[m] = f[m * T];
  uint16_t code[8];
  if (csg) {
    const uint16_t *mc = a.mcode + static_cast<size_t>(c) * N + t;
#pragma unroll
    for (int m = 0; m < 8; ++m) code[m] = mc[m * T];
  }

  cd v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = cvt<double>(raw[m]);
  if (a.mode & Y_INV) {
    fft_line<N, +1, double, WL, false>(v, ll, t, twb);
    const double k = a.kvec[b];
    if (a.mat.mode == MAT_NONE) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = scale(v[m], a.inv_norm);
    } else if (csg) {
      const cd *pf = static_cast<const cd *>(a.mat.part_fac) + b * (a.mat.n_parts + 1);
      unsigned crossing = 0;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (code[m] & MAP_CROSSING)
          crossing |= 1u << m;
        else
          v[m] = v[m] * pf[code[m]];
      }
      // Border cells (rare): the screen of the crossing integral's (delta,
      // beta).  The field values wait in the lane's own canonical LDS slots
      // (no other lane touches them between the two FFTs) so that they are
      // not live in registers across the out-of-line exp/sincos.
      if (__ballot(crossing != 0) != 0) {
#pragma unroll
        for (int m = 0; m < 8; ++m) ll[lds_pos(t + m * T)] = v[m];
        const cd *xdb = a.mxdb + static_cast<size_t>(c) * N + t;
        unsigned pending = crossing;
        while (__ballot(pending != 0) != 0) {
          if (pending) {
            const int m = __builtin_ctz(pending);
            pending &= pending - 1;
            const cd db = xdb[m * T];
            cd &slot = ll[lds_pos(t + m * T)];
            slot = slot * screen_from_values(k, a.dz, db.x, db.y, a.inv_norm);
          }
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = ll[lds_pos(t + m * T)];
      }
    } else {  // MAT_PLANES: effective (delta, beta) of this step per cell
      const int p = a.mat.z_invariant ? 0 : static_cast<int>(a.step) - a.mat.first_plane;
      const double *pl = a.mat.planes + static_cast<size_t>(p) * a.mat.plane_stride +
                         2 * static_cast<size_t>(c) * N;
      // one cell at a time through the lane's own LDS slots, as above: eight
      // interleaved FP64 exp/sincos chains would double the register count
#pragma unroll
      for (int m = 0; m < 8; ++m) ll[lds_pos(t + m * T)] = v[m];
#pragma unroll 1
      for (int m = 0; m < 8; ++m) {
        const int r = t + m * T;
        cd &slot = ll[lds_pos(r)];
        slot = slot * screen_from_values(k, a.dz, pl[2 * r], pl[2 * r + 1], a.inv_norm);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = ll[lds_pos(t + m * T)];
    }
    if (a.absy) {
      const double ax = a.absx[c];
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = scale(scale(v[m], a.absy[t + m * T]), ax);
    }
  }
  if (a.mode & Y_FWD) fft_line<N, -1, double, WL>(v, ll, t, twb);
  store8(v, f, T);
}

// blockIdx.y = field of the batch
template <int N, class R, int L>
__global__ __launch_bounds__(L *N / 8) void k_pass_y(FftArgs<R> a)
{
  __shared__ cd lds[L * LdsLine<N>::STRIDE];
  pass_y_body<N, R, L>(a, static_cast<int>(blockIdx.y), lds);
}

// Arguments of group member i (batch-1 plans), with the launch's pass-Y mode
template <class R>
__device__ __forceinline__ FftArgs<R> member_args(const GroupArgs<R> &g, int i)
{
  FftArgs<R> a = g.m[i];
  a.mode = g.mode;
  return a;
}

// blockIdx.y = group member
template <int N, class R, int L>
__global__ __launch_bounds__(L *N / 8) void k_pass_y_group(GroupArgs<R> g)
{
  __shared__ cd lds[L * LdsLine<N>::STRIDE];
  const FftArgs<R> a = member_args(g, static_cast<int>(blockIdx.y));
  pass_y_body<N, R, L>(a, 0, lds);
}

// ---------------------------------------------------------------------------
// Material map of z-step `a.step` (src/fftSolver3D.cpp:122-156), computed once
// for every field of the batch: per cell a 16-bit code holding the deciding
// part at z1, with MAP_CROSSING set when the material differs between z0 and
// z1, and for those border cells the crossing integral's (delta, beta) in
// mxdb.  Pass Y of the same step reads it.  A work-group covers the cells of
// the block [xlo, xhi] x [ylo, yhi] (column and row indices, parts culled
// against it); lane cell m is cell(m) = (column, row), m < CPL.  The field-0
// work-groups of pass X build it for their rows; k_material_map does it in a
// launch of its own for plans with few fields.
// LDS scratch a wavefront of `lanes` lanes needs for its border cells.
template <int CPL>
constexpr int map_scratch_bytes(int lanes)
{
  return lanes * CPL * 8 + lanes * 16;
}

// Border cells of a wavefront (listed in `list`, nb of them): the interior
// sample points of their crossing integrals (S per cell) are spread over the
// lanes, 64 points per CSG walk, instead of S walks per cell by its own lane;
// then one lane per cell sums its values in the order of crossing_integral
// (z0, z1, then the samples), so each result is the per-cell integral.
template <class R>
__device__ __forceinline__ void map_border_cells(const FftArgs<R> &a, const uint64_t *mask, int nw, int n,
                                                 const ushort4 *list, int nb, cd *vals, int lanes, double z0,
                                                 double z1, cd *mxdb)
{
  const int lane = static_cast<int>(__lane_id());
  const int last = (a.quirks & 2u) ? static_cast<int>(a.n_refr) : static_cast<int>(a.n_refr) - 1;
  const int S = last > 1 ? last - 1 : 0;  // interior samples per cell
  const double h = (z1 - z0) / a.n_refr;
  const int G = S > 0 ? (lanes / S > 0 ? lanes / S : 0) : lanes;  // cells per walk
  if (S > 0 && G == 0) {
    // more samples than lanes: each lane walks its cells' samples in turn
    for (int c = lane; c < nb; c += lanes) {
      const ushort4 en = list[c];
      double d, b;
      crossing_integral<1>(a.mat, mask, nw, a.n_refr, a.quirks, coord(a.xmin, en.x, a.dx),
                           coord(a.ymin, en.y, a.dy), z0, z1, d, b);
      mxdb[static_cast<size_t>(en.x) * n + en.y] = cd{d, b};
    }
    return;
  }
  for (int g0 = 0; g0 < nb; g0 += G) {
    if (S > 0 && lane < G * S) {
      const int c = g0 + lane / S;
      if (c < nb) {
        const ushort4 en = list[c];
        const double z = z0 + static_cast<double>(lane % S + 1) * h;
        double d, b;
        part_values(a.mat, csg_eval(a.mat, mask, nw, coord(a.xmin, en.x, a.dx), coord(a.ymin, en.y, a.dy), z), d,
                    b);
        vals[lane] = cd{d, b};
      }
    }
    lds_sync<true>();
    if (lane < G && g0 + lane < nb) {
      const ushort4 en = list[g0 + lane];
      double d0, b0, d1, b1;
      part_values(a.mat, en.w, d0, b0);
      part_values(a.mat, en.z, d1, b1);
      double delta = 0.0, beta = 0.0;
      delta += 1.0 * d0;
      beta += 1.0 * b0;
      delta += 1.0 * d1;
      beta += 1.0 * b1;
      for (int q = 0; q < S; ++q) {
        const cd v = vals[lane * S + q];
        delta += 2.0 * v.x;
        beta += 2.0 * v.y;
      }
      delta /= (2.0 * a.n_refr);
      beta /= (2.0 * a.n_refr);
      mxdb[static_cast<size_t>(en.x) * n + en.y] = cd{delta, beta};
    }
    lds_sync<true>();
  }
}

// wscr: this wavefront's LDS scratch, map_scratch_bytes<CPL>(lanes) bytes,
// free for the map's use (every lane of the work-group has passed a barrier
// since any other use of it).
template <int CPL, class R, class CellOf>
__device__ __forceinline__ void material_map_cells(const FftArgs<R> &a, uint64_t *mask, int n, int xlo, int xhi,
                                                   int ylo, int yhi, CellOf cell, char *wscr, int lanes)
{
  constexpr int K = CPL < 2 ? CPL : 2;
  const double z1 = coord(a.zmin, a.step, a.dz);
  const double z0 = coord(a.zmin, a.step - 1, a.dz);
  // the previous step's codes (the z0 plane) are loaded before the culling so
  // that their latency overlaps it
  uint16_t prev[CPL];
  if (a.map_prev_valid) {
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int2 cr = cell(q);
      prev[q] = a.mcode[static_cast<size_t>(cr.x) * n + cr.y];
    }
  }
  int nw = (a.mat.n_parts + 63) >> 6;
  csg_cull(a.mat, mask, nw, coord(a.xmin, xlo, a.dx), coord(a.xmin, xhi, a.dx), coord(a.ymin, ylo, a.dy),
           coord(a.ymin, yhi, a.dy), fmin(z0, z1), fmax(z0, z1));
  bool any = false;
  for (int w = 0; w < nw; ++w) any |= mask[w] != 0;
  if (!any || a.mat.is_ref) nw = 0;  // no part near this block: every cell is vacuum
  ushort4 *list = reinterpret_cast<ushort4 *>(wscr);
  cd *vals = reinterpret_cast<cd *>(wscr + lanes * CPL * 8);
  int nb = 0;  // border cells listed so far (uniform)
#pragma unroll 1
  for (int m0 = 0; m0 < CPL; m0 += K) {
    double xv[K], yv[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int2 cr = cell(m0 + q);
      xv[q] = coord(a.xmin, cr.x, a.dx);
      yv[q] = coord(a.ymin, cr.y, a.dy);
    }
    int e1[K], e0[K];
    if (a.map_prev_valid) {
      // the z0 plane is the previous step's z1 plane: its parts are in the map
      csg_eval_x2<K, false>(a.mat, mask, nw, xv, yv, z1, z0, e1, e0);
#pragma unroll
      for (int q = 0; q < K; ++q) e0[q] = prev[m0 + q] & (MAP_CROSSING - 1u);
    } else {
      csg_eval_x2<K>(a.mat, mask, nw, xv, yv, z1, z0, e1, e0);
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      unsigned code = static_cast<unsigned>(e1[q]);
      bool border = false;
      if (e1[q] != e0[q]) {
        double d1, b1, d0, b0;
        part_values(a.mat, e1[q], d1, b1);
        part_values(a.mat, e0[q], d0, b0);
        border = fabs(d1 - d0) > 1e-10 || fabs(b1 - b0) > 1e-10;
      }
      const int2 cr = cell(m0 + q);
      // list the border cells of the wavefront (slots in lane order)
      const uint64_t bal = __ballot(border);
      if (border) {
        code |= MAP_CROSSING;
        const int slot = nb + __popcll(bal & ((1ull << __lane_id()) - 1ull));
        list[slot] = make_ushort4(static_cast<unsigned short>(cr.x), static_cast<unsigned short>(cr.y),
                                  static_cast<unsigned short>(e1[q]), static_cast<unsigned short>(e0[q]));
      }
      nb += __popcll(bal);
      a.mcode[static_cast<size_t>(cr.x) * n + cr.y] = static_cast<uint16_t>(code);
    }
  }
  if (nb > 0) {
    lds_sync<true>();
    map_border_cells(a, mask, nw, n, list, nb, vals, lanes, z0, z1, a.mxdb);
  }
}

// The maps of the nz z-steps a.step .. a.step + nz - 1 in one pass over the
// cells: step a.step + k goes to mcode + k n^2 and mxdb + k n^2.  The parts are
// culled once against the block over the whole z range, and each step's z0
// classification is the previous step's z1 classification, kept in registers
// (the first step's from mprev when map_prev_valid, else evaluated).  Culling a
// superset of a step's parts changes no classification (the first part whose
// shapes contain a point decides, and a part away from the block contains none
// of its points), so every code and border value is the one material_map_cells
// computes for that step.
template <int CPL, class R, class CellOf>
__device__ __forceinline__ void material_map_planes(const FftArgs<R> &a, uint64_t *mask, int n, int xlo, int xhi,
                                                    int ylo, int yhi, CellOf cell, char *wscr, int lanes, int nz)
{
  constexpr int K = CPL < 2 ? CPL : 2;
  const size_t plane = static_cast<size_t>(n) * n;
  const double zbegin = coord(a.zmin, a.step - 1, a.dz);
  const double zend = coord(a.zmin, a.step + nz - 1, a.dz);
  int e0[CPL];  // classification at the current step's z0
  if (a.map_prev_valid) {
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int2 cr = cell(q);
      e0[q] = a.mprev[static_cast<size_t>(cr.x) * n + cr.y] & (MAP_CROSSING - 1u);
    }
  }
  int nw = (a.mat.n_parts + 63) >> 6;
  csg_cull(a.mat, mask, nw, coord(a.xmin, xlo, a.dx), coord(a.xmin, xhi, a.dx), coord(a.ymin, ylo, a.dy),
           coord(a.ymin, yhi, a.dy), fmin(zbegin, zend), fmax(zbegin, zend));
  bool any = false;
  for (int w = 0; w < nw; ++w) any |= mask[w] != 0;
  if (!any || a.mat.is_ref) nw = 0;  // no part near this block: every cell is vacuum
  if (!a.map_prev_valid) {
#pragma unroll 1
    for (int m0 = 0; m0 < CPL; m0 += K) {
      double xv[K], yv[K];
      int ez[K], unused[K];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const int2 cr = cell(m0 + q);
        xv[q] = coord(a.xmin, cr.x, a.dx);
        yv[q] = coord(a.ymin, cr.y, a.dy);
      }
      csg_eval_x2<K, false>(a.mat, mask, nw, xv, yv, zbegin, zbegin, ez, unused);
#pragma unroll
      for (int q = 0; q < K; ++q) e0[m0 + q] = ez[q];
    }
  }
  ushort4 *list = reinterpret_cast<ushort4 *>(wscr);
  cd *vals = reinterpret_cast<cd *>(wscr + lanes * CPL * 8);
#pragma unroll 1
  for (int k = 0; k < nz; ++k) {
    const double z1 = coord(a.zmin, a.step + k, a.dz);
    const double z0 = coord(a.zmin, a.step + k - 1, a.dz);
    uint16_t *mcode = a.mcode + k * plane;
    int nb = 0;  // border cells listed so far (uniform)
#pragma unroll
    for (int m0 = 0; m0 < CPL; m0 += K) {
      double xv[K], yv[K];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const int2 cr = cell(m0 + q);
        xv[q] = coord(a.xmin, cr.x, a.dx);
        yv[q] = coord(a.ymin, cr.y, a.dy);
      }
      int e1[K], unused[K];
      csg_eval_x2<K, false>(a.mat, mask, nw, xv, yv, z1, z1, e1, unused);
#pragma unroll
      for (int q = 0; q < K; ++q) {
        unsigned code = static_cast<unsigned>(e1[q]);
        bool border = false;
        if (e1[q] != e0[m0 + q]) {
          double d1, b1, d0, b0;
          part_values(a.mat, e1[q], d1, b1);
          part_values(a.mat, e0[m0 + q], d0, b0);
          border = fabs(d1 - d0) > 1e-10 || fabs(b1 - b0) > 1e-10;
        }
        const int2 cr = cell(m0 + q);
        const uint64_t bal = __ballot(border);
        if (border) {
          code |= MAP_CROSSING;
          const int slot = nb + __popcll(bal & ((1ull << __lane_id()) - 1ull));
          list[slot] = make_ushort4(static_cast<unsigned short>(cr.x), static_cast<unsigned short>(cr.y),
                                    static_cast<unsigned short>(e1[q]), static_cast<unsigned short>(e0[m0 + q]));
        }
        nb += __popcll(bal);
        mcode[static_cast<size_t>(cr.x) * n + cr.y] = static_cast<uint16_t>(code);
        e0[m0 + q] = e1[q];
      }
    }
    if (nb > 0) {
      lds_sync<true>();
      map_border_cells(a, mask, nw, n, list, nb, vals, lanes, z0, z1, a.mxdb + k * plane);
    }
    lds_sync<true>();  // the next step's list reuses the scratch
  }
}

// The same map in a launch of its own, for plans with few fields (inside pass
// X the field-0 work-groups would carry all of it): a work-group is one
// wavefront over 64 * CPL rows of one column, CPL cells per lane.
template <class R, int CPL>
__global__ __launch_bounds__(64) void k_material_map(FftArgs<R> a, int n)
{
  __shared__ uint64_t mask[MAX_MASK_WORDS];
  __shared__ __align__(16) char scr[map_scratch_bytes<CPL>(64)];
  const int lanes = static_cast<int>(blockDim.x);
  const int c = blockIdx.y, r0 = blockIdx.x * lanes * CPL;
  const int r = r0 + static_cast<int>(threadIdx.x);
  material_map_cells<CPL>(a, mask, n, c, c, r0, r0 + lanes * CPL - 1,
                          [&](int m) { return make_int2(c, r + m * lanes); }, scr, lanes);
}

// The maps of nz consecutive steps in one launch (material_map_planes), with the
// strips of k_material_map.  Work-groups along z take ppw consecutive steps each
// (small grids: more work-groups, shorter chains); with more than one group every
// group classifies its first z0 plane itself, at the coordinate the group before
// used for its last z1 plane, so the maps are the same bits (and no group reads
// a previous map that another group of the launch overwrites).
template <class R, int CPL>
__global__ __launch_bounds__(64) void k_material_map_planes(FftArgs<R> a, int n, int nz, int ppw)
{
  __shared__ uint64_t mask[MAX_MASK_WORDS];
  __shared__ __align__(16) char scr[map_scratch_bytes<CPL>(64)];
  const int lanes = static_cast<int>(blockDim.x);
  const int c = blockIdx.y, r0 = blockIdx.x * lanes * CPL;
  const int r = r0 + static_cast<int>(threadIdx.x);
  const int k0 = static_cast<int>(blockIdx.z) * ppw;
  const size_t off = static_cast<size_t>(k0) * n * n;
  a.step += static_cast<unsigned>(k0);
  a.mcode += off;
  a.mxdb += off;
  if (gridDim.z > 1) a.map_prev_valid = 0;
  material_map_planes<CPL>(a, mask, n, c, c, r0, r0 + lanes * CPL - 1,
                           [&](int m) { return make_int2(c, r + m * lanes); }, scr, lanes, min(ppw, nz - k0));
}

// The same for the members of a group, blockIdx.z = member (members without a
// CSG material have no map)
template <class R, int CPL>
__global__ __launch_bounds__(64) void k_material_map_group(GroupArgs<R> g, int n)
{
  __shared__ uint64_t mask[MAX_MASK_WORDS];
  __shared__ __align__(16) char scr[map_scratch_bytes<CPL>(64)];
  const FftArgs<R> a = member_args(g, static_cast<int>(blockIdx.z));
  if (a.mat.mode != MAT_CSG) return;
  const int lanes = static_cast<int>(blockDim.x);
  const int c = blockIdx.y, r0 = blockIdx.x * lanes * CPL;
  const int r = r0 + static_cast<int>(threadIdx.x);
  material_map_cells<CPL>(a, mask, n, c, c, r0, r0 + lanes * CPL - 1,
                          [&](int m) { return make_int2(c, r + m * lanes); }, scr, lanes);
}

// The same for line lengths n that are not a multiple of the strip: cells past
// the last row are clamped to row n - 1, so those lanes repeat that cell's work
// and write the same code (and border values) to it.
template <class R, int CPL>
__global__ __launch_bounds__(64) void k_material_map_clamped(FftArgs<R> a, int n)
{
  __shared__ uint64_t mask[MAX_MASK_WORDS];
  __shared__ __align__(16) char scr[map_scratch_bytes<CPL>(64)];
  const int lanes = static_cast<int>(blockDim.x);
  const int c = blockIdx.y, r0 = blockIdx.x * lanes * CPL;
  const int r = r0 + static_cast<int>(threadIdx.x);
  const int rlast = r0 + lanes * CPL - 1 < n - 1 ? r0 + lanes * CPL - 1 : n - 1;
  material_map_cells<CPL>(a, mask, n, c, c, r0, rlast,
                          [&](int m) { return make_int2(c, r + m * lanes < n ? r + m * lanes : n - 1); }, scr,
                          lanes);
}

// ---------------------------------------------------------------------------
// Pass X: strided x-lines (matrix rows).  Global loads and stores run with the
// lines fastest across lanes (mapping A), so that each instruction touches L
// consecutive rows of a column.  FFT_x, multiply by Kx(c) Ky(r), IFFT_x.  For
// N <= 512 the first exchange of FFT_x moves every line into one wavefront
// (mapping B) and the last exchange of IFFT_x moves it back: the exchanges in
// between need no work-group barrier.  With a CSG material the field-0
// work-groups then build the material map of the step for their rows.
// LDS stride between the lines of a pass X work-group.  In mapping A the lanes
// hold the L lines side by side, so eight consecutive lanes (the 128 bytes of
// 16-byte accesses one LDS cycle serves) touch L lines at 8/L consecutive
// canonical positions each; a stride of 8/L modulo 8 puts those eight accesses
// in eight different bank groups (an odd stride does for L >= 8).
template <int N, int L>
constexpr int pass_x_stride()
{
  return LdsLine<N>::SIZE + (L >= 8 ? 1 : 8 / L);
}

template <int N, class R, int L>
__device__ __forceinline__ void pass_x_body(const FftArgs<R> &a, int b, cd *lds)
{
  constexpr int T = N / 8;
  const int tid = threadIdx.x;
  const int la = tid % L, ta = tid / L;  // mapping A
  // Row group of this work-group.  Work-groups are dealt to the 8 XCDs in
  // turn; the groups that share a 128-byte line of a column (128 / (L s)
  // neighbours) are given to work-groups of one XCD, so that one L2 serves
  // them.  Every group is still covered exactly once.
  constexpr int G = N / L;
  const int bx = blockIdx.x;
  const int g = (G % 8 == 0) ? (bx % 8) * (G / 8) + bx / 8 : bx;
  const int ra = g * L + la;
  cx<R> *f = a.field + b * a.fstride + ra + static_cast<size_t>(ta) * N;
  cd *lla = lds + la * pass_x_stride<N, L>();
  cd v[8];
  load8(v, f, static_cast<size_t>(T) * N);
  const cd *kx = a.kx + b * N;
  if constexpr (T <= 64) {
    const int lb = tid / T, tb = tid % T;  // mapping B
    cd *llb = lds + lb * pass_x_stride<N, L>();
    const TwBase<N, double> twa = tw_base<N>(a.tw, ta), twb = tw_base<N>(a.tw, tb);
    fft_line_remap_in<N, -1>(v, lla, llb, ta, tb, twb);
    const cd kyr = a.ky[b * N + g * L + lb];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = v[m] * (kx[tb + m * T] * kyr);
    fft_line_remap_out<N, +1>(v, lla, llb, ta, tb, twa, twb);
  } else {
    const TwBase<N, double> twa = tw_base<N>(a.tw, ta);
    fft_line<N, -1, double, false, false>(v, lla, ta, twa);
    const cd kyr = a.ky[b * N + ra];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = v[m] * (kx[ta + m * T] * kyr);
    fft_line<N, +1>(v, lla, ta, twa);
  }
  store8(v, f, static_cast<size_t>(T) * N);
  if (a.map_in_pass_x && a.mat.mode == MAT_CSG && b == 0) {
    __shared__ uint64_t mask[MAX_MASK_WORDS];
    const int r0 = g * L;
    // the FFT's LDS is free (the culling's barriers follow every lane's last
    // use of it); each wavefront takes a slice for its border cells
    constexpr int WL = L * T < 64 ? L * T : 64;
    static_assert((L * T / WL) * map_scratch_bytes<8>(WL) <= L * pass_x_stride<N, L>() * static_cast<int>(sizeof(cd)),
                  "map scratch must fit the FFT's LDS");
    char *wscr = reinterpret_cast<char *>(lds) + (tid / WL) * map_scratch_bytes<8>(WL);
    material_map_cells<8>(a, mask, N, 0, N - 1, r0, r0 + L - 1,
                             [&](int m) { return make_int2(ta + m * T, ra); }, wscr, WL);
  }
}

// blockIdx.y = field of the batch
template <int N, class R, int L>
__global__ __launch_bounds__(L *N / 8) void k_pass_x(FftArgs<R> a)
{
  __shared__ cd lds[L * pass_x_stride<N, L>()];
  pass_x_body<N, R, L>(a, static_cast<int>(blockIdx.y), lds);
}

// blockIdx.y = group member (the host clears map_in_pass_x: batch-1 plans)
template <int N, class R, int L>
__global__ __launch_bounds__(L *N / 8) void k_pass_x_group(GroupArgs<R> g)
{
  __shared__ cd lds[L * pass_x_stride<N, L>()];
  const FftArgs<R> a = member_args(g, static_cast<int>(blockIdx.y));
  pass_x_body<N, R, L>(a, 0, lds);
}

// ---------------------------------------------------------------------------
// FFTSolver2D step (src/fftSolver2D.cpp:15-74): one field per line,
// FFT -> x exp(-i dz/(2k) kx^2) -> IFFT -> x exp(-k(beta + i delta) dz)/N with
// the material at z_step + dz/2 (screen_2d); the result is also stored as
// solution column.
template <int N, class R, int L>
__global__ __launch_bounds__(L *N / 8) void k_fft2d_step(FftArgs<R> a)
{
  constexpr int T = N / 8;
  __shared__ cd lds[L * LdsLine<N>::STRIDE];
  __shared__ uint64_t mask[MAX_MASK_WORDS];
  const int tid = threadIdx.x, line = tid / T, t = tid % T;
  const int b = blockIdx.x * L + line;
  cx<R> *f = a.field + b * a.fstride + t;
  cd *ll = lds + line * LdsLine<N>::STRIDE;
  cd v[8];
  load8(v, f, T);
  constexpr bool WL = (N / 8) <= 64;
  const TwBase<N, double> twb = tw_base<N>(a.tw, t);
  fft_line<N, -1, double, WL>(v, ll, t, twb);
  const cd *kx = a.kx + b * N;
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = v[m] * kx[t + m * T];
  fft_line<N, +1, double, WL>(v, ll, t, twb);
  const double k = a.kvec[b];
  const double z = coord(a.zmin, a.step, a.dz) + 0.5 * a.dz;
  const int nw = a.mat.mode == MAT_CSG ? (a.mat.n_parts + 63) >> 6 : 0;
  if (a.mat.mode == MAT_CSG) csg_cull(a.mat, mask, nw, a.xmin, coord(a.xmin, N - 1, a.dx), 0.0, 0.0, z, z);
#pragma unroll 1
  for (int m = 0; m < 8; ++m) {
    const int i = t + m * T;
    ll[lds_pos(i)] = screen_2d(a.mat, mask, nw, a.step, i, coord(a.xmin, i, a.dx), z, k, a.dz, a.inv_norm);
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = v[m] * ll[lds_pos(t + m * T)];
  store8(v, f, T);
  if (a.sol) store8(v, a.sol + b * a.sol_fstride + static_cast<size_t>(a.step) * N + t, T);
}

// ---------------------------------------------------------------------------
// Solver3D::copyCurrentSolution block average (src/solver3D.cpp:78-125): each
// output cell sums its block column by column, then divides by the block area.
template <class R>
__global__ void k_cube_store(const cx<R> *field, size_t fstride, int n, int cube_n,
                             cx<double> *cube, size_t cube_fstride, int slice)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= cube_n * cube_n) return;
  const int row = i % cube_n, col = i / cube_n;
  const double d = static_cast<double>(n) / cube_n;
  const unsigned r0 = static_cast<unsigned>(row * d), r1 = static_cast<unsigned>(row * d + d - 1);
  const unsigned c0 = static_cast<unsigned>(col * d), c1 = static_cast<unsigned>(col * d + d - 1);
  const cx<R> *f = field + b * fstride;
  double tx = 0.0, ty = 0.0;
  for (unsigned cc = c0; cc <= c1; ++cc) {
    double sx = 0.0, sy = 0.0;
    for (unsigned rr = r0; rr <= r1; ++rr) {
      const cx<R> u = f[static_cast<size_t>(cc) * n + rr];
      sx += static_cast<double>(u.x);
      sy += static_cast<double>(u.y);
    }
    tx += sx;
    ty += sy;
  }
  cube[b * cube_fstride + static_cast<size_t>(slice) * cube_n * cube_n + i] = {tx / (d * d),
                                                                             ty / (d * d)};
}

// ---------------------------------------------------------------------------
bool fft_size_supported(int n) { return n >= 16 && n <= 4096 && (n & (n - 1)) == 0; }

// Lines per work-group: at most 1024 lanes and lds_cap bytes of FP64 line
// storage (gfx950: 160 KiB of LDS per CU and per work-group).
int max_lines(int n, int lds_cap)
{
  if (n < 8) return 1;
  int l = 1024 / (n / 8);
  while (l > 1 && l * (n + n / 8 + 1) * static_cast<int>(sizeof(cd)) > lds_cap) l /= 2;
  return l > 16 ? 16 : l;
}

#define PXG_LINES_CASE(KER, NN, LL)                                                      \
  case LL:                                                                               \
    if constexpr (LL * NN / 8 <= 1024 &&                                                 \
                  LL * (NN + NN / 8 + 1) * sizeof(cd) <= 163840 - 1024) {               \
      hipLaunchKernelGGL((KER<NN, R, LL>), grid_of(NN, LL), dim3(LL * NN / 8), 0, s, a); \
      return hipGetLastError();                                                          \
    }                                                                                    \
    return hipErrorInvalidValue;

#define PXG_SIZE_CASE(KER, NN)            \
  case NN:                                \
    switch (lines) {                      \
      PXG_LINES_CASE(KER, NN, 1)          \
      PXG_LINES_CASE(KER, NN, 2)          \
      PXG_LINES_CASE(KER, NN, 4)          \
      PXG_LINES_CASE(KER, NN, 8)          \
      PXG_LINES_CASE(KER, NN, 16)         \
    default:                              \
      return hipErrorInvalidValue;        \
    }

#define PXG_SIZE_SWITCH(KER)       \
  switch (n) {                     \
    PXG_SIZE_CASE(KER, 16)         \
    PXG_SIZE_CASE(KER, 32)         \
    PXG_SIZE_CASE(KER, 64)         \
    PXG_SIZE_CASE(KER, 128)        \
    PXG_SIZE_CASE(KER, 256)        \
    PXG_SIZE_CASE(KER, 512)        \
    PXG_SIZE_CASE(KER, 1024)       \
    PXG_SIZE_CASE(KER, 2048)       \
    PXG_SIZE_CASE(KER, 4096)       \
  default:                         \
    return hipErrorInvalidValue;   \
  }

template <class R>
hipError_t launch_pass_y(int n, int lines, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (a.blue_m) return launch_blue_pass_y<R>(n, a, batch, s);
  auto grid_of = [&](int nn, int ll) { return dim3(nn / ll, batch); };
  PXG_SIZE_SWITCH(k_pass_y)
}

template <class R>
hipError_t launch_pass_x(int n, int lines, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (a.blue_m) return launch_blue_pass_x<R>(n, a, batch, s);
  auto grid_of = [&](int nn, int ll) { return dim3(nn / ll, batch); };
  PXG_SIZE_SWITCH(k_pass_x)
}

template <class R>
hipError_t launch_material_map(int n, const FftArgs<R> &a, hipStream_t s)
{
  // One wavefront per work-group: short column strips keep the culled part
  // lists short.  From n = 128 on a strip is 128 rows, two cells per lane:
  // one culling per 128 cells instead of 64.
  // (four cells per lane took 98 registers, 4 waves per SIMD, and was slower)
  if (a.blue_m) {
    // other lengths: strips of 128 rows, the last one clamped to row n - 1
    hipLaunchKernelGGL((k_material_map_clamped<R, 2>), dim3((n + 127) / 128, n), dim3(64), 0, s, a, n);
  } else if (n >= 128) {
    hipLaunchKernelGGL((k_material_map<R, 2>), dim3(n / 128, n), dim3(64), 0, s, a, n);
  } else {
    const int rows = n < 64 ? n : 64;  // n is a power of two >= 16
    hipLaunchKernelGGL((k_material_map<R, 1>), dim3(n / rows, n), dim3(rows), 0, s, a, n);
  }
  return hipGetLastError();
}

template <class R>
hipError_t launch_material_map_planes(int n, const FftArgs<R> &a, int nz, hipStream_t s)
{
  // the strips of launch_material_map (power-of-two n only)
  if (!fft_size_supported(n) || nz < 1 || nz > MAP_RING) return hipErrorInvalidValue;
  // steps per work-group: all nz when the strips alone give 16384 work-groups
  // (one wavefront each), fewer on smaller grids
  const int strips = n >= 128 ? (n / 128) * n : (n / (n < 64 ? n : 64)) * n;
  const int groups = std::max(1, std::min(nz, 16384 / strips));
  const int ppw = (nz + groups - 1) / groups;
  const unsigned gz = static_cast<unsigned>((nz + ppw - 1) / ppw);
  if (n >= 128) {
    hipLaunchKernelGGL((k_material_map_planes<R, 2>), dim3(n / 128, n, gz), dim3(64), 0, s, a, n, nz, ppw);
  } else {
    const int rows = n < 64 ? n : 64;
    hipLaunchKernelGGL((k_material_map_planes<R, 1>), dim3(n / rows, n, gz), dim3(rows), 0, s, a, n, nz, ppw);
  }
  return hipGetLastError();
}

template <class R>
hipError_t launch_pass_y_group(int n, int lines, const GroupArgs<R> &a, int members, hipStream_t s)
{
  auto grid_of = [&](int nn, int ll) { return dim3(nn / ll, members); };
  PXG_SIZE_SWITCH(k_pass_y_group)
}

template <class R>
hipError_t launch_pass_x_group(int n, int lines, const GroupArgs<R> &a, int members, hipStream_t s)
{
  auto grid_of = [&](int nn, int ll) { return dim3(nn / ll, members); };
  PXG_SIZE_SWITCH(k_pass_x_group)
}

template <class R>
hipError_t launch_material_map_group(int n, const GroupArgs<R> &a, int members, hipStream_t s)
{
  // the strips of launch_material_map (power-of-two n only)
  if (!fft_size_supported(n)) return hipErrorInvalidValue;
  if (n >= 128) {
    hipLaunchKernelGGL((k_material_map_group<R, 2>), dim3(n / 128, n, members), dim3(64), 0, s, a, n);
  } else {
    const int rows = n < 64 ? n : 64;
    hipLaunchKernelGGL((k_material_map_group<R, 1>), dim3(n / rows, n, members), dim3(rows), 0, s, a, n);
  }
  return hipGetLastError();
}

template <class R>
hipError_t launch_fft2d_step(int n, const FftArgs<R> &a, int batch, hipStream_t s)
{
  if (a.blue_m) return launch_blue_fft2d<R>(n, a, batch, s);
  // one field per line; 4 lines per work-group when the batch divides evenly
  const int lines = (batch % 4 == 0 && max_lines(n, 81920) >= 4) ? 4 : 1;
  auto grid_of = [&](int, int ll) { return dim3(batch / ll, 1); };
  PXG_SIZE_SWITCH(k_fft2d_step)
}

template <class R>
hipError_t launch_cube_store(const cx<R> *field, size_t fstride, int n, int cube_n,
                             cx<double> *cube, size_t cube_fstride, int slice, int batch,
                             hipStream_t s)
{
  const int cells = cube_n * cube_n;
  hipLaunchKernelGGL(k_cube_store<R>, dim3((cells + 255) / 256, batch), dim3(256), 0, s, field,
                     fstride, n, cube_n, cube, cube_fstride, slice);
  return hipGetLastError();
}

template hipError_t launch_pass_y<float>(int, int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_pass_y<double>(int, int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_pass_x<float>(int, int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_pass_x<double>(int, int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_material_map<float>(int, const FftArgs<float> &, hipStream_t);
template hipError_t launch_material_map<double>(int, const FftArgs<double> &, hipStream_t);
template hipError_t launch_material_map_planes<float>(int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_material_map_planes<double>(int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_pass_y_group<float>(int, int, const GroupArgs<float> &, int, hipStream_t);
template hipError_t launch_pass_y_group<double>(int, int, const GroupArgs<double> &, int, hipStream_t);
template hipError_t launch_pass_x_group<float>(int, int, const GroupArgs<float> &, int, hipStream_t);
template hipError_t launch_pass_x_group<double>(int, int, const GroupArgs<double> &, int, hipStream_t);
template hipError_t launch_material_map_group<float>(int, const GroupArgs<float> &, int, hipStream_t);
template hipError_t launch_material_map_group<double>(int, const GroupArgs<double> &, int, hipStream_t);
template hipError_t launch_fft2d_step<float>(int, const FftArgs<float> &, int, hipStream_t);
template hipError_t launch_fft2d_step<double>(int, const FftArgs<double> &, int, hipStream_t);
template hipError_t launch_cube_store<float>(const cx<float> *, size_t, int, int, cx<double> *,
                                             size_t, int, int, hipStream_t);
template hipError_t launch_cube_store<double>(const cx<double> *, size_t, int, int, cx<double> *,
                                              size_t, int, int, hipStream_t);

}  // namespace pxg
