// Device-side material evaluation: the per-cell getXrayMatProp call-back of the
// reference, restated for the GPU.  Materials are evaluated in FP64 so that
// membership tests decide exactly as the reference's double-precision code.
#pragma once

#include "pxg_common.hpp"

namespace pxg {

enum MatMode { MAT_NONE = 0, MAT_CSG = 1, MAT_PLANES = 2, MAT_CURVED = 3 };

// One shape: rows 0..2 of its 4x4 world->shape transform and its parameters.
// translate: the 3x3 part of T is exactly the identity, so T applied to a finite
// point (x, y, z) is (x + T[3], y + T[7], z + T[11]) with the same values as the
// full product (1*x = x, 0*y = +-0, x + +-0 = x).  r2lo, r2hi: for a sphere of
// radius p[0] > 0, radius_less's band r^2 (1 -+ 1e-14), products formed on the
// host as the device forms them.
struct DShape {
  int type, op, translate, pad_;
  double T[12];
  double p[3];
  double r2lo, r2hi;
};

// One part in *evaluation order*: modules in order, each module's parts last
// to first (src/materialFunction.cpp:3-13, src/geometry.cpp:212-234); the first
// entry whose shapes contain the point decides.
struct DPart {
  int first_shape, n_shapes, op, pad;
  double delta, beta;
};

// World-space bounding box of a part's union shapes (culling only).
struct DBox {
  double lo[3], hi[3];
};

struct DCurved {
  double R, width, clad_delta, clad_beta, in_delta, in_beta, wglength, dx, dz;
  int use_smoothed;
};

struct MatDev {
  int mode;
  // CSG
  const DShape *shapes;
  const DPart *parts;
  const DBox *boxes;
  int n_parts;
  int is_ref;
  const void *part_fac;   // cx<R>[batch][n_parts + 1]; entry n_parts = vacuum
  // host-sampled planes: [plane][cell][2] doubles
  const double *planes;
  int n_planes, first_plane, z_invariant;
  size_t plane_stride;    // doubles per plane
  // curved waveguide
  DCurved wg;
};

// sqrt(s) < r decided without the square root away from the boundary: with a
// correctly rounded sqrt, s < r^2 (1 - 1e-14) implies sqrt(s) < r and
// s > r^2 (1 + 1e-14) implies sqrt(s) > r; only the band between them takes
// the square root, so the decision equals the reference's sqrt(s) < r.
__device__ __forceinline__ bool radius_less(double s, double r)
{
#pragma clang fp contract(off)
  if (!(r > 0.0)) return sqrt(s) < r;
  const double r2 = r * r;
  if (s < r2 * (1.0 - 1e-14)) return true;
  if (s > r2 * (1.0 + 1e-14)) return false;
  return sqrt(s) < r;
}

// Shape membership (src/shape.cpp:126-187), strict inequalities, FP64.
__device__ __forceinline__ bool shape_inside(const DShape &s, double x, double y, double z)
{
#pragma clang fp contract(off)
  double px, py, pz;
  if (s.translate) {
    px = x + s.T[3];
    py = y + s.T[7];
    pz = z + s.T[11];
  } else {
    px = s.T[0] * x + s.T[1] * y + s.T[2] * z + s.T[3];
    py = s.T[4] * x + s.T[5] * y + s.T[6] * z + s.T[7];
    pz = s.T[8] * x + s.T[9] * y + s.T[10] * z + s.T[11];
  }
  if (s.type == 0) {
    const double q = px * px + py * py + pz * pz;
    if (!(s.p[0] > 0.0)) return sqrt(q) < s.p[0];
    if (q < s.r2lo) return true;
    if (q > s.r2hi) return false;
    return sqrt(q) < s.p[0];
  }
  if (s.type == 1)
    return (px > -s.p[0] / 2.0) && (px < s.p[0] / 2.0) && (py > -s.p[1] / 2.0) &&
           (py < s.p[1] / 2.0) && (pz > -s.p[2] / 2.0) && (pz < s.p[2] / 2.0);
  double h = s.p[2];
  if (!((pz < h / 2.0) && (pz > -h / 2.0))) return false;
  double slope = (s.p[1] - s.p[0]) / h;
  double radius = slope * (pz + h / 2.0) + s.p[0];
  return radius_less(px * px + py * py, radius);
}

// Part membership: shapes last to first, first containing shape decides
// (src/geometry.cpp:91-108).
__device__ __forceinline__ bool part_inside(const MatDev &m, const DPart &p, double x, double y,
                                            double z)
{
  for (int i = p.n_shapes - 1; i >= 0; --i) {
    const DShape &s = m.shapes[p.first_shape + i];
    if (shape_inside(s, x, y, z)) return s.op == 0;
  }
  return false;
}

// Evaluation-order index of the deciding part, or n_parts when no part claims
// the point (vacuum).  `mask` selects the parts that survived culling
// (bit e of word e/64), nw words.
__device__ __forceinline__ int csg_eval(const MatDev &m, const uint64_t *mask, int nw, double x,
                                        double y, double z)
{
  if (m.is_ref) return m.n_parts;
  for (int w = 0; w < nw; ++w) {
    // The mask is the same for every lane of the work-group: moving it to
    // scalar registers makes the part/shape descriptors scalar loads.
    const uint64_t raw = mask[w];
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw >> 32));
    uint64_t bits = (static_cast<uint64_t>(hi) << 32) | lo;
    while (bits) {
      int e = w * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      if (part_inside(m, m.parts[e], x, y, z)) return e;
    }
  }
  return m.n_parts;
}

// csg_eval for K points (x[k], y[k], za) and (x[k], y[k], zb): each culled part
// and each of its shapes is loaded once (scalar loads, the loop is uniform) and
// tested against all 2K points.  ea/eb receive the indices csg_eval returns.
// With ZB = false only the za points are classified (eb left at n_parts).
template <int K, bool ZB = true>
__device__ __forceinline__ void csg_eval_x2(const MatDev &m, const uint64_t *mask, int nw,
                                            const double (&x)[K], const double (&y)[K], double za,
                                            double zb, int (&ea)[K], int (&eb)[K])
{
#pragma unroll
  for (int k = 0; k < K; ++k) ea[k] = eb[k] = m.n_parts;
  if (m.is_ref) return;
  for (int w = 0; w < nw; ++w) {
    const uint64_t raw = mask[w];
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw >> 32));
    uint64_t bits = (static_cast<uint64_t>(hi) << 32) | lo;
    while (bits) {
      const int e = w * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      const DPart p = m.parts[e];
      int sa[K], sb[K];  // 0 undecided, 1 inside (union shape), 2 outside (difference)
#pragma unroll
      for (int k = 0; k < K; ++k) sa[k] = sb[k] = 0;
      for (int i = p.n_shapes - 1; i >= 0; --i) {
        const DShape s = m.shapes[p.first_shape + i];
        const int verdict = s.op == 0 ? 1 : 2;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (sa[k] == 0 && shape_inside(s, x[k], y[k], za)) sa[k] = verdict;
          if (ZB && sb[k] == 0 && shape_inside(s, x[k], y[k], zb)) sb[k] = verdict;
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (ea[k] == m.n_parts && sa[k] == 1) ea[k] = e;
        if (eb[k] == m.n_parts && sb[k] == 1) eb[k] = e;
      }
    }
  }
}

// csg_eval for K points on one (x, y) column at heights z[k], in one walk over
// the culled parts (each part's shapes loaded once for all K points).
template <int K>
__device__ __forceinline__ void csg_eval_column(const MatDev &m, const uint64_t *mask, int nw, double x, double y,
                                                const double (&z)[K], int (&e)[K])
{
#pragma unroll
  for (int k = 0; k < K; ++k) e[k] = m.n_parts;
  if (m.is_ref) return;
  for (int w = 0; w < nw; ++w) {
    const uint64_t raw = mask[w];
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(raw >> 32));
    uint64_t bits = (static_cast<uint64_t>(hi) << 32) | lo;
    while (bits) {
      const int p = w * 64 + __builtin_ctzll(bits);
      bits &= bits - 1;
      const DPart part = m.parts[p];
      int s[K];  // 0 undecided, 1 inside (union shape), 2 outside (difference)
#pragma unroll
      for (int k = 0; k < K; ++k) s[k] = 0;
      for (int i = part.n_shapes - 1; i >= 0; --i) {
        const DShape sh = m.shapes[part.first_shape + i];
        const int verdict = sh.op == 0 ? 1 : 2;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (s[k] == 0 && shape_inside(sh, x, y, z[k])) s[k] = verdict;
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (e[k] == m.n_parts && s[k] == 1) e[k] = p;
    }
  }
}

__device__ __forceinline__ void part_values(const MatDev &m, int e, double &d, double &b)
{
  if (e >= m.n_parts || m.parts[e].op != 0) {
    d = 0.0;
    b = 0.0;
  } else {
    d = m.parts[e].delta;
    b = m.parts[e].beta;
  }
}

__device__ __forceinline__ bool box_hits(const DBox &bx, double xlo, double xhi, double ylo,
                                         double yhi, double zlo, double zhi)
{
  return bx.lo[0] <= xhi && bx.hi[0] >= xlo && bx.lo[1] <= yhi && bx.hi[1] >= ylo &&
         bx.lo[2] <= zhi && bx.hi[2] >= zlo;
}

// Work-group culling: parts whose boxes meet the region are marked in the LDS
// mask (nw words, cleared here).  Must be called by every lane; its barriers
// order LDS only (outstanding global stores of the caller stay in flight).
__device__ __forceinline__ void csg_cull(const MatDev &m, uint64_t *mask, int nw, double xlo,
                                         double xhi, double ylo, double yhi, double zlo,
                                         double zhi)
{
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int w = tid; w < nw; w += nt) mask[w] = 0;
  lds_barrier();
  for (int e = tid; e < m.n_parts; e += nt)
    if (box_hits(m.boxes[e], xlo, xhi, ylo, yhi, zlo, zhi))
      atomicOr(reinterpret_cast<unsigned long long *>(&mask[e >> 6]), 1ull << (e & 63));
  lds_barrier();
}

// Curved waveguide membership and material (src/curvedWaveGuide2D.cpp:49-52,
// 179-205; src/waveGuideFDSimulation.cpp:219-265).
__device__ __forceinline__ bool curved_inside(const DCurved &w, double x, double z)
{
#pragma clang fp contract(off)
  return (2.0 * x * w.R + z * z > 0.0) && (2.0 * x * w.R + z * z < 2.0 * w.width * w.R);
}

__device__ __forceinline__ void curved_material(const DCurved &w, double x, double z, double &d,
                                                double &b)
{
#pragma clang fp contract(off)
  if (w.use_smoothed) {
    double ww = 0.1 * w.width;
    double dd = 2.0 * w.R * x + z * z;
    double f1 = 1.0 / (1.0 + exp(-dd / ww));
    double f2 = 1.0 / (1.0 + exp((2 * w.R * w.width - dd) / ww));
    double f = 1.0 - (f1 - f2);
    d = w.clad_delta * f;
    b = w.clad_beta * f;
    return;
  }
  if (z > w.wglength) {
    d = 0.0;
    b = 0.0;
    return;
  }
  bool inside = curved_inside(w, x, z);
  bool nb = curved_inside(w, x + w.dx, z) && curved_inside(w, x - w.dx, z) &&
            curved_inside(w, x, z + w.dz) && curved_inside(w, x, z - w.dz);
  if (inside && nb) {
    d = w.in_delta;
    b = w.in_beta;
  } else if (inside) {
    d = 0.5 * (w.clad_delta + w.in_delta);
    b = 0.5 * (w.clad_beta + w.in_beta);
  } else {
    d = w.clad_delta;
    b = w.clad_beta;
  }
}

}  // namespace pxg
