// Internal structures shared by the C-ABI layer (pxg_api.hip) and the kernel
// translation units.  Not part of the public interface.
#pragma once

#include "../../include/paxgpu.h"
#include "pxg_common.hpp"
#include "pxg_material.hpp"

namespace pxg {

// Arguments of the split-step passes (FFT3D and FFT2D).  Fields are stored as
// cx<R> (complex128 or complex64); every table and all arithmetic are FP64.
template <class R>
struct FftArgs {
  cx<R> *field;            // batch of fields
  size_t fstride;          // elements between consecutive fields
  const cd *tw;            // tw[m] = exp(-2 pi i m / N), m < N
  const cd *kx;            // [batch][N] exp(-i dz/(2k) kx^2), column (x) index
  const cd *ky;            // [batch][N] exp(-i dz/(2k) ky^2), row (y) index
  const double *absx;      // [N] absorber factor along x, or null
  const double *absy;      // [N] absorber factor along y, or null
  const double *kvec;      // [batch] wavenumbers
  MatDev mat;
  double xmin, dx, ymin, dy, zmin, dz;
  double inv_norm;         // 1/(number of cells)
  unsigned step;           // z-step index being applied by the screen
  unsigned n_refr;         // crossing-integral samples
  unsigned quirks;
  int mode;                // pass-Y mode bits (Y_INV | Y_FWD)
  cx<R> *sol;              // FFT2D: solution columns (Nx x Nz per field) or null
  size_t sol_fstride;
  // FFT3D CSG material map of the current z-step, shared by the batch (built
  // by pass X, read by pass Y):
  uint16_t *mcode;         // [N*N] deciding part at z1; bit 15: border crossing
  const uint16_t *mprev;   // the previous step's codes (the map's z0 plane when map_prev_valid)
  cd *mxdb;                // [N*N] (delta, beta) of the crossing integral, border cells only
  int map_in_pass_x;       // build the map inside pass X (large batches) or in its own launch
  int map_prev_valid;      // mcode holds the map of step - 1: reuse its parts as the z0 plane
  // Line lengths without a power-of-two kernel (pxg_bluestein.hip): N-point DFTs as
  // Bluestein convolutions of M = blue_m points (tw then holds the M-point table)
  int blue_m;              // 0: power-of-two path
  const cd *chirp;         // [N] exp(-i pi j^2 / N), j^2 reduced mod 2N
  const cd *bf, *bi;       // [M] FFT_M of the forward / inverse convolution kernel, / M
};

// Batch-1 FFT3D plans of one line length advanced in shared launches
// (pxg_step_group): the members' arguments travel by value with each launch
// (kernel arguments keep their buffers known as global memory); MAX_GROUP
// members keep them within 4 KiB.
constexpr int MAX_GROUP = 8;
template <class R>
struct GroupArgs {
  FftArgs<R> m[MAX_GROUP];
  int mode;
};
static_assert(sizeof(GroupArgs<double>) <= 4096, "group kernel arguments within 4 KiB");

// Bit of a material-map code marking a border cell (material differs between
// the two planes of the step); the other bits hold the deciding part index.
constexpr unsigned MAP_CROSSING = 0x8000u;

enum { Y_INV = 1, Y_FWD = 2 };

// Launchers (pxg_fft3d.hip).  `lines` = lines per work-group.
template <class R>
hipError_t launch_pass_y(int n, int lines, const FftArgs<R> &a, int batch, hipStream_t s);
template <class R>
hipError_t launch_pass_x(int n, int lines, const FftArgs<R> &a, int batch, hipStream_t s);
template <class R>
hipError_t launch_material_map(int n, const FftArgs<R> &a, hipStream_t s);
// The maps of nz <= MAP_RING consecutive steps a.step .. a.step + nz - 1 in one
// launch, step a.step + k into a.mcode + k n^2 and a.mxdb + k n^2 (power-of-two n)
constexpr int MAP_RING = 8;
template <class R>
hipError_t launch_material_map_planes(int n, const FftArgs<R> &a, int nz, hipStream_t s);
// The same three launches over the members of a group (grid y or z = member)
template <class R>
hipError_t launch_pass_y_group(int n, int lines, const GroupArgs<R> &a, int members, hipStream_t s);
template <class R>
hipError_t launch_pass_x_group(int n, int lines, const GroupArgs<R> &a, int members, hipStream_t s);
template <class R>
hipError_t launch_material_map_group(int n, const GroupArgs<R> &a, int members, hipStream_t s);
template <class R>
hipError_t launch_fft2d_step(int n, const FftArgs<R> &a, int batch, hipStream_t s);
// Bluestein variants of the passes for other line lengths (pxg_bluestein.hip);
// the launchers above hand over to them when n has no power-of-two kernel.
template <class R>
hipError_t launch_blue_pass_y(int n, const FftArgs<R> &a, int batch, hipStream_t s);
template <class R>
hipError_t launch_blue_pass_x(int n, const FftArgs<R> &a, int batch, hipStream_t s);
template <class R>
hipError_t launch_blue_fft2d(int n, const FftArgs<R> &a, int batch, hipStream_t s);
// Bluestein transform length for n points: the power of two >= 2n - 1 (at least 16)
int blue_length(int n);
template <class R>
hipError_t launch_cube_store(const cx<R> *field, size_t fstride, int n, int cube_n,
                             cx<double> *cube, size_t cube_fstride, int slice, int batch,
                             hipStream_t s);
// Far field of a 3D field (pxg_farfield.hip).  Windows are index ranges of the
// fftshifted M-point spectrum; M = D * NP with NP the transform length.
struct FfArgs {
  const void *field;       // n x n, column-major
  int field_c64;           // field stored as complex64 (else complex128)
  const void *ref;         // subtracted first (vacuum run), or null
  int ref_c64;
  int n;                   // rows = columns of the field
  int M, D;                // signal length, M / NP
  const cd *tw_m;          // exp(-2 pi i m / M), m < M
  const cd *tw_np;         // exp(-2 pi i m / NP), m < NP
  int ylo, ny;             // row window (y): shifted indices ylo .. ylo + ny - 1
  int xlo, nx;             // column window (x)
  cd *tmp;                 // column-pass output, ny x n, column-major
  double *out;             // ny x nx intensities, column-major
  double inv_len;          // 1 / M
  int pad_const;           // constant padding: entries n .. M-1 of a line repeat its entry 0
  // linear-fit padding: two least-squares lines through a line's entries 0..9, one per
  // padded segment; per fit grid the deviations x_k - mean(x), mean(x) and mean of the
  // squared deviations (the fit's x-only terms), the grid step and the end point
  int pad_linear;
  int lin_L;               // entries per padded segment, (M - n) / 2
  double fit1_dx[10], fit1_mx, fit1_mdx2;
  double fit2_dx[10], fit2_mx, fit2_mdx2;
  double lin_dxf;          // step of the first fit grid
  double lin_x2end;        // last point of the second fit grid
  int lin_wrap;            // first segment's ramp with the reference's unsigned index difference
};
hipError_t launch_far_field(const FfArgs &a, int np, hipStream_t s);
// Signal lengths that are not a power of two: the window's DFT sums directly.
hipError_t launch_far_field_direct(const FfArgs &a, hipStream_t s);
hipError_t launch_ff_resize(const double *in, int rows, int cols, double *out, int er, int ec,
                            hipStream_t s);
// Trapezoid sums of |u|^2 over per-step index ranges of a stored 2D solution
// (pxg_post.hip): one work-group per entry.
hipError_t launch_trapz_intensity(const void *sol, int c64, int nx, int n, const int *steps,
                                  const int *ixs, const int *ixe, double *sums, hipStream_t s);
// Field quantities of n cells (pxg_post.hip): q = 0 |u|, 1 |u|^2, 2 arg u, 3 Re u.
hipError_t launch_field_quantity(const void *src, int c64, size_t n, int q, double *out, hipStream_t s);
// Work-groups of the extrema pass of launch_field_uint8; `part` holds 2 doubles per group.
int extrema_groups(size_t n);
// 8-bit |u| (linear) or log|u| images of n cells; lo / hi are logarithms.
hipError_t launch_field_uint8(const void *src, int c64, size_t n, double *part, int log_scale,
                              int lo_set, double lo, int hi_set, double hi, unsigned char *out,
                              hipStream_t s);
// Lines per work-group for an LDS budget of lds_cap bytes (FP64 lines).
int max_lines(int n, int lds_cap);
// Stream of a plan and device of a context (server internals).
hipStream_t plan_stream(pxg_plan plan);
int ctx_device(pxg_ctx ctx);
bool fft_size_supported(int n);

}  // namespace pxg
