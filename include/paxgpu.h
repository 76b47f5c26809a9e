/*
 * paxgpu.h -- C ABI of the MI355X (gfx950) propagation hot path of PaxPro.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = PXG_OK), no exceptions across the boundary, no torch types.  The
 * reference exposes this path as a C++ virtual class API; the C++ drop-in
 * classes in paxpro_amd/cpp (Solver, Solver2D, Solver3D, FFTSolver2D,
 * FFTSolver3D, CrankNicholson, ADI, ThomasAlgorithm) call these entry points
 * from their solveStep()/solve()/step() overrides.  Each entry point names the
 * reference interface it replaces.
 *
 * Data layout follows Armadillo (column-major, interleaved complex<double>),
 * so cx_vec/cx_mat memptr() maps 1:1:
 *   2D field : Nx complex values.
 *   3D field : Ny x Nx matrix (Ny == Nx; include/solver3D.hpp:56, src/solver3D.cpp:24),
 *              element (row r = y, col c = x) at c*Ny + r.
 * Host buffers are caller-owned and always complex128; device buffers are
 * library-owned and hold complex128 (PXG_PREC_FP64) or complex64
 * (PXG_PREC_MIXED) fields.
 *
 * Threading: a ctx may be shared by threads (the server does); a plan is used
 * by one thread at a time and owns one HIP stream (default priority).
 */
#ifndef PAXGPU_H
#define PAXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  PXG_OK = 0,
  PXG_ERR_INVALID = 1,      /* bad argument / shape mismatch (reference: std::runtime_error) */
  PXG_ERR_UNSUPPORTED = 2,  /* size or option not implemented on the GPU path */
  PXG_ERR_BUDGET = 3,       /* HBM budget of the context exceeded */
  PXG_ERR_DEVICE = 4,       /* HIP runtime error (message in pxg_last_error) */
  PXG_ERR_STATE = 5,        /* call order violated (e.g. step before set_field) */
  PXG_ERR_CANCELLED = 6,    /* server job cancelled */
  PXG_ERR_INJECTED = 7      /* failure forced from the host for testing */
};

typedef struct pxg_ctx_s *pxg_ctx;
typedef struct pxg_plan_s *pxg_plan;

/* ---- solver kinds and options --------------------------------------------- */
enum { PXG_FFT2D = 0, PXG_FFT3D = 1, PXG_CN = 2, PXG_ADI = 3 };
/* FP64: complex128 fields and FP64 arithmetic throughout (the plain path).
 * MIXED: complex64 fields in HBM (half the bytes per pass); every kernel
 * widens to FP64 on load and computes in FP64 with FP64 tables, rounding to
 * complex64 once per store.  FP32 arithmetic was measured to accumulate a
 * coherent error (the rounded twiddles and propagator act as a fixed operator
 * perturbation, growing linearly with steps), FP64 arithmetic does not.
 * CN plans keep complex128 state in both precisions: their march is
 * latency-bound, and a slowly varying 1D field rounded to complex64 every step
 * accumulates a coherent error (measured 5.7e-5 after 1000 steps). */
enum { PXG_PREC_FP64 = 0, PXG_PREC_MIXED = 1 };
enum { PXG_BC_DIRICHLET = 0, PXG_BC_TRANSPARENT = 1 };   /* Solver2D::BC_t */
/* Quirk switches: 0 reproduces the reference; a set bit selects the physics form. */
enum {
  PXG_FIX_ADI_Y_DIAG = 1u << 0,    /* src/alternatingDirectionSolver.cpp:103 parenthesis */
  PXG_FIX_REFR_SAMPLES = 1u << 1,  /* src/fftSolver3D.cpp:262 skipped trapezoid sample */
  PXG_FIX_LINEAR_PAD = 1u << 2     /* src/farFieldPost.cpp:288 unsigned index difference */
};

typedef struct {
  int kind;              /* PXG_FFT2D / PXG_FFT3D / PXG_CN / PXG_ADI */
  int precision;         /* PXG_PREC_* */
  int nx;                /* transverse nodes; 3D grids are nx x nx */
  int nz;                /* longitudinal nodes (steps run 1..nz-1) */
  int batch;             /* independent fields sharing this geometry (>= 1) */
  double xmin, dx;       /* transverse discretisation */
  double ymin, dy;       /* vertical discretisation (3D) */
  double zmin, dz;       /* longitudinal discretisation */
  double k;              /* wavenumber (nm^-1); per-field values via pxg_set_wavenumbers */
  int bc;                /* CN boundary condition, PXG_BC_* */
  int use_tbc;           /* ADI::useTBC */
  unsigned absorb_npix_x, absorb_npix_y;   /* FFTSolver3D::absorbingBC (0 = off); CN uses x */
  double absorb_inv_x, absorb_inv_y;       /* inverse damping length in 1/pixel */
  unsigned quirks;       /* PXG_FIX_* */
  unsigned n_refr;       /* samples of the border-crossing integral (reference: 10) */
  int store_dsx;         /* solution storage: transverse block size (0 = none) */
  int store_dsz;         /* solution storage: every store_dsz-th step */
} pxg_problem;

/* ---- CSG material (layout shared with the oracle) -------------------------- */
enum { PXG_SHAPE_SPHERE = 0, PXG_SHAPE_BOX = 1, PXG_SHAPE_CYLINDER = 2 };
enum { PXG_OP_UNION = 0, PXG_OP_DIFFERENCE = 1 };
typedef struct {
  int type, op;          /* PXG_SHAPE_*, PXG_OP_* (operation inside its part) */
  double T[16];          /* 4x4 world->shape transform, row-major */
  double p[3];           /* sphere r | box Lx,Ly,Lz | cylinder r1,r2,height */
} pxg_shape;
typedef struct {
  int first_shape, n_shapes, op, pad_;
  double delta, beta;
} pxg_part;
typedef struct {
  int first_part, n_parts;
} pxg_module;

typedef struct {
  double R, width;
  double clad_delta, clad_beta, in_delta, in_beta;
  double wglength;
  double dx, dz;
  int use_smoothed;
} pxg_curved_wg;

/* ---- context ---------------------------------------------------------------- */
/* Replaces nothing in the reference (it has no device): owns the device choice,
 * the HBM budget used for admission, and the last error string. */
int pxg_ctx_create(int device, size_t hbm_budget_bytes, pxg_ctx *out);
/* Plans keep their context alive: a context destroyed while plans of it live
 * is freed when the last of them is destroyed (no new plans can be made). */
int pxg_ctx_destroy(pxg_ctx ctx);
/* Bytes currently reserved by live plans of this context. */
size_t pxg_ctx_reserved(pxg_ctx ctx);
/* Message of the last failing call on this thread (valid until the next call). */
const char *pxg_last_error(void);
/* Library build information (gfx target, kernels present). */
const char *pxg_version(void);

/* ---- plans ------------------------------------------------------------------ */
/* Replaces Solver2D::setSimulator (src/solver2D.cpp:21-42) and
 * Solver3D::setSimulator (src/solver3D.cpp:19-44): allocates the device fields
 * and solution storage for the problem. */
int pxg_plan_create(pxg_ctx ctx, const pxg_problem *prob, pxg_plan *out);
int pxg_plan_destroy(pxg_plan plan);
/* Reconfigures a plan for a new problem of the same kind, precision, nx, batch
 * and storage shape without reallocating device memory (grids, wavenumber,
 * absorber and quirks may change); the material resets to none and the field
 * must be set again.  Extension used by the server to recycle plans. */
int pxg_plan_reset(pxg_plan plan, const pxg_problem *prob);
/* Bytes of device memory the plan reserves (for admission by HBM budget). */
int pxg_plan_bytes(const pxg_problem *prob, size_t *bytes);
/* Work-group shape of the two FFT3D passes: lines per work-group of the x pass
 * and of the y pass (powers of two 1..16 within the LDS of one CU; 0 keeps the
 * current value).  Results do not depend on it.  FFT3D plans only.  No
 * reference counterpart: FFTW chooses its own plan shapes
 * (src/fftSolver3D.cpp:61-62 plans with FFTW_ESTIMATE). */
int pxg_plan_set_lines(pxg_plan plan, int lines_x, int lines_y);
int pxg_plan_get_lines(pxg_plan plan, int *lines_x, int *lines_y);
/* Tile shapes of the ADI half-step kernels (tile-Jacobi path): index of the
 * x-implicit shape (lines x positions: 0: 16x128, 1: 8x256, 2: 16x256, 3: 32x64,
 * 4: 32x128, 5: 8x128) and of the y-implicit shape (0: 4x256, 1: 1x256, 2: 2x256,
 * 3: 8x256, 4: 2x512, 5: 1x2048, the whole line up to 2048 points); -1 keeps the
 * current one.  Results do not depend on them.  ADI plans only.  No reference
 * counterpart (src/alternatingDirectionSolver.cpp:8-134 sweeps one line at a
 * time). */
#define PXG_ADI_TILES 6
int pxg_plan_set_adi_tiles(pxg_plan plan, int x_tile, int y_tile);
int pxg_plan_get_adi_tiles(pxg_plan plan, int *x_tile, int *y_tile);
/* Autotuner, in the manner of FFTW's measured planning: times every supported
 * shape of both FFT3D passes (lines per work-group) and, for batches that can
 * march as two halves on two streams (16 or more fields), steps of the march both
 * ways; or every tile shape of both ADI half-steps; on this plan with its current
 * material, and keeps the fastest.  Host-sampled planes that change with z are
 * timed at the first steps they cover (none covered: nothing is timed).  The
 * choice is cached in the context for plans of the same kind, nx, batch,
 * precision and material kind, and no choice changes a result.  Overwrites
 * (zeroes) the field: call it after the material and before pxg_set_field.
 * reps <= 0 means 5 timed launches per shape.  A no-op for the other kinds. */
int pxg_plan_autotune(pxg_plan plan, int reps);
/* FFT3D z-step loops as HIP graphs: with enable != 0, pxg_step captures the
 * kernels of a call into a graph on first use and replays it when a later call has
 * the same steps, work-group shapes, material state and kernel arguments (up to 8
 * shapes per plan are kept).  Results are identical to the stream enqueue.  No
 * reference counterpart: FFTW executes its plans from the host
 * (src/fftSolver3D.cpp:95,119). */
int pxg_plan_set_graphs(pxg_plan plan, int enable);
/* Job plan: n independent problems of one kind in one plan, each with its own
 * grid, wavenumber and material (uneven Nx, photon energies, waveguides), all
 * advanced by one launch per pxg_step.  Field/solution index b = job b.
 * Supported kind: PXG_CN (one work-group per job).  No reference counterpart:
 * the reference solves one problem per Solver object; this batches many
 * CrankNicholson::solveStep marches (src/crancNicholson.cpp:17-105). */
int pxg_plan_create_jobs(pxg_ctx ctx, const pxg_problem *probs, int n, pxg_plan *out);
/* Device-evaluated CurvedWaveGuideFD material of one job of a job plan. */
int pxg_set_job_material_curved_wg(pxg_plan plan, int job, const pxg_curved_wg *wg);
/* Solution dimensions of field/job b (job plans have per-job sizes). */
int pxg_job_solution_dims(pxg_plan plan, int job, int dims[3]);
/* Per-field wavenumbers of a batched plan (photon energies); n == batch. */
int pxg_set_wavenumbers(pxg_plan plan, const double *k, int n);

/* Replaces Solver2D::setInitialConditions / setLeftBC (src/solver2D.cpp:51-69) and
 * Solver3D::setInitialConditions (src/solver3D.cpp:127-146); stores slice/column 0
 * of the solution like copyCurrentSolution(0). */
int pxg_set_field(pxg_plan plan, int b, const double *c128);
/* Replaces getLastSolution / getLastSolution3D (include/solver2D.hpp:69,
 * include/solver3D.hpp:36): copies the current field of batch entry b. */
int pxg_get_field(pxg_plan plan, int b, double *c128);
/* Replaces getSolution (2D: Nx x Nz matrix, include/solver2D.hpp:41) and
 * getSolution3D (cube of (nx/dsx)^2 x (nz/dsz + 1), src/solver3D.cpp:201-204). */
int pxg_get_solution(pxg_plan plan, int b, double *c128, size_t n_complex);
int pxg_solution_dims(pxg_plan plan, int dims[3]);

/* ---- materials (the per-cell getXrayMatProp call-back, evaluated on device) -- */
/* Replaces CSGMaterial::getXrayMatProp (src/materialFunction.cpp:3-13,
 * src/geometry.cpp:91-108,212-234, src/shape.cpp:126-187). */
int pxg_set_material_csg(pxg_plan plan, const pxg_shape *shapes, int n_shapes,
                         const pxg_part *parts, int n_parts, const pxg_module *modules,
                         int n_modules, int is_reference_run);
/* Replaces CurvedWaveGuideFD / WaveGuideFDSimulation::getXrayMatProp
 * (src/waveGuideFDSimulation.cpp:219-265, src/curvedWaveGuide2D.cpp:49-52,179-205). */
int pxg_set_material_curved_wg(pxg_plan plan, const pxg_curved_wg *wg);
/* Host-sampled material for any MaterialFunction subclass the device cannot
 * evaluate.  db holds n_planes planes of (delta, beta) pairs per cell, in the
 * field's layout.  Plane p is what the solver consumes as its material input:
 *   CN    : plane p = material at z = zmin + p*dz (step iz uses p = iz, iz-1);
 *   FFT2D : plane p = material at z_p + dz/2 for step p;
 *   FFT3D : plane p = effective (delta, beta) of step p after the crossing integral;
 *   ADI   : plane 2p = x-pass of step p (z_p), plane 2p+1 = y-pass (z_p + dz/2).
 * z_invariant != 0 means one plane serves every step. */
int pxg_set_material_planes(pxg_plan plan, const double *db, int n_planes, int first_plane,
                            int z_invariant);
int pxg_set_material_none(pxg_plan plan);
/* Replaces Solver2D::setEquation (src/solver2D.cpp:44-49) with a ParaxialEquation
 * other than the default, as CrankNicholson::solveStep evaluates it
 * (src/crancNicholson.cpp:41-60,83-88): the caller samples it on the host.  coef holds
 * n_planes planes of 5 values per cell: F(x,z), G(x,z), H(x + dx/2, z), H(x - dx/2, z),
 * J(x,z), x = xmin + i dx; plane p is z = zmin + p dz (step iz uses planes iz and
 * iz - 1); z_invariant != 0: one plane for every z.  coef = NULL restores the default
 * (F = G = H = 1, J = 0).  Single CN plans only. */
int pxg_set_equation_planes(pxg_plan plan, const double *coef, int n_planes, int first_plane,
                            int z_invariant);

/* ---- stepping ----------------------------------------------------------------- */
/* Replaces Solver2D::step/solve (src/solver2D.cpp:224-250) and Solver3D::step/solve
 * (src/solver3D.cpp:148-199): advances every field of the plan through steps
 * first_step .. first_step+nsteps-1, storing solution slices as configured.
 * Enqueued on the plan's stream; returns after enqueue (use pxg_sync). */
int pxg_step(pxg_plan plan, unsigned first_step, unsigned nsteps);
/* Several separate FFT3D plans advanced in shared launches: plan i through steps
 * first_steps[i] .. first_steps[i]+nsteps-1, one launch per pass and z-step covering
 * every plan, each plan with its own grid, wavenumber, material and absorber.  The
 * plans share a context, a power-of-two nx and the precision, and have batch 1 and no
 * solution storage; at most 8.  Each plan's field ends as pxg_step(plans[i],
 * first_steps[i], nsteps) leaves it (the same kernels per cell).  Enqueued on
 * plans[0]'s stream after every plan's earlier work; every plan's stream waits for
 * it.
 * No counterpart in the reference, which runs one simulation per object
 * (src/paraxialSimulation.cpp:40-48): for servers that batch separate requests. */
int pxg_step_group(pxg_plan *plans, int n_plans, const unsigned *first_steps, unsigned nsteps);
int pxg_sync(pxg_plan plan);
/* Same as pxg_step, bracketing every kernel launch with HIP events on the plan's
 * stream; reports the mean duration per kernel name (ms) and the wall time of
 * the whole enqueue-to-completion interval.  names[i] point to static strings. */
typedef struct { const char *name; double mean_ms; int launches; double bytes_per_launch; } pxg_kernel_stat;
int pxg_step_profiled(pxg_plan plan, unsigned first_step, unsigned nsteps, pxg_kernel_stat *stats,
                      int max_stats, int *n_stats, double *total_ms);
/* Device pointer of field b (advanced use: zero-copy serving, benchmarks). */
int pxg_field_device_ptr(pxg_plan plan, int b, void **dptr);

/* ---- far field (3D plans: FFT3D, ADI) ----------------------------------------- */
/* Replaces post::FarField::result(const Solver&, arma::mat&) (src/farFieldPost.cpp:141-230)
 * as GenericScattering::getFarField calls it (src/genericScattering.cpp:235-238): the
 * scattered intensity of field b minus field ref_b of plan ref (the vacuum run; ref may
 * be NULL), zero padding, signal length M = max(pad, nx), angular window
 * [phi_min, phi_max] degrees mapped by farFieldAngleToIndx (src/farFieldPost.cpp:80-103)
 * with dy for rows and dx for columns and the wavenumber of field b, |FFT|^2 / M, then
 * PostProcessingModule::resizeMatrix (src/postProcessing.cpp:29-49) when export_rows and
 * export_cols are both nonzero (GenericScattering sets 512 x 512).  M must be a power of
 * two (the reference accepts any length; other lengths return PXG_ERR_UNSUPPORTED).
 * pxg_far_field_dims gives the result size (rows, cols); pxg_far_field writes it to out
 * (host, column-major doubles) after the plans' pending work. */
int pxg_far_field_dims(pxg_plan plan, int b, unsigned pad, double phi_min, double phi_max,
                       unsigned export_rows, unsigned export_cols, int dims[2]);
int pxg_far_field(pxg_plan plan, int b, pxg_plan ref, int ref_b, unsigned pad, double phi_min,
                  double phi_max, unsigned export_rows, unsigned export_cols, double *out);
/* The same with post::FarField::setPadding (src/farFieldPost.cpp:249-321), applied to every
 * column and to every conjugated row of the column pass's window: PXG_PAD_ZERO (the default,
 * = pxg_far_field); PXG_PAD_CONSTANT (entries n .. M-1 repeat the line's entry 0);
 * PXG_PAD_LINEAR_FIT (linearPadding, :265-312, which the reference marks untested: least-
 * squares lines through entries 0..9 (gsl_fit_linear) extend the line over its last
 * (M-n)/2 entries and over (M-n)/2 entries after n, the second with the fit of the same
 * entries against the grid's last ten x positions; needs nx >= 10).  The first segment's
 * abscissa is dxf*(i - n_elem + 1) with Armadillo's unsigned 64-bit n_elem (:288): by
 * default it wraps as in the reference (entries M-L .. M-2 sit near dxf*2^64, entry M-1
 * at 0); the plan's quirk bit PXG_FIX_LINEAR_PAD takes the signed ramp dxf*(i - M + 1). */
enum { PXG_PAD_ZERO = 0, PXG_PAD_LINEAR_FIT = 1, PXG_PAD_CONSTANT = 2 };
int pxg_far_field_padded(pxg_plan plan, int b, pxg_plan ref, int ref_b, unsigned pad, int padding,
                         double phi_min, double phi_max, unsigned export_rows, unsigned export_cols,
                         double *out);

/* ---- per-step reductions (2D plans storing the full solution) ----------------- */
/* Replaces post::Transmittivity::trapezoidalIntegrateIntensityZ (src/transmittivity.cpp:30-42),
 * which post::Transmittivity::compute (src/transmittivity.cpp:44-83) applies after every step
 * of CurvedWaveGuideFD::solve (src/curvedWaveGuide2D.cpp:207-219), for many steps at once:
 * sums[i] = |u(s)|^2 + |u(e)|^2 + 2 sum_{s<j<e} |u(j)|^2 over column steps[i] of the stored
 * solution of field/job b, s = ix_start[i], e = ix_end[i].  The caller applies the
 * reference's 0.5 (x_e - x_s)/(e - s) and the normalisation.  Needs store_dsz = 1 and
 * store_dsx = 1 (every step at full resolution), 0 <= s, e < nx and 0 <= steps[i] < nz. */
int pxg_intensity_trapezoid(pxg_plan plan, int b, int n, const int *steps, const int *ix_start,
                            const int *ix_end, double *sums);

/* ---- field quantities of the post-processing modules ------------------------- */
/* Replace the result() methods of post::Intensity, Phase, ExitField, ExitIntensity,
 * ExitPhase, IntensityUint8 and LogIntensityUint8 (src/postProcessMod.cpp:7-151), computed
 * on the device from what the solver holds there:
 *   source PXG_SRC_FIELD     the current field b: Solver::getLastSolution (2D, a vector
 *                            of nx) or getLastSolution3D (3D, an nx x nx matrix);
 *   source PXG_SRC_SOLUTION  the stored solution of field/job b: Solver::getSolution (2D,
 *                            nx x stored steps) or getSolution3D (3D, the block-averaged
 *                            cube); needs store_dsz > 0.
 * quantity: PXG_Q_ABS |u| (Intensity), PXG_Q_ABS2 |u|^2 (ExitIntensity), PXG_Q_ARG
 * atan2(Im u, Re u) (Phase, ExitPhase), PXG_Q_REAL Re u (ExitField).  When the result is
 * a matrix and export_rows, export_cols are both nonzero,
 * PostProcessingModule::resizeMatrix (src/postProcessing.cpp:29-49) follows: block means
 * down to export_rows x export_cols, unless either size reaches the matrix's (then it is
 * kept).  Vectors and cubes are never resized (export sizes ignored), as in the modules.
 * pxg_field_quantity_dims gives (rows, cols, slices); pxg_field_quantity writes that many
 * doubles to out (host, column-major) after the plan's pending work.
 * pxg_field_uint8 writes rows*cols*slices bytes: log_scale 0 is IntensityUint8,
 * 255 |u| / max|u|; log_scale 1 is LogIntensityUint8, log|u| mapped from [bot, top] to
 * [0, 255] and clamped, top = log max|u| lowered to log(*max_value) when max_value is
 * given (setMaxValue), bot = log min|u| raised to log(*min_value) when min_value is given
 * (setMinValue); NULL = not set.  Values truncate to unsigned char as the reference's
 * assignment does; where the reference converts a NaN (a field of zeros, top = bot) the
 * byte is 0.  No resize (the reference's 8-bit modules have none). */
enum { PXG_SRC_FIELD = 0, PXG_SRC_SOLUTION = 1 };
enum { PXG_Q_ABS = 0, PXG_Q_ABS2 = 1, PXG_Q_ARG = 2, PXG_Q_REAL = 3 };
int pxg_field_quantity_dims(pxg_plan plan, int b, int source, unsigned export_rows,
                            unsigned export_cols, int dims[3]);
int pxg_field_quantity(pxg_plan plan, int b, int source, int quantity, unsigned export_rows,
                       unsigned export_cols, double *out);
int pxg_field_uint8(pxg_plan plan, int b, int source, int log_scale, const double *min_value,
                    const double *max_value, unsigned char *out);

/* ---- batched tridiagonal solve ---------------------------------------------- */
/* Replaces ThomasAlgorithm::solve<cdouble> (include/thomasAlgorithm.tpp:2-32) for
 * many independent systems: diag/sub/rhs are host complex128 arrays of nsys
 * systems of n rows (sub has n-1 entries per system, sup == sub); layout 0 =
 * system-contiguous, 1 = interleaved (row i of system s at i*nsys + s).  The
 * solution is written to diag; rhs is clobbered, as in the reference. */
enum { PXG_LAYOUT_CONTIGUOUS = 0, PXG_LAYOUT_INTERLEAVED = 1 };
int pxg_thomas_batched(pxg_ctx ctx, int layout, int nsys, int n, double *diag, const double *sub,
                       double *rhs);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pxg_thomas_batched on DEVICE arrays diag/sub/rhs (same
 * layouts and results, bit for bit), enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) and returned without waiting.  Its scratch (the eliminated
 * super-diagonal, 16 bytes per row) is a static device array of
 * PXG_THOMAS_NOALLOC_MAX_ROWS rows (16 MiB) that each device holds for the life
 * of the process, so the call allocates nothing; nsys * n above that returns
 * PXG_ERR_UNSUPPORTED.  All calls on one device share that array: calls that
 * overlap on one device (other streams, other threads) race. */
#define PXG_THOMAS_NOALLOC_MAX_ROWS (1u << 20)
int pxg_thomas_batched_noalloc(pxg_ctx ctx, int layout, int nsys, int n, double *diag,
                               const double *sub, double *rhs, void *stream);

/* ---- serving ------------------------------------------------------------------ */
/* A persistent request loop (no counterpart in the reference, which runs one
 * simulation per process): jobs of mixed kinds and grid sizes are admitted by
 * HBM budget into slots (one plan and one stream each), advance in chunks of
 * z-steps so that the work of different jobs overlaps on the device, and leave
 * as soon as their result is on the host; queued jobs are admitted as slots
 * and memory free up.  Latency is measured from submission to result. */
typedef struct pxg_server_s *pxg_server;
enum { PXG_JOB_QUEUED = 0, PXG_JOB_RUNNING = 1, PXG_JOB_DONE = 2, PXG_JOB_FAILED = 3, PXG_JOB_CANCELLED = 4 };
enum { PXG_MAT_NONE = 0, PXG_MAT_CSG = 1, PXG_MAT_CURVED_WG = 2 };
typedef struct {
  pxg_problem prob;              /* batch must be 1 */
  int material;                  /* PXG_MAT_* */
  const pxg_shape *shapes;       /* CSG arrays, copied at submission */
  int n_shapes;
  const pxg_part *parts;
  int n_parts;
  const pxg_module *modules;
  int n_modules;
  pxg_curved_wg wg;
  const double *field;           /* initial field (complex128); read at admission */
  double *result;                /* exit field (complex128); written at completion */
  unsigned first_step, nsteps;
  int inject_failure_after_chunks;  /* < 0: none; >= 0 fails the job from the host (tests) */
  int inject_plan_failure;          /* PXG_INJECT_*: a plan failure forced from the host (tests) */
} pxg_job;
/* Host-forced plan failures of one job: its plan creation, or the recycling
 * (pxg_plan_reset) of a pooled plan for it, fails with PXG_ERR_INJECTED. */
enum { PXG_INJECT_NONE = 0, PXG_INJECT_CREATE = 1, PXG_INJECT_RESET = 2 };
typedef struct {
  uint64_t submitted, completed, failed, cancelled, rejected, running, queued;
  double p50_ms, p99_ms, mean_ms, max_ms;   /* submission -> result on host, completed jobs */
  double cell_steps;                        /* sum over completed jobs */
  double busy_s;                            /* wall time with at least one job running */
  size_t hbm_reserved, hbm_budget;
  uint64_t chunks;                          /* job chunks enqueued */
  uint64_t grouped_chunks;                  /* of them, in shared launches (pxg_step_group) */
} pxg_server_stats;
/* slots: jobs running at once; hbm_budget: bytes of plans (0 = 64 GiB);
 * chunk_steps: z-steps enqueued per job before the next job is served.
 * A job whose plan cannot be made or recycled fails with that error (its
 * reserved bytes are released); a job that does not fit the budget now waits
 * for running jobs to free memory.  The worker thread takes the server lock
 * only to move jobs between the queue and the running set: copies, plan
 * creation and synchronisation run without it, so submit/status/wait/cancel
 * never wait for device work. */
int pxg_server_create(pxg_ctx ctx, int slots, size_t hbm_budget, unsigned chunk_steps, pxg_server *out);
int pxg_server_submit(pxg_server s, const pxg_job *job, uint64_t *id);
int pxg_server_status(pxg_server s, uint64_t id, int *state);
/* Blocks until the job ends (timeout_s <= 0: no timeout); returns its status. */
int pxg_server_wait(pxg_server s, uint64_t id, double *latency_ms, double timeout_s);
int pxg_server_cancel(pxg_server s, uint64_t id);
int pxg_server_stats_get(pxg_server s, pxg_server_stats *st);
/* Chunks of up to max_group running FFT3D jobs (batch 1, no storage, one
 * power-of-two nx <= max_nx and precision, chunks of equal length) that are ready
 * together are enqueued as one pxg_step_group; a ready job of such a shape waits
 * for a partner whose chunk is in flight.  max_group 1 runs every chunk alone;
 * max_nx 0: any nx.  Defaults: 8 (the most a group holds) and 512 (larger grids
 * fill the device with one job's launches).  Results do not depend on them. */
int pxg_server_set_grouping(pxg_server s, int max_group, int max_nx);
int pxg_server_destroy(pxg_server s);

#ifdef __cplusplus
}
#endif
#endif
