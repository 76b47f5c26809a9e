/*
 * paxoracle.h -- CPU FP64 restatement of the PaxPro propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * paxpro_amd/ and the timed CPU baseline of bench.py.  Only tests/, the smoke
 * check of __graft_entry__.py and the cpu_baseline leg of bench.py may load it;
 * the product library (libpaxgpu.so) never links or calls it.
 *
 * Every routine restates the behaviour of a reference file (cited per function
 * in paxoracle.c) in plain C99 with double _Complex arithmetic, keeping the
 * reference's loop structure: OpenMP over cells, a single-threaded FFT (the
 * reference links fftw3_omp but never enables FFTW threads), a serial Thomas
 * sweep and a material lookup through a function pointer per cell (the
 * reference's virtual MaterialFunction::getXrayMatProp).
 *
 * Pinning status of each function is listed in the header comment of
 * paxoracle.c; functions with no independent check say "parity unpinned".
 *
 * Layout conventions (Armadillo, column-major):
 *   2D field    : Nx complex values, index ix.
 *   3D field    : matrix of Ny rows (y) x Nx cols (x); element (r,c) at c*Ny + r.
 *   2D solution : Nx rows x Nz cols; element (ix,iz) at iz*Nx + ix.
 *   3D cube     : (nx' rows, nx' cols, slices); element (r,c,s) at s*nx'*nx' + c*nx' + r.
 */
#ifndef PAXORACLE_H
#define PAXORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_c;

/* ---- material ------------------------------------------------------------ */
/* Per-point (delta, beta) provider; n = 1 - delta + i beta.  2D solvers call it
 * with y = 0 (ParaxialSimulation::getXrayMatProp(x,z), src/paraxialSimulation.cpp:513-524). */
typedef void (*orc_matfn)(const void *ctx, double x, double y, double z,
                          double *delta, double *beta);
typedef struct {
  orc_matfn fn;      /* NULL => delta = beta = 0 everywhere */
  const void *ctx;
} orc_material;

/* ---- CSG (src/materialFunction.cpp, src/geometry.cpp, src/shape.cpp) ----- */
enum { ORC_SHAPE_SPHERE = 0, ORC_SHAPE_BOX = 1, ORC_SHAPE_CYLINDER = 2 };
enum { ORC_OP_UNION = 0, ORC_OP_DIFFERENCE = 1 };

typedef struct {
  int type;          /* ORC_SHAPE_* */
  int op;            /* operation of this shape inside its part */
  double T[16];      /* 4x4 world->shape transform, row-major T[4*i+j] */
  double p[3];       /* sphere: r ; box: Lx,Ly,Lz ; cylinder: r1,r2,height */
} orc_shape;

typedef struct {
  int first_shape, n_shapes;
  int op;            /* operation of this part inside its module */
  double delta, beta;
} orc_part;

typedef struct {
  int first_part, n_parts;
} orc_module;

typedef struct {
  const orc_shape *shapes;
  const orc_part *parts;
  const orc_module *modules;
  int n_modules;
  int is_reference_run;   /* CSGMaterial::isReferenceRun */
} orc_csg;

void orc_mat4_identity(double T[16]);
void orc_shape_translate(double T[16], double x, double y, double z);
void orc_shape_rotate(double T[16], double angle_deg, int axis /*0=X,1=Y,2=Z*/);
void orc_shape_scale(double T[16], double factor, int axis);
void orc_shape_transform(const double T[16], double *x, double *y, double *z);
int  orc_shape_inside(const orc_shape *s, double x, double y, double z);
int  orc_part_inside(const orc_csg *g, const orc_part *p, double x, double y, double z);
void orc_csg_eval(const orc_csg *g, double x, double y, double z, double *delta, double *beta);
void orc_csg_matfn(const void *ctx, double x, double y, double z, double *d, double *b);

/* ---- curved 2D waveguide (src/curvedWaveGuide2D.cpp, src/waveGuideFDSimulation.cpp) */
typedef struct {
  double R, width;
  double clad_delta, clad_beta;
  double in_delta, in_beta;    /* inside material (0,0 when none is set) */
  double wglength;             /* default 1e10 */
  double dx, dz;               /* grid steps used for the neighbour test */
  int use_smoothed;
} orc_curved_wg;
int  orc_curved_inside(const orc_curved_wg *w, double x, double z);
void orc_curved_matfn(const void *ctx, double x, double y, double z, double *d, double *b);

/* ---- straight slab, the MaterialFunction of Examples/waveguide2D.cpp:9-27 - */
typedef struct {
  double x_lo, x_hi;           /* core is x_lo < x < x_hi */
  double core_delta, core_beta, clad_delta, clad_beta;
} orc_slab;
void orc_slab_matfn(const void *ctx, double x, double y, double z, double *d, double *b);

/* ---- discretisation (src/paraxialSimulation.cpp:68-120,250-263) --------- */
unsigned orc_nodes(double min, double max, double step);

/* ---- sources (src/gaussianBeam.cpp, src/planeWave.cpp) ------------------- */
typedef struct {
  double k;                  /* wavenumber of the source */
  double amplitude, z0;
  double waist, cx, cy, angle_x_deg, angle_y_deg;
} orc_gauss;
orc_c orc_gauss_2d(const orc_gauss *g, double x, double z);
orc_c orc_gauss_3d(const orc_gauss *g, double x, double y, double z);
typedef struct { double k, amplitude, z0, angle_rad; } orc_plane;
orc_c orc_plane_wave(const orc_plane *p, double x, double z);
/* out-pointer forms for FFI callers; fill3d writes values(j,i) = src(x_i, y_j, 0)
 * for an n x n column-major field (src/paraxialSimulation.cpp:381-395). */
void orc_gauss_2d_v(const orc_gauss *g, double x, double z, double out[2]);
void orc_plane_wave_v(const orc_plane *p, double x, double z, double out[2]);
void orc_gauss_fill3d(const orc_gauss *g, double xmin, double dx, double ymin, double dy,
                      unsigned n, orc_c *out);

/* ---- FFT (FFTW semantics: unnormalised, sign -1 forward) ----------------- */
void orc_fft1d(orc_c *a, unsigned n, int sign);
void orc_fft2d(orc_c *a, unsigned nrows, unsigned ncols, int sign);

/* ---- Thomas (include/thomasAlgorithm.tpp:2-32) ---------------------------- */
void orc_thomas(orc_c *diag, const orc_c *sub, orc_c *rhs, unsigned n);

/* ---- grids ---------------------------------------------------------------- */
typedef struct {
  double xmin, xmax, dx;
  double ymin, ymax, dy;
  double zmin, zmax, dz;
  double k;
} orc_grid;

/* Quirk switches: 0 reproduces the reference, a set bit selects the physics form. */
enum {
  ORC_FIX_ADI_Y_DIAG   = 1u << 0,  /* src/alternatingDirectionSolver.cpp:103 parenthesis */
  ORC_FIX_REFR_SAMPLES = 1u << 1   /* src/fftSolver3D.cpp:262 skips sample n-1 */
};

/* ---- FFT split-step 3D (src/fftSolver3D.cpp:52-156,253-273,327-374) ------ */
typedef struct { unsigned npix_x, npix_y; double inv_x, inv_y; } orc_absorber3d;
/* field: Ny x Nx (Ny = Nx), propagated from step first_step-1 to first_step+nsteps-1.
 * cube (optional): block averages stored when step % dsz == 0 (src/solver3D.cpp:78-125). */
void orc_fft3d_run(const orc_grid *g, const orc_material *m, const orc_absorber3d *abs,
                   unsigned n_refr, unsigned quirks, orc_c *field, unsigned n,
                   unsigned first_step, unsigned nsteps,
                   orc_c *cube, unsigned cube_n, unsigned cube_slices, unsigned dsz);
void orc_cube_store(const orc_c *field, unsigned n, orc_c *cube, unsigned cube_n,
                    unsigned cube_slices, unsigned dsz, unsigned step);

/* ---- FFT split-step 2D (src/fftSolver2D.cpp:15-74) ------------------------ */
void orc_fft2d_run(const orc_grid *g, const orc_material *m, orc_c *field, unsigned nx,
                   unsigned first_step, unsigned nsteps, orc_c *solution, unsigned nz);

/* ---- Crank-Nicolson (src/crancNicholson.cpp:17-148) ----------------------- */
enum { ORC_BC_DIRICHLET = 0, ORC_BC_TRANSPARENT = 1 };
/* absorb_npix > 0 enables the edge absorber extension (Absorber::apply
 * semantics, src/absorber.cpp:4-12) after every step; 0 = pure reference. */
void orc_cn_run(const orc_grid *g, const orc_material *m, int bc, orc_c *field, unsigned nx,
                unsigned first_step, unsigned nsteps, orc_c *solution, unsigned nz,
                unsigned absorb_npix, double absorb_inv);
/* One CN step on explicit per-cell material planes (d1,b1 at z; d0,b0 at z-dz). */
void orc_cn_step_planes(double rho, double r, int bc, const orc_c *prev, orc_c *cur, unsigned nx,
                        const double *d1, const double *b1, const double *d0, const double *b0);

/* ParaxialEquation (include/paraxialEquation.hpp:16-35): F, G, H, J at (x, z). */
typedef void (*orc_eqfn)(const void *ctx, double x, double z, double out[4]);
typedef struct {
  orc_eqfn fn;     /* NULL: the default equation F = G = H = 1, J = 0 */
  const void *ctx;
} orc_equation;
/* Coefficients c0 + cx x + cz z + cinv / x for F, G, H, J (rows of c[4][4]). */
typedef struct { double c[4][4]; } orc_poly_eq;
void orc_poly_eqfn(const void *ctx, double x, double z, double out[4]);
/* Per-row equation values of one plane, 5 per row: F(x,z), G(x,z),
 * H(x + dx/2, z), H(x - dx/2, z), J(x,z); NULL = defaults.  e1 at z, e0 at z - dz. */
void orc_cn_step_eq(double rho, double r, int bc, const orc_c *prev, orc_c *cur, unsigned nx,
                    const double *d1, const double *b1, const double *d0, const double *b0,
                    const double *e1, const double *e0);
/* orc_cn_run with a ParaxialEquation. */
void orc_cn_run_eq(const orc_grid *g, const orc_material *m, const orc_equation *eq, int bc,
                   orc_c *field, unsigned nx, unsigned first_step, unsigned nsteps,
                   orc_c *solution, unsigned nz, unsigned absorb_npix, double absorb_inv);

/* ---- ADI (src/alternatingDirectionSolver.cpp:8-234) ----------------------- */
void orc_adi_run(const orc_grid *g, const orc_material *m, int use_tbc, unsigned quirks,
                 orc_c *field, unsigned n, unsigned first_step, unsigned nsteps,
                 orc_c *cube, unsigned cube_n, unsigned cube_slices, unsigned dsz);

/* ---- helpers -------------------------------------------------------------- */
void orc_absorb_line(orc_c *f, unsigned len, unsigned stride, unsigned npix, double inv);
double orc_rel_l2(const orc_c *a, const orc_c *ref, uint64_t n);
double orc_rel_l2_intensity(const orc_c *a, const orc_c *ref, uint64_t n);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
