/*
 * paxoracle.c -- CPU FP64 restatement of the PaxPro propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see paxoracle.h).  Written from the behaviour of
 * the reference files cited per function; no reference source is included.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_fft1d / orc_fft2d   brute-force O(N^2) DFT, numpy/pocketfft golden
 *                           vectors (tests/golden/fft_*.npz), Parseval, impulse.
 *   orc_thomas              the reference's own ThomasAlgorithm compiled from
 *                           include/thomasAlgorithm.tpp into oracle/_ref
 *                           (bit-exact), dense numpy solves, residuals.
 *   CSG transforms          the reference's known-answer tests
 *                           unittest/transformTest.cpp:4-65, analytic membership.
 *   orc_fft3d_run           free-space composition (n steps of dz == one step of
 *                           n*dz), homogeneous closed form u0*exp(-k(b+id)z),
 *                           brute-force crossing integral on a 1-cell sphere case.
 *   orc_fft2d_run           same closed forms in 1D.
 *   orc_cn_run              Cayley-transform norm conservation (delta=beta=0,
 *                           Dirichlet) and the uniform-field closed form.
 *   orc_adi_run             uniform-field closed forms per half-step, including
 *                           the y-pass parenthesis quirk.
 *   sources                 closed forms at z = z0.
 *   orc_cube_store          block average checked against numpy reshape-mean.
 *   orc_curved_matfn        analytic membership 0 < 2xR + z^2 < 2wR.
 * Parity unpinned: exact FFTW rounding (tolerance-based parity instead) and the
 * VISA low-pass filter (downsampling ratio > 1 in 2D; not restated).
 */
#include "paxoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_PI = 3.14159265358979323846264338327950288;

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ========================================================================== */
/* Discretisation: N = floor((max - min)/step + 1.0)  (src/paraxialSimulation.cpp:107-120) */
unsigned orc_nodes(double min, double max, double step)
{
  return (unsigned)((max - min) / step + 1.0);
}

/* ========================================================================== */
/* 4x4 transforms (src/shape.cpp:16-70).  translate and rotate right-multiply,
 * scale left-multiplies; a point maps to shape coordinates as p' = T p.      */
void orc_mat4_identity(double T[16])
{
  memset(T, 0, 16 * sizeof(double));
  T[0] = T[5] = T[10] = T[15] = 1.0;
}

static void mat4_mul(const double A[16], const double B[16], double C[16])
{
  double R[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[4 * i + k] * B[4 * k + j];
      R[4 * i + j] = s;
    }
  memcpy(C, R, sizeof(R));
}

void orc_shape_translate(double T[16], double x, double y, double z)
{
  double M[16];
  orc_mat4_identity(M);
  M[3] = -x; M[7] = -y; M[11] = -z;
  mat4_mul(T, M, T);
}

void orc_shape_rotate(double T[16], double angle_deg, int axis)
{
  double a = angle_deg * acos(-1.0) / 180.0;
  double M[16];
  orc_mat4_identity(M);
  double c = cos(a), s = sin(a);
  if (axis == 0) { M[5] = c; M[6] = s; M[9] = -s; M[10] = c; }
  else if (axis == 1) { M[0] = c; M[2] = -s; M[8] = s; M[10] = c; }
  else { M[0] = c; M[1] = s; M[4] = -s; M[5] = c; }
  mat4_mul(T, M, T);
}

void orc_shape_scale(double T[16], double factor, int axis)
{
  double M[16];
  orc_mat4_identity(M);
  M[5 * axis] = 1.0 / factor;
  mat4_mul(M, T, T);
}

void orc_shape_transform(const double T[16], double *x, double *y, double *z)
{
  double v[4] = {*x, *y, *z, 1.0}, o[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 4; ++j) s += T[4 * i + j] * v[j];
    o[i] = s;
  }
  *x = o[0]; *y = o[1]; *z = o[2];
}

/* Shape membership (src/shape.cpp:126-187): strict inequalities. */
int orc_shape_inside(const orc_shape *s, double x, double y, double z)
{
  orc_shape_transform(s->T, &x, &y, &z);
  switch (s->type) {
  case ORC_SHAPE_SPHERE:
    return sqrt(x * x + y * y + z * z) < s->p[0];
  case ORC_SHAPE_BOX:
    return (x > -s->p[0] / 2.0) && (x < s->p[0] / 2.0) && (y > -s->p[1] / 2.0) &&
           (y < s->p[1] / 2.0) && (z > -s->p[2] / 2.0) && (z < s->p[2] / 2.0);
  case ORC_SHAPE_CYLINDER: {
    double h = s->p[2];
    if (!((z < h / 2.0) && (z > -h / 2.0))) return 0;
    double slope = (s->p[1] - s->p[0]) / h;
    double radius = slope * (z + h / 2.0) + s->p[0];
    return sqrt(x * x + y * y) < radius;
  }
  }
  return 0;
}

/* Part: shapes scanned last -> first; the first containing shape decides
 * (UNION -> inside, DIFFERENCE -> outside)  (src/geometry.cpp:91-108). */
int orc_part_inside(const orc_csg *g, const orc_part *p, double x, double y, double z)
{
  for (int i = p->n_shapes - 1; i >= 0; --i) {
    const orc_shape *s = &g->shapes[p->first_shape + i];
    if (orc_shape_inside(s, x, y, z)) return s->op == ORC_OP_UNION;
  }
  return 0;
}

/* Module: parts scanned last -> first; UNION part -> its (delta,beta),
 * DIFFERENCE part -> 0; either way the module claims the point
 * (src/geometry.cpp:212-234).  CSGMaterial: first claiming module wins,
 * reference run -> 0 (src/materialFunction.cpp:3-13). */
void orc_csg_eval(const orc_csg *g, double x, double y, double z, double *delta, double *beta)
{
  *delta = 0.0;
  *beta = 0.0;
  if (g->is_reference_run) return;
  for (int m = 0; m < g->n_modules; ++m) {
    const orc_module *mod = &g->modules[m];
    for (int i = mod->n_parts - 1; i >= 0; --i) {
      const orc_part *p = &g->parts[mod->first_part + i];
      if (orc_part_inside(g, p, x, y, z)) {
        if (p->op == ORC_OP_UNION) { *delta = p->delta; *beta = p->beta; }
        else { *delta = 0.0; *beta = 0.0; }
        return;
      }
    }
  }
}

void orc_csg_matfn(const void *ctx, double x, double y, double z, double *d, double *b)
{
  orc_csg_eval((const orc_csg *)ctx, x, y, z, d, b);
}

/* ========================================================================== */
/* Curved waveguide: inside iff 0 < 2xR + z^2 < 2wR (src/curvedWaveGuide2D.cpp:49-52) */
int orc_curved_inside(const orc_curved_wg *w, double x, double z)
{
  return (2.0 * x * w->R + z * z > 0.0) && (2.0 * x * w->R + z * z < 2.0 * w->width * w->R);
}

static double curved_smoothed(const orc_curved_wg *w, double x, double z)
{
  /* src/curvedWaveGuide2D.cpp:179-194 */
  double ww = 0.1 * w->width;
  double d = 2.0 * w->R * x + z * z;
  double f1 = 1.0 / (1.0 + exp(-d / ww));
  double f2 = 1.0 / (1.0 + exp((2 * w->R * w->width - d) / ww));
  return 1.0 - (f1 - f2);
}

/* WaveGuideFDSimulation::getXrayMatProp (src/waveGuideFDSimulation.cpp:219-265)
 * and the smoothed override (src/curvedWaveGuide2D.cpp:196-205). */
void orc_curved_matfn(const void *ctx, double x, double y, double z, double *d, double *b)
{
  (void)y;
  const orc_curved_wg *w = (const orc_curved_wg *)ctx;
  if (w->use_smoothed) {
    double f = curved_smoothed(w, x, z);
    *d = w->clad_delta * f;
    *b = w->clad_beta * f;
    return;
  }
  if (z > w->wglength) { *b = 0.0; *d = 0.0; return; }
  int inside = orc_curved_inside(w, x, z);
  int nb = orc_curved_inside(w, x + w->dx, z) && orc_curved_inside(w, x - w->dx, z) &&
           orc_curved_inside(w, x, z + w->dz) && orc_curved_inside(w, x, z - w->dz);
  if (inside && nb) { *b = w->in_beta; *d = w->in_delta; return; }
  if (inside && !nb) {
    *d = 0.5 * (w->clad_delta + w->in_delta);
    *b = 0.5 * (w->clad_beta + w->in_beta);
    return;
  }
  *d = w->clad_delta;
  *b = w->clad_beta;
}

/* Examples/waveguide2D.cpp:15-27 with configurable bounds. */
void orc_slab_matfn(const void *ctx, double x, double y, double z, double *d, double *b)
{
  (void)y; (void)z;
  const orc_slab *s = (const orc_slab *)ctx;
  if ((x < s->x_hi) && (x > s->x_lo)) { *d = s->core_delta; *b = s->core_beta; }
  else { *d = s->clad_delta; *b = s->clad_beta; }
}

static inline void mat_eval(const orc_material *m, double x, double y, double z, double *d, double *b)
{
  if (m && m->fn) m->fn(m->ctx, x, y, z, d, b);
  else { *d = 0.0; *b = 0.0; }
}

/* ========================================================================== */
/* Sources (src/gaussianBeam.cpp:7-63, src/planeWave.cpp:5-14) */
static double gb_zr(const orc_gauss *g)
{
  double lambda = 2.0 * ORC_PI / g->k;
  return ORC_PI * g->waist * g->waist / lambda;
}
static double gb_spot(const orc_gauss *g, double z)
{
  double zr = gb_zr(g);
  return g->waist * sqrt(1.0 + pow(z / zr, 2));
}

orc_c orc_gauss_2d(const orc_gauss *g, double x, double z)
{
  x -= g->cx;
  double zz = z - g->z0;
  double zr = gb_zr(g);
  double wr = g->waist / gb_spot(g, zz);
  double gf = exp(-pow(x / gb_spot(g, zz), 2));
  double invR = zz / (zz * zz + zr * zr);
  double guoy = atan(zz / zr);
  double ph = -(g->k * g->z0);
  ph = ph + 0.5 * g->k * x * x * invR;
  ph = ph - guoy;
  orc_c phase = cexp(I * ph);
  double ang = g->angle_x_deg * ORC_PI / 180.0;
  orc_c angx = cexp(I * (g->k * sin(ang) * x));
  return g->amplitude * wr * gf * phase * angx;
}

orc_c orc_gauss_3d(const orc_gauss *g, double x, double y, double z)
{
  y -= g->cy;
  double gy = exp(-pow(y / gb_spot(g, z - g->z0), 2));
  double ang = g->angle_y_deg * ORC_PI / 180.0;
  orc_c angy = cexp(I * (g->k * sin(ang) * y));
  return orc_gauss_2d(g, x, z) * gy * angy;
}

orc_c orc_plane_wave(const orc_plane *p, double x, double z)
{
  (void)z;
  return p->amplitude * cexp(I * (p->k * x * sin(p->angle_rad)) - I * (p->k * p->z0));
}

void orc_gauss_2d_v(const orc_gauss *g, double x, double z, double out[2])
{
  orc_c v = orc_gauss_2d(g, x, z);
  out[0] = creal(v); out[1] = cimag(v);
}

void orc_plane_wave_v(const orc_plane *p, double x, double z, double out[2])
{
  orc_c v = orc_plane_wave(p, x, z);
  out[0] = creal(v); out[1] = cimag(v);
}

void orc_gauss_fill3d(const orc_gauss *g, double xmin, double dx, double ymin, double dy,
                      unsigned n, orc_c *out)
{
#pragma omp parallel for
  for (long idx = 0; idx < (long)n * n; ++idx) {
    unsigned i = (unsigned)(idx / n), j = (unsigned)(idx % n);
    double x = xmin + i * dx, y = ymin + j * dy;
    out[(size_t)i * n + j] = orc_gauss_3d(g, x, y, 0.0);
  }
}

/* ========================================================================== */
/* FFT with FFTW's conventions: X_k = sum_j x_j exp(sign*2*pi*i*j*k/N), no scaling.
 * Power-of-two sizes use an iterative radix-2 transform with FP64 twiddles from
 * cos/sin; other sizes use the O(N^2) definition. */
static int is_pow2(unsigned n) { return n && !(n & (n - 1)); }

static void fft_twiddles(orc_c *w, unsigned n, int sign)
{
  for (unsigned j = 0; j < n / 2; ++j) {
    double a = 2.0 * ORC_PI * (double)j / (double)n;
    w[j] = cos(a) + I * ((double)sign * sin(a));
  }
}

static void fft_pow2_inplace(orc_c *a, unsigned n, const orc_c *w)
{
  /* bit reversal */
  for (unsigned i = 1, j = 0; i < n; ++i) {
    unsigned bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { orc_c t = a[i]; a[i] = a[j]; a[j] = t; }
  }
  for (unsigned len = 2; len <= n; len <<= 1) {
    unsigned half = len >> 1, step = n / len;
    for (unsigned i = 0; i < n; i += len) {
      for (unsigned j = 0; j < half; ++j) {
        orc_c u = a[i + j];
        orc_c v = a[i + j + half] * w[j * step];
        a[i + j] = u + v;
        a[i + j + half] = u - v;
      }
    }
  }
}

/* The definition, term by term in order; the factor of term (j, k) is the
 * table entry of (j k mod n), cos + i sin of sign 2 pi (j k mod n)/n, i.e. the
 * same value as evaluating it in place. */
static void dft_table(orc_c *w, unsigned n, int sign)
{
  for (unsigned m = 0; m < n; ++m) {
    double ang = (double)sign * 2.0 * ORC_PI * (double)m / (double)n;
    w[m] = cos(ang) + I * sin(ang);
  }
}

static void dft_naive_w(orc_c *a, unsigned n, const orc_c *w)
{
  orc_c *o = (orc_c *)malloc(sizeof(orc_c) * n);
  for (unsigned k = 0; k < n; ++k) {
    orc_c s = 0.0;
    for (unsigned j = 0; j < n; ++j) s += a[j] * w[((unsigned long long)j * k) % n];
    o[k] = s;
  }
  memcpy(a, o, sizeof(orc_c) * n);
  free(o);
}

static void dft_naive(orc_c *a, unsigned n, int sign)
{
  orc_c *w = (orc_c *)malloc(sizeof(orc_c) * n);
  dft_table(w, n, sign);
  dft_naive_w(a, n, w);
  free(w);
}

void orc_fft1d(orc_c *a, unsigned n, int sign)
{
  if (n <= 1) return;
  if (!is_pow2(n)) { dft_naive(a, n, sign); return; }
  orc_c *w = (orc_c *)malloc(sizeof(orc_c) * (n / 2));
  fft_twiddles(w, n, sign);
  fft_pow2_inplace(a, n, w);
  free(w);
}

/* 2D transform of a column-major nrows x ncols matrix.  FFTW is handed the
 * column-major buffer as a row-major [nrows][ncols] array (src/fftSolver3D.cpp:59-62);
 * for the square grids the solver forces this is the same full 2D DFT. */
void orc_fft2d(orc_c *a, unsigned nrows, unsigned ncols, int sign)
{
  if (!is_pow2(nrows) || !is_pow2(ncols)) {
    /* lines are independent: OpenMP over them changes no result */
    orc_c *wr = (orc_c *)malloc(sizeof(orc_c) * nrows);
    orc_c *wc = (orc_c *)malloc(sizeof(orc_c) * ncols);
    dft_table(wr, nrows, sign);
    dft_table(wc, ncols, sign);
#pragma omp parallel for schedule(dynamic, 4)
    for (long c = 0; c < (long)ncols; ++c) dft_naive_w(a + (size_t)c * nrows, nrows, wr);
#pragma omp parallel
    {
      orc_c *line = (orc_c *)malloc(sizeof(orc_c) * ncols);
#pragma omp for schedule(dynamic, 4)
      for (long r = 0; r < (long)nrows; ++r) {
        for (unsigned c = 0; c < ncols; ++c) line[c] = a[(size_t)c * nrows + r];
        dft_naive_w(line, ncols, wc);
        for (unsigned c = 0; c < ncols; ++c) a[(size_t)c * nrows + r] = line[c];
      }
      free(line);
    }
    free(wr);
    free(wc);
    return;
  }
  orc_c *wr = (orc_c *)malloc(sizeof(orc_c) * (nrows / 2 + 1));
  orc_c *wc = (orc_c *)malloc(sizeof(orc_c) * (ncols / 2 + 1));
  orc_c *line = (orc_c *)malloc(sizeof(orc_c) * ncols);
  fft_twiddles(wr, nrows, sign);
  fft_twiddles(wc, ncols, sign);
  if (nrows > 1)
    for (unsigned c = 0; c < ncols; ++c) fft_pow2_inplace(a + (size_t)c * nrows, nrows, wr);
  if (ncols > 1)
    for (unsigned r = 0; r < nrows; ++r) {
      for (unsigned c = 0; c < ncols; ++c) line[c] = a[(size_t)c * nrows + r];
      fft_pow2_inplace(line, ncols, wc);
      for (unsigned c = 0; c < ncols; ++c) a[(size_t)c * nrows + r] = line[c];
    }
  free(wr); free(wc); free(line);
}

/* ========================================================================== */
/* Thomas algorithm without pivoting, symmetric off-diagonals (sup = sub),
 * solution returned in diag, rhs overwritten (include/thomasAlgorithm.tpp:2-32). */
void orc_thomas(orc_c *diag, const orc_c *sub, orc_c *rhs, unsigned n)
{
  if (n == 1) { diag[0] = rhs[0] / diag[0]; return; }
  orc_c *sup = (orc_c *)malloc(sizeof(orc_c) * (n - 1));
  for (unsigned i = 0; i < n - 1; ++i) sup[i] = sub[i];
  sup[0] /= diag[0];
  rhs[0] /= diag[0];
  for (unsigned i = 1; i < n; ++i) {
    if (i < n - 1) sup[i] /= (diag[i] - sub[i - 1] * sup[i - 1]);
    rhs[i] = (rhs[i] - sub[i - 1] * rhs[i - 1]) / (diag[i] - sub[i - 1] * sup[i - 1]);
  }
  diag[n - 1] = rhs[n - 1];
  for (int i = (int)n - 2; i >= 0; --i) diag[i] = rhs[i] - sup[i] * diag[i + 1];
  free(sup);
}

/* ========================================================================== */
/* Absorber::apply (src/absorber.cpp:4-12): both ends of a line, factor
 * exp(-(npix - i - 1) * inv) for i < npix (innermost cell factor 1). */
void orc_absorb_line(orc_c *f, unsigned len, unsigned stride, unsigned npix, double inv)
{
  for (unsigned i = 0; i < npix; ++i) {
    double fac = exp(-(double)(npix - i - 1) * inv);
    f[(size_t)i * stride] *= fac;
    f[(size_t)(len - i - 1) * stride] *= fac;
  }
}

/* ========================================================================== */
/* Solver3D::copyCurrentSolution block average (src/solver3D.cpp:78-125). */
void orc_cube_store(const orc_c *field, unsigned n, orc_c *cube, unsigned cube_n,
                    unsigned cube_slices, unsigned dsz, unsigned step)
{
  if (!cube || step % dsz != 0) return;
  unsigned s = step / dsz;
  if (s >= cube_slices) return;
  double dX = (double)n / (double)cube_n, dY = dX;
  orc_c *slice = cube + (size_t)s * cube_n * cube_n;
#pragma omp parallel for
  for (long i = 0; i < (long)cube_n * cube_n; ++i) {
    unsigned row = (unsigned)(i % cube_n), col = (unsigned)(i / cube_n);
    unsigned r0 = (unsigned)(row * dX), r1 = (unsigned)(row * dX + dX - 1);
    unsigned c0 = (unsigned)(col * dY), c1 = (unsigned)(col * dY + dY - 1);
    orc_c tot = 0.0;
    for (unsigned c = c0; c <= c1; ++c) {
      orc_c colsum = 0.0;
      for (unsigned r = r0; r <= r1; ++r) colsum += field[(size_t)c * n + r];
      tot += colsum;
    }
    slice[(size_t)col * cube_n + row] = tot / (dX * dY);
  }
}

/* ========================================================================== */
/* FFTSolver3D (src/fftSolver3D.cpp).  kernel: exp(-i dz/(2k) (kx^2+ky^2)) with
 * kx from the column index and dx, ky from the row index and dy, folded as
 * 2*pi*min(i, N-i)/(d*N) (:24-50). */
static double spatial_freq(unsigned idx, unsigned size, double step)
{
  if (idx > size / 2) idx = size - idx;
  return 2.0 * ORC_PI * idx / (step * size);
}

/* refractionIntegral (:253-273): trapezoid with n samples that skips i = n-1. */
static void refr_integral(const orc_material *m, unsigned n_refr, unsigned quirks, double x,
                          double y, double za, double zb, double *delta, double *beta)
{
  double dt = 0.0, bt = 0.0;
  mat_eval(m, x, y, za, delta, beta);
  mat_eval(m, x, y, zb, &dt, &bt);
  *delta += dt;
  *beta += bt;
  double dz = (zb - za) / n_refr;
  unsigned last = (quirks & ORC_FIX_REFR_SAMPLES) ? n_refr : n_refr - 1;
  for (unsigned i = 1; i < last; ++i) {
    double z = za + i * dz;
    mat_eval(m, x, y, z, &dt, &bt);
    *delta += 2.0 * dt;
    *beta += 2.0 * bt;
  }
  *delta /= (2.0 * n_refr);
  *beta /= (2.0 * n_refr);
}

void orc_fft3d_run(const orc_grid *g, const orc_material *m, const orc_absorber3d *ab,
                   unsigned n_refr, unsigned quirks, orc_c *field, unsigned n,
                   unsigned first_step, unsigned nsteps,
                   orc_c *cube, unsigned cube_n, unsigned cube_slices, unsigned dsz)
{
  size_t nn = (size_t)n * n;
  orc_c *prev = field;
  orc_c *cur = (orc_c *)malloc(sizeof(orc_c) * nn);
  const double dz = g->dz, k = g->k;
  const orc_c A = I * (0.5 * dz / k);
  double *kxs = (double *)malloc(sizeof(double) * n);
  double *kys = (double *)malloc(sizeof(double) * n);
  for (unsigned i = 0; i < n; ++i) {
    kxs[i] = spatial_freq(i, n, g->dx);
    kys[i] = spatial_freq(i, n, g->dy);
  }
  const double norm = (double)n * (double)n;
  for (unsigned st = first_step; st < first_step + nsteps; ++st) {
    /* propagate (:87-120): cur = FFT(prev); cur *= kernel; prev = IFFT(cur) */
    memcpy(cur, prev, sizeof(orc_c) * nn);
    orc_fft2d(cur, n, n, -1);
#pragma omp parallel for
    for (long i = 0; i < (long)nn; ++i) {
      unsigned row = (unsigned)(i % n), col = (unsigned)(i / n);
      double kx = kxs[col], ky = kys[row];
      cur[i] *= cexp(-A * (kx * kx + ky * ky));
    }
    memcpy(prev, cur, sizeof(orc_c) * nn);
    orc_fft2d(prev, n, n, +1);
    /* refraction (:122-156) */
    double z1 = g->zmin + st * g->dz;
    double z0 = g->zmin + (st - 1) * g->dz;
#pragma omp parallel for
    for (long i = 0; i < (long)nn; ++i) {
      unsigned row = (unsigned)(i % n), col = (unsigned)(i / n);
      double x = g->xmin + col * g->dx;
      double y = g->ymin + row * g->dy;
      double d1, b1, d0, b0;
      mat_eval(m, x, y, z1, &d1, &b1);
      mat_eval(m, x, y, z0, &d0, &b0);
      if ((fabs(d1 - d0) > 1e-10) || (fabs(b1 - b0) > 1e-10))
        refr_integral(m, n_refr, quirks, x, y, z0, z1, &d1, &b1);
      cur[i] = prev[i] * cexp(-k * (b1 + I * d1) * dz) / norm;
    }
    /* applyAbsorbingBC (:351-374): along each column (y) first, then rows (x) */
    if (ab && ab->npix_y)
      for (unsigned c = 0; c < n; ++c) orc_absorb_line(cur + (size_t)c * n, n, 1, ab->npix_y, ab->inv_y);
    if (ab && ab->npix_x)
      for (unsigned r = 0; r < n; ++r) orc_absorb_line(cur + r, n, n, ab->npix_x, ab->inv_x);
    /* copyCurrentSolution: prev = cur; cube slice */
    memcpy(prev, cur, sizeof(orc_c) * nn);
    orc_cube_store(prev, n, cube, cube_n, cube_slices, dsz, st);
  }
  free(cur); free(kxs); free(kys);
}

/* ========================================================================== */
/* FFTSolver2D (src/fftSolver2D.cpp:15-74): material sampled at z_step + dz/2. */
void orc_fft2d_run(const orc_grid *g, const orc_material *m, orc_c *field, unsigned nx,
                   unsigned first_step, unsigned nsteps, orc_c *solution, unsigned nz)
{
  orc_c *prev = field;
  orc_c *cur = (orc_c *)malloc(sizeof(orc_c) * nx);
  const orc_c A = I * (0.5 * g->dz / g->k);
  for (unsigned st = first_step; st < first_step + nsteps; ++st) {
    memcpy(cur, prev, sizeof(orc_c) * nx);
    orc_fft1d(cur, nx, -1);
    for (unsigned i = 0; i < nx; ++i) {
      double kx = spatial_freq(i, nx, g->dx);
      cur[i] *= cexp(-A * kx * kx);
    }
    memcpy(prev, cur, sizeof(orc_c) * nx);
    orc_fft1d(prev, nx, +1);
    double z = g->zmin + st * g->dz + 0.5 * g->dz;
    for (unsigned i = 0; i < nx; ++i) {
      double d, b;
      mat_eval(m, g->xmin + i * g->dx, 0.0, z, &d, &b);
      double norm = nx;
      cur[i] = prev[i] * cexp(-g->k * (b + I * d) * g->dz) / norm;
    }
    memcpy(prev, cur, sizeof(orc_c) * nx);
    if (solution && st < nz) memcpy(solution + (size_t)st * nx, cur, sizeof(orc_c) * nx);
  }
  free(cur);
}

/* ========================================================================== */
/* CrankNicholson::solveStep (src/crancNicholson.cpp:17-105) followed by applyTBC
 * (:107-148) and the Thomas solve.  Per row ix: diag = F + 0.25 (H+ + H-) G i rho
 * + 0.5 (beta r + i delta r) with delta -= J; sub[ix] (the coupling of ix and ix+1,
 * sup = sub) = -0.25 i rho H- G of row ix; rhs = 0.25 i rho (u_{ix-1} H-' G' +
 * u_{ix+1} H+' G') - 0.25 u i rho (H+' + H-') G + (F' - 0.5 (beta' r + i delta' r)) u,
 * primes at z - dz; note G (not G') in the centre term (:87).  The transparent
 * boundary uses H-, G, H-', G' of the last row for both sides (:146-147). */
void orc_poly_eqfn(const void *ctx, double x, double z, double out[4])
{
  const orc_poly_eq *e = (const orc_poly_eq *)ctx;
  for (int q = 0; q < 4; ++q)
    out[q] = e->c[q][0] + e->c[q][1] * x + e->c[q][2] * z + (e->c[q][3] != 0.0 ? e->c[q][3] / x : 0.0);
}

void orc_cn_step_eq(double rho, double r, int bc, const orc_c *prev, orc_c *cur, unsigned nx,
                    const double *d1, const double *b1, const double *d0, const double *b0,
                    const double *e1, const double *e0)
{
  orc_c *sub = (orc_c *)calloc(nx > 1 ? nx - 1 : 1, sizeof(orc_c));
  orc_c *rhs = (orc_c *)malloc(sizeof(orc_c) * nx);
  orc_c *diag = (orc_c *)malloc(sizeof(orc_c) * nx);
  double Hp = 1.0, Hm = 1.0, gv = 1.0, HpP = 1.0, HmP = 1.0, gP = 1.0;
  for (unsigned ix = 0; ix < nx; ++ix) {
    double fv = 1.0, jv = 0.0, fP = 1.0, jP = 0.0;
    if (e1) {
      fv = e1[5 * ix]; gv = e1[5 * ix + 1]; Hp = e1[5 * ix + 2]; Hm = e1[5 * ix + 3]; jv = e1[5 * ix + 4];
    }
    if (e0) {
      fP = e0[5 * ix]; gP = e0[5 * ix + 1]; HpP = e0[5 * ix + 2]; HmP = e0[5 * ix + 3]; jP = e0[5 * ix + 4];
    }
    double delta = d1[ix] - jv, beta = b1[ix];
    double deltaP = d0[ix] - jP, betaP = b0[ix];
    diag[ix] = fv * 1.0 + 0.25 * (Hp + Hm) * gv * I * rho + 0.5 * (beta * r + I * delta * r);
    if (ix < nx - 1) sub[ix] = -0.25 * I * rho * Hm * gv;
    orc_c left = ix > 0 ? prev[ix - 1] : 0.0;
    orc_c center = prev[ix];
    orc_c right = ix < nx - 1 ? prev[ix + 1] : 0.0;
    orc_c v = left * HmP * gP;
    v += right * HpP * gP;
    v *= (0.25 * I * rho);
    v -= 0.25 * center * I * rho * (HpP + HmP) * gv;
    v += (1.0 * fP - 0.5 * (betaP * r + I * deltaP * r)) * center;
    rhs[ix] = v;
  }
  if (bc == ORC_BC_TRANSPARENT && nx >= 2) {
    const double ZERO = 1e-16;
    unsigned sides[2][2] = {{0, 1}, {nx - 1, nx - 2}};
    for (int s = 0; s < 2; ++s) {
      unsigned outer = sides[s][0], inner = sides[s][1];
      if (cabs(prev[inner]) > ZERO) {
        orc_c ratio = prev[outer] / prev[inner];
        orc_c kdx = clog(ratio) / I;
        if (creal(kdx) < 0.0) kdx = 0.0 + I * cimag(kdx);
        diag[outer] -= 0.25 * I * rho * Hm * gv * cexp(I * kdx);
        rhs[outer] += 0.25 * I * rho * HmP * gP * cexp(I * kdx) * prev[outer];
      }
    }
  }
  orc_thomas(diag, sub, rhs, nx);
  for (unsigned ix = 0; ix < nx; ++ix) cur[ix] = diag[ix];
  free(sub); free(rhs); free(diag);
}

void orc_cn_step_planes(double rho, double r, int bc, const orc_c *prev, orc_c *cur, unsigned nx,
                        const double *d1, const double *b1, const double *d0, const double *b0)
{
  orc_cn_step_eq(rho, r, bc, prev, cur, nx, d1, b1, d0, b0, NULL, NULL);
}

/* The equation values of one plane at z, 5 per row (orc_cn_step_eq). */
static void eq_plane(const orc_equation *eq, const orc_grid *g, unsigned nx, double z, double *e)
{
  for (unsigned ix = 0; ix < nx; ++ix) {
    double x = g->xmin + ix * g->dx, v[4], vp[4], vm[4];
    eq->fn(eq->ctx, x, z, v);
    eq->fn(eq->ctx, x + 0.5 * g->dx, z, vp);
    eq->fn(eq->ctx, x - 0.5 * g->dx, z, vm);
    e[5 * ix] = v[0];
    e[5 * ix + 1] = v[1];
    e[5 * ix + 2] = vp[2];
    e[5 * ix + 3] = vm[2];
    e[5 * ix + 4] = v[3];
  }
}

void orc_cn_run_eq(const orc_grid *g, const orc_material *m, const orc_equation *eq, int bc,
                   orc_c *field, unsigned nx, unsigned first_step, unsigned nsteps,
                   orc_c *solution, unsigned nz, unsigned absorb_npix, double absorb_inv)
{
  orc_c *cur = (orc_c *)malloc(sizeof(orc_c) * nx);
  double *d1 = (double *)malloc(sizeof(double) * nx * 4);
  double *b1 = d1 + nx, *d0 = d1 + 2 * nx, *b0 = d1 + 3 * nx;
  const int custom = eq && eq->fn;
  double *e1 = custom ? (double *)malloc(sizeof(double) * nx * 10) : NULL;
  double *e0 = custom ? e1 + 5 * nx : NULL;
  const double rho = g->dz / (g->k * g->dx * g->dx);
  const double r = g->k * g->dz;
  for (unsigned st = first_step; st < first_step + nsteps; ++st) {
    double z = g->zmin + (double)st * g->dz;
    for (unsigned ix = 0; ix < nx; ++ix) {
      double x = g->xmin + ix * g->dx;
      mat_eval(m, x, 0.0, z, &d1[ix], &b1[ix]);
      mat_eval(m, x, 0.0, z - g->dz, &d0[ix], &b0[ix]);
    }
    if (custom) {
      eq_plane(eq, g, nx, z, e1);
      eq_plane(eq, g, nx, z - g->dz, e0);
    }
    orc_cn_step_eq(rho, r, bc, field, cur, nx, d1, b1, d0, b0, e1, e0);
    if (absorb_npix) orc_absorb_line(cur, nx, 1, absorb_npix, absorb_inv);
    memcpy(field, cur, sizeof(orc_c) * nx);
    if (solution && st < nz) memcpy(solution + (size_t)st * nx, cur, sizeof(orc_c) * nx);
  }
  free(cur); free(d1); free(e1);
}

void orc_cn_run(const orc_grid *g, const orc_material *m, int bc, orc_c *field, unsigned nx,
                unsigned first_step, unsigned nsteps, orc_c *solution, unsigned nz,
                unsigned absorb_npix, double absorb_inv)
{
  orc_cn_run_eq(g, m, NULL, bc, field, nx, first_step, nsteps, solution, nz, absorb_npix, absorb_inv);
}

/* ========================================================================== */
/* ADI (src/alternatingDirectionSolver.cpp).  Per-cell material lookups (the
 * reference shares delta/beta across OpenMP threads at :22,29,36; the race-free
 * per-cell semantics are restated). */
static void adi_tbc(const orc_c *prev, orc_c *diag, orc_c *rhs, unsigned Nx, unsigned Ny,
                    double k, double dx, double dy, int implicit_x)
{
  /* :157-234 */
  const double zero = 1e-6;
  if (implicit_x) {
    for (unsigned i = 0; i < Ny; ++i) {
      orc_c kd = -I * clog(prev[(size_t)(Nx - 1) * Ny + i] / prev[(size_t)(Nx - 2) * Ny + i]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      size_t idx = (size_t)i * Nx + Nx - 1;
      if (cabs(prev[(size_t)(Nx - 2) * Ny + i]) > zero) diag[idx] -= 0.5 * I * cexp(I * kd) / (k * dx * dx);
      idx = (size_t)i * Nx;
      kd = -I * clog(prev[i] / prev[(size_t)Ny + i]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      if (cabs(prev[(size_t)Ny + i]) > zero) diag[idx] -= 0.5 * I * cexp(I * kd) / (k * dx * dx);
    }
    for (unsigned i = 0; i < Nx; ++i) {
      orc_c kd = -I * clog(prev[(size_t)i * Ny + Ny - 1] / prev[(size_t)i * Ny + Ny - 2]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      size_t idx = (size_t)(Ny - 1) * Nx + i;
      if (cabs(prev[(size_t)i * Ny + Ny - 2]) > zero)
        rhs[idx] += 0.5 * I * cexp(I * kd) * prev[(size_t)i * Ny + Ny - 1] / (k * dy * dy);
      idx = i;
      kd = -I * clog(prev[(size_t)i * Ny] / prev[(size_t)i * Ny + 1]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      if (cabs(prev[(size_t)i * Ny + 1]) > zero)
        rhs[idx] += 0.5 * I * cexp(I * kd) * prev[(size_t)i * Ny] / (k * dy * dy);
    }
  } else {
    for (unsigned i = 0; i < Nx; ++i) {
      orc_c kd = -I * clog(prev[(size_t)i * Ny + Ny - 1] / prev[(size_t)i * Ny + Ny - 2]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      size_t idx = (size_t)i * Ny + Ny - 1;
      if (cabs(prev[(size_t)i * Ny + Ny - 2]) > zero) diag[idx] -= 0.5 * I * cexp(I * kd) / (k * dy * dy);
      idx = (size_t)i * Ny;
      kd = -I * clog(prev[(size_t)i * Ny] / prev[(size_t)i * Ny + 1]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      if (cabs(prev[(size_t)i * Ny + 1]) > zero) diag[idx] -= 0.5 * I * cexp(I * kd) / (k * dy * dy);
    }
    for (unsigned i = 0; i < Ny; ++i) {
      orc_c kd = -I * clog(prev[(size_t)(Nx - 1) * Ny + i] / prev[(size_t)(Nx - 2) * Ny + i]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      size_t idx = (size_t)(Nx - 1) * Ny + i;
      if (cabs(prev[(size_t)(Nx - 2) * Ny + i]) > zero)
        rhs[idx] += 0.5 * I * cexp(I * kd) * prev[(size_t)(Nx - 1) * Ny + i] / (k * dx * dx);
      idx = i;
      kd = -I * clog(prev[i] / prev[(size_t)Ny + i]);
      if (creal(kd) < 0.0) kd = 0.0 + I * cimag(kd);
      if (cabs(prev[(size_t)Ny + i]) > zero)
        rhs[idx] += 0.5 * I * cexp(I * kd) * prev[i] / (k * dx * dx);
    }
  }
}

static void adi_half(const orc_grid *g, const orc_material *m, int use_tbc, unsigned quirks,
                     const orc_c *prev, orc_c *cur, unsigned n, unsigned step, int implicit_x,
                     orc_c *diag, orc_c *sub, orc_c *rhs)
{
  const unsigned Nx = n, Ny = n;
  const double dx = g->dx, dy = g->dy, k = g->k;
  const double dz = g->dz / 2.0;
  size_t N = (size_t)Nx * Ny;
  memset(sub, 0, sizeof(orc_c) * (N - 1));
  if (implicit_x) {
    /* xImplicit (:8-76): system index indx = y*Nx + x */
    double z = g->zmin + step * g->dz;
#pragma omp parallel for
    for (long indx = 0; indx < (long)N; ++indx) {
      unsigned j = (unsigned)(indx % Nx), i = (unsigned)(indx / Nx);
      double x = g->xmin + j * dx, y = g->ymin + i * dy, delta, beta;
      mat_eval(m, x, y, z, &delta, &beta);
      diag[indx] = 1.0 / dz + I / (k * dx * dx) + beta * k + I * delta * k;
      orc_c v = (1.0 / dz - I / (k * dy * dy)) * prev[(size_t)j * Ny + i];
      if (j > 0) sub[indx - 1] = -0.5 * I / (k * dx * dx);
      if (i > 0) v += 0.5 * I * prev[(size_t)j * Ny + i - 1] / (k * dy * dy);
      if (i < Ny - 1) v += 0.5 * I * prev[(size_t)j * Ny + i + 1] / (k * dy * dy);
      rhs[indx] = v;
    }
  } else {
    /* yImplicit (:78-134): system index indx = x*Ny + y */
    double z = g->zmin + step * g->dz + dz;
    int fix = (quirks & ORC_FIX_ADI_Y_DIAG) != 0;
#pragma omp parallel for
    for (long indx = 0; indx < (long)N; ++indx) {
      unsigned i = (unsigned)(indx / Ny), j = (unsigned)(indx % Ny);
      double x = g->xmin + i * dx, y = g->ymin + j * dy, delta, beta;
      mat_eval(m, x, y, z, &delta, &beta);
      if (fix) diag[indx] = 1.0 / dz + I / (k * dy * dy) + beta * k + I * delta * k;
      else diag[indx] = 1.0 / dz + I / (k * dy * dy + beta * k + I * delta * k);
      orc_c v = (-I / (k * dx * dx) + 1.0 / dz) * prev[(size_t)i * Ny + j];
      if (j > 0) sub[indx - 1] = -0.5 * I / (k * dy * dy);
      if (i > 0) v += 0.5 * I * prev[(size_t)(i - 1) * Ny + j] / (k * dx * dx);
      if (i < Ny - 1) v += 0.5 * I * prev[(size_t)(i + 1) * Ny + j] / (k * dx * dx);
      rhs[indx] = v;
    }
  }
  if (use_tbc) adi_tbc(prev, diag, rhs, Nx, Ny, k, dx, dy, implicit_x);
  orc_thomas(diag, sub, rhs, (unsigned)N);
#pragma omp parallel for
  for (long indx = 0; indx < (long)N; ++indx) {
    if (implicit_x) {
      unsigned i = (unsigned)(indx / Nx), j = (unsigned)(indx % Nx);
      cur[(size_t)j * Ny + i] = diag[indx];
    } else {
      cur[indx] = diag[indx];
    }
  }
}

void orc_adi_run(const orc_grid *g, const orc_material *m, int use_tbc, unsigned quirks,
                 orc_c *field, unsigned n, unsigned first_step, unsigned nsteps,
                 orc_c *cube, unsigned cube_n, unsigned cube_slices, unsigned dsz)
{
  size_t N = (size_t)n * n;
  orc_c *cur = (orc_c *)malloc(sizeof(orc_c) * N);
  orc_c *diag = (orc_c *)malloc(sizeof(orc_c) * N);
  orc_c *sub = (orc_c *)malloc(sizeof(orc_c) * N);
  orc_c *rhs = (orc_c *)malloc(sizeof(orc_c) * N);
  for (unsigned st = first_step; st < first_step + nsteps; ++st) {
    adi_half(g, m, use_tbc, quirks, field, cur, n, st, 1, diag, sub, rhs);
    memcpy(field, cur, sizeof(orc_c) * N);
    adi_half(g, m, use_tbc, quirks, field, cur, n, st, 0, diag, sub, rhs);
    memcpy(field, cur, sizeof(orc_c) * N);
    orc_cube_store(field, n, cube, cube_n, cube_slices, dsz, st);
  }
  free(cur); free(diag); free(sub); free(rhs);
}

/* ========================================================================== */
double orc_rel_l2(const orc_c *a, const orc_c *ref, uint64_t n)
{
  double num = 0.0, den = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    orc_c d = a[i] - ref[i];
    num += creal(d) * creal(d) + cimag(d) * cimag(d);
    den += creal(ref[i]) * creal(ref[i]) + cimag(ref[i]) * cimag(ref[i]);
  }
  return den > 0 ? sqrt(num / den) : sqrt(num);
}

double orc_rel_l2_intensity(const orc_c *a, const orc_c *ref, uint64_t n)
{
  double num = 0.0, den = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    double ia = creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
    double ir = creal(ref[i]) * creal(ref[i]) + cimag(ref[i]) * cimag(ref[i]);
    num += (ia - ir) * (ia - ir);
    den += ir * ir;
  }
  return den > 0 ? sqrt(num / den) : sqrt(num);
}
