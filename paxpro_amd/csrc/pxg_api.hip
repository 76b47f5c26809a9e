// C-ABI layer of libpaxgpu: contexts, plans, device memory, material upload and
// the z-step drivers that enqueue the kernels of pxg_fft3d.hip, pxg_cn.hip and
// pxg_adi.hip.  No exception crosses the boundary: every entry point catches,
// records a thread-local message and returns a PXG_ERR_* code.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/paxgpu.h"
#include "pxg_internal.hpp"
#include "pxg_tridiag.hpp"

using namespace pxg;

// Scratch of pxg_thomas_batched_noalloc: the eliminated super-diagonal of every
// row of one call (16 bytes per row), PXG_THOMAS_NOALLOC_MAX_ROWS rows, 16 MiB
// per device for the life of the process.  Host code never uses this name as a
// device address: pxg_ctx_create looks up each device's copy once
// (hipGetSymbolAddress with that device current) into g_scratch_dev, and only
// those addresses reach kernels.
static __device__ __align__(256) char scratch[static_cast<size_t>(PXG_THOMAS_NOALLOC_MAX_ROWS) * 16];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_dev;  // per device: address of its copy of scratch (null: not looked up)

thread_local std::string g_err;

struct Failure {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Failure{code, buf};
}

void hip_check(hipError_t e, const char *what)
{
  if (e != hipSuccess) fail(PXG_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

template <class F>
int guarded(F &&f)
{
  try {
    f();
    g_err.clear();
    return PXG_OK;
  } catch (const Failure &e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return PXG_ERR_BUDGET;
  } catch (const std::exception &e) {
    g_err = e.what();
    return PXG_ERR_INVALID;
  }
}

const double PI = 3.14159265358979323846264338327950288;

// Device buffer owned by a plan.
// Re-allocation to a size within the current capacity reuses the buffer, so a
// plan that is reconfigured (pxg_plan_reset) does not call hipMalloc/hipFree,
// which would synchronise the device while other plans run.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  size_t cap = 0;
  void alloc(size_t n)
  {
    if (n <= cap && p) {
      bytes = n;
      return;
    }
    release();
    if (n == 0) return;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
    cap = n;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    cap = 0;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
  ~DBuf() { release(); }
};

}  // namespace

struct pxg_ctx_s {
  int device = 0;
  size_t budget = 0;
  std::atomic<size_t> reserved{0};
  std::mutex mu;
  // live plans; a destroyed context is freed when its last plan is destroyed
  int plans = 0;
  bool closed = false;
  // autotuned FFT3D shapes: (nx, batch, precision, material) -> (lines_x, lines_y,
  // batch marched in one launch per pass instead of two halves)
  std::map<std::array<int, 4>, std::array<int, 3>> tuned;
  // autotuned ADI tile shapes: (nx, batch, precision, material) -> (x tile, y tile)
  std::map<std::array<int, 4>, std::pair<int, int>> tuned_adi;
};

struct pxg_plan_s {
  pxg_ctx ctx = nullptr;
  pxg_problem p{};
  hipStream_t stream = nullptr;
  bool mixed = false;
  int esz = 16;           // bytes per complex field element
  size_t cells = 0;       // per field
  size_t reserved = 0;
  int lines_x = 1, lines_y = 1;
  bool one_launch = false;  // a batch that could march in halves runs one launch per pass (autotuned)
  DBuf field, tw, kx, ky, absx, absy, kvec;
  // FFT kinds with nx not a power of two in [16, 4096]: Bluestein length and tables
  int blue_m = 0;
  DBuf chirp, blue_bf, blue_bi;
  // CN with a ParaxialEquation other than the default (pxg_set_equation_planes)
  DBuf eq_planes;
  int eq_n = 0, eq_first = 0, eq_inv = 0;
  // far field: twiddle tables of the last signal length and transform length used
  DBuf ff_twm, ff_twn;
  unsigned ff_twm_len = 0, ff_twn_len = 0;
  DBuf mcode, mxdb;  // FFT3D: material map of the current z-step (CSG materials)
  long map_step = -1;  // step whose map mcode holds (-1: none)
  // FFT3D batches marched as two halves on two streams (split_fft3d): the
  // second half's map and the step it holds
  DBuf mcode2, mxdb2;
  long map_step2 = -1;
  std::vector<double> k_host;
  // material
  int mat_mode = MAT_NONE;
  DBuf shapes, parts, boxes, part_fac, planes;
  int adi_tile[2] = {0, 0};  // ADI: tile shapes of the x- and y-implicit half-steps
  int n_parts = 0, is_ref = 0;
  std::vector<DPart> parts_host;
  int n_planes = 0, first_plane = 0, z_invariant = 0;
  DCurved wg{};
  // solution storage
  DBuf sol;
  int sol_dims[3] = {0, 0, 0};
  // FFT3D z-step loops captured as HIP graphs (pxg_plan_set_graphs): replayed when
  // the call's shape and every kernel argument match the capture
  bool use_graphs = false;
  struct Graph {
    std::vector<char> key;
    long map_step_out;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  size_t sol_fstride = 0;  // elements per field
  // tridiagonal scratch (CN, ADI)
  DBuf scratch;
  bool field_set = false;
  // per-field geometry: job plans (pxg_plan_create_jobs) and CN job tables
  bool is_jobs = false;
  std::vector<pxg_problem> jp;      // one problem per job (job plans only)
  std::vector<size_t> foff, soff;   // element offsets of each field / stored solution
  std::vector<size_t> fcells;       // cells of each field
  std::vector<int> job_mat;         // per-job material mode (job plans)
  std::vector<DCurved> job_wg;
  DBuf jobs_dev;      // CN jobs marched by k_cn_jobs (chunks above cn_reg_rows() rows)
  DBuf jobs_reg_dev;  // CN jobs marched by k_cn_reg
  DBuf jobs_big_dev;  // CN jobs marched by k_cn_big
  int njobs_glob = 0, njobs_reg = 0, njobs_big = 0, max_P_reg = 1, max_P_big = 1;
  bool reg_fixed = false;  // register-class jobs march with the factored matrix (k_cn_fixed)
  // extra streams: the CN job sets march concurrently
  hipStream_t aux = nullptr, aux2 = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, join2_ev = nullptr;
  bool jobs_dirty = true;
  int max_P = 1;
  size_t sol_total = 0;             // elements of solution storage
  // profiling
  std::vector<hipEvent_t> ev;
  // pinned staging for field transfers: two chunks, DMA of one overlaps the
  // host-side copy / precision conversion of the other
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // group stepping (pxg_step_group), batch-1 FFT3D plans: the event after this
  // plan's earlier work, and the event after a group's launches that it led
  hipEvent_t ready_ev = nullptr, group_done_ev = nullptr;
  struct Rec {
    const char *name;
    int e0, e1;
    double bytes;
  };
  std::vector<Rec> recs;
  bool profiling = false;
};

namespace {

// -------------------------------------------------------------- bookkeeping
// complex64 field storage: MIXED plans of the HBM-bound kinds.  CN plans keep
// complex128 state in both precisions: the march is latency-bound (complex64
// saves no time), and rounding a slowly varying 1D field to complex64 every
// step was measured to accumulate coherently (5.7e-5 relative L2 after 1000
// steps on curved guides).
bool field_c64(const pxg_problem &p) { return p.precision == PXG_PREC_MIXED && p.kind != PXG_CN; }

size_t table_bytes(const pxg_problem &p, int esz)
{
  const size_t n = p.nx;
  size_t b = n * esz * 3 + 2 * n * 8 + p.batch * 8 + 2 * n * p.batch * esz;
  return b + 4096;
}

// FFT3D batches of 16 fields or more march as two halves on two streams
// (run_fft3d_halves): power-of-two lines, even batch, no solution storage.
bool fft3d_halves(const pxg_problem &p)
{
  return p.kind == PXG_FFT3D && fft_size_supported(p.nx) && p.batch >= 16 && p.batch % 2 == 0 &&
         p.store_dsz <= 0;
}

// FFT3D plans whose map has launches of its own (fewer fields than pass X
// builds it for) keep MAP_RING maps: consecutive steps' maps are built in one
// launch (run_fft3d).  Slot 0 is the single map of the other paths.
constexpr int MAP_IN_PASS_X_BATCH = 8;  // fields from which pass X builds the map
int map_slots(const pxg_problem &p)
{
  return p.kind == PXG_FFT3D && fft_size_supported(p.nx) && p.batch < MAP_IN_PASS_X_BATCH ? MAP_RING : 1;
}

size_t plan_bytes(const pxg_problem &p)
{
  const int esz = field_c64(p) ? 8 : 16;
  const size_t n = p.nx;
  const bool is3d = p.kind == PXG_FFT3D || p.kind == PXG_ADI;
  const size_t cells = is3d ? n * n : n;
  size_t b = cells * p.batch * esz;
  b += table_bytes(p, esz);
  if (p.store_dsz > 0) {
    if (is3d) {
      const size_t cn = n / std::max(1, p.store_dsx);
      const size_t slices = p.nz / p.store_dsz + 1;
      b += cn * cn * slices * 16 * p.batch;
    } else {
      b += cells * p.nz * esz * p.batch;
    }
  }
  if (p.kind == PXG_FFT3D)  // CSG material maps: one per half (batch halves), MAP_RING (maps ahead), else one
    b += cells * (sizeof(uint16_t) + sizeof(cd)) * (fft3d_halves(p) ? 2 : map_slots(p));
  if (p.kind == PXG_ADI) b += cells * p.batch * 16 * 4;   // coefficient planes
  if (p.kind == PXG_CN) b += cells * p.batch * 16 * 6;
  return b;
}

void validate(const pxg_problem &p)
{
  if (p.kind < PXG_FFT2D || p.kind > PXG_ADI) fail(PXG_ERR_INVALID, "unknown solver kind %d", p.kind);
  if (p.precision != PXG_PREC_FP64 && p.precision != PXG_PREC_MIXED)
    fail(PXG_ERR_INVALID, "unknown precision %d", p.precision);
  if (p.nx < 2) fail(PXG_ERR_INVALID, "nx must be >= 2 (got %d)", p.nx);
  if (p.nz < 1) fail(PXG_ERR_INVALID, "nz must be >= 1");
  if (p.batch < 1) fail(PXG_ERR_INVALID, "batch must be >= 1");
  if (!(p.dx > 0) || !(p.dz > 0) || !(p.k > 0)) fail(PXG_ERR_INVALID, "dx, dz and k must be > 0");
  if ((p.kind == PXG_FFT3D || p.kind == PXG_FFT2D) && p.nx > 4096)
    fail(PXG_ERR_UNSUPPORTED, "FFT solvers take nx up to 4096 (got %d)", p.nx);
  if ((p.kind == PXG_FFT3D || p.kind == PXG_ADI) && !(p.dy > 0)) fail(PXG_ERR_INVALID, "dy must be > 0");
  if (p.store_dsz < 0 || p.store_dsx < 0) fail(PXG_ERR_INVALID, "negative storage ratio");
  if (p.store_dsz > 0 && p.store_dsx < 1) fail(PXG_ERR_INVALID, "store_dsx must be >= 1");
  if ((p.kind == PXG_FFT2D || p.kind == PXG_CN) && p.store_dsz > 0 && p.store_dsx != 1)
    fail(PXG_ERR_UNSUPPORTED,
         "2D transverse downsampling needs the VISA low-pass filter, which is not implemented");
  if ((p.kind == PXG_FFT3D || p.kind == PXG_ADI) && p.store_dsz > 0 && p.nx / p.store_dsx < 1)
    fail(PXG_ERR_INVALID, "storage downsampling larger than the grid");
}

template <class R>
void upload_cx(DBuf &dst, const std::vector<cx<double>> &src, hipStream_t s)
{
  std::vector<cx<R>> tmp(src.size());
  for (size_t i = 0; i < src.size(); ++i) tmp[i] = cvt<R>(src[i]);
  dst.alloc(tmp.size() * sizeof(cx<R>));
  hip_check(hipMemcpyAsync(dst.p, tmp.data(), dst.bytes, hipMemcpyHostToDevice, s), "upload");
  hip_check(hipStreamSynchronize(s), "upload sync");
}

template <class R>
void upload_real(DBuf &dst, const std::vector<double> &src, hipStream_t s)
{
  std::vector<R> tmp(src.begin(), src.end());
  dst.alloc(tmp.size() * sizeof(R));
  hip_check(hipMemcpyAsync(dst.p, tmp.data(), dst.bytes, hipMemcpyHostToDevice, s), "upload");
  hip_check(hipStreamSynchronize(s), "upload sync");
}

// Absorber::apply along one line of length n (src/absorber.cpp:4-12) applied to
// a line of ones: the per-position factor table.
std::vector<double> absorber_table(unsigned n, unsigned npix, double inv)
{
  std::vector<double> f(n, 1.0);
  for (unsigned i = 0; i < npix && i < n; ++i) {
    const double fac = std::exp(-static_cast<double>(npix - i - 1) * inv);
    f[i] *= fac;
    f[n - i - 1] *= fac;
  }
  return f;
}

// FFTSolver3D::spatialFreqX/Y (src/fftSolver3D.cpp:31-50).
double spatial_freq(unsigned idx, unsigned size, double step)
{
  if (idx > size / 2) idx = size - idx;
  return 2.0 * PI * idx / (step * size);
}

// exp(-i dz/(2k) q^2), the half-kernel of FFTSolver3D::kernel (:24-29).
cx<double> kernel_half(double q, double dz, double k)
{
  const double ph = -(0.5 * dz / k) * (q * q);
  return {std::cos(ph), std::sin(ph)};
}

// exp(-k (beta + i delta) dz) * inv_norm in FP64.
cx<double> screen_factor(double k, double dz, double delta, double beta, double inv_norm)
{
  const double amp = std::exp((-k * beta) * dz);
  const double ph = (-k * delta) * dz;
  return {std::cos(ph) * amp * inv_norm, std::sin(ph) * amp * inv_norm};
}

// In-place FP64 radix-2 DFT of a power-of-two length, X_k = sum_j x_j e^{sign 2 pi i jk/n},
// with directly evaluated twiddles: the Bluestein kernel tables.
void host_fft(std::vector<cx<double>> &a, int sign)
{
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t k = 0; k < len / 2; ++k) {
      const double ang = sign * 2.0 * PI * (static_cast<double>(k) / len);
      const cx<double> w{std::cos(ang), std::sin(ang)};
      for (size_t i = 0; i < n; i += len) {
        const cx<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

// Bluestein tables of an n-point DFT on M = blue_length(n) points: the chirp
// c_j = e^{-i pi j^2/n} (j^2 reduced mod 2n, so the angle is exact before its
// one rounding) and the transformed convolution kernels FFT_M(conj c)/M
// (forward) and FFT_M(c)/M (inverse), each kernel placed cyclically (entries
// m and M - m for |m| < n).
void build_bluestein(pxg_plan P, unsigned n, hipStream_t s)
{
  const unsigned M = static_cast<unsigned>(blue_length(static_cast<int>(n)));
  P->blue_m = static_cast<int>(M);
  std::vector<cx<double>> tw(M), chirp(n), kf(M, cx<double>{0.0, 0.0}), ki(M, cx<double>{0.0, 0.0});
  for (unsigned m = 0; m < M; ++m) {
    const double a = 2.0 * PI * (static_cast<double>(m) / M);
    tw[m] = {std::cos(a), -std::sin(a)};
  }
  for (unsigned j = 0; j < n; ++j) {
    const uint64_t r = (static_cast<uint64_t>(j) * j) % (2ull * n);
    const double a = PI * (static_cast<double>(r) / n);
    chirp[j] = {std::cos(a), -std::sin(a)};
  }
  for (unsigned m = 0; m < n; ++m) {
    const cx<double> c = chirp[m];
    kf[m] = conj(c);
    ki[m] = c;
    if (m) {
      kf[M - m] = conj(c);
      ki[M - m] = c;
    }
  }
  host_fft(kf, -1);
  host_fft(ki, -1);
  for (unsigned m = 0; m < M; ++m) {
    kf[m] = {kf[m].x / M, kf[m].y / M};
    ki[m] = {ki[m].x / M, ki[m].y / M};
  }
  upload_cx<double>(P->tw, tw, s);
  upload_cx<double>(P->chirp, chirp, s);
  upload_cx<double>(P->blue_bf, kf, s);
  upload_cx<double>(P->blue_bi, ki, s);
}

void build_tables(pxg_plan P)
{
  const pxg_problem &p = P->p;
  const unsigned n = p.nx;
  hipStream_t s = P->stream;
  if (p.kind == PXG_FFT2D || p.kind == PXG_FFT3D) {
    if (fft_size_supported(static_cast<int>(n))) {
      P->blue_m = 0;
      std::vector<cx<double>> tw(n);
      for (unsigned m = 0; m < n; ++m) {
        const double a = 2.0 * PI * (static_cast<double>(m) / n);
        tw[m] = {std::cos(a), -std::sin(a)};
      }
      upload_cx<double>(P->tw, tw, s);
    } else {
      build_bluestein(P, n, s);
    }
    std::vector<cx<double>> kx(static_cast<size_t>(n) * p.batch), ky(kx.size());
    for (int b = 0; b < p.batch; ++b)
      for (unsigned i = 0; i < n; ++i) {
        kx[b * n + i] = kernel_half(spatial_freq(i, n, p.dx), p.dz, P->k_host[b]);
        ky[b * n + i] = kernel_half(spatial_freq(i, n, p.kind == PXG_FFT3D ? p.dy : p.dx), p.dz,
                                    P->k_host[b]);
      }
    upload_cx<double>(P->kx, kx, s);
    upload_cx<double>(P->ky, ky, s);
    if (p.kind == PXG_FFT3D && (p.absorb_npix_x || p.absorb_npix_y)) {
      upload_real<double>(P->absx, absorber_table(n, p.absorb_npix_x, p.absorb_inv_x), s);
      upload_real<double>(P->absy, absorber_table(n, p.absorb_npix_y, p.absorb_inv_y), s);
    } else {
      P->absx.release();
      P->absy.release();
    }
  }
  P->kvec.alloc(sizeof(double) * p.batch);
  hip_check(hipMemcpyAsync(P->kvec.p, P->k_host.data(), P->kvec.bytes, hipMemcpyHostToDevice, s),
            "upload k");
  hip_check(hipStreamSynchronize(s), "upload sync");
}

void build_part_factors(pxg_plan P)
{
  const pxg_problem &p = P->p;
  const double inv_norm = 1.0 / static_cast<double>(P->cells);
  const int np = P->n_parts;
  std::vector<cx<double>> pf(static_cast<size_t>(p.batch) * (np + 1));
  for (int b = 0; b < p.batch; ++b) {
    for (int e = 0; e < np; ++e) {
      const DPart &q = P->parts_host[e];
      const double d = q.op == 0 ? q.delta : 0.0, be = q.op == 0 ? q.beta : 0.0;
      pf[b * (np + 1) + e] = screen_factor(P->k_host[b], p.dz, d, be, inv_norm);
    }
    pf[b * (np + 1) + np] = screen_factor(P->k_host[b], p.dz, 0.0, 0.0, inv_norm);
  }
  upload_cx<double>(P->part_fac, pf, P->stream);
}

void rebuild_tables(pxg_plan P)
{
  build_tables(P);
  if (P->mat_mode == MAT_CSG) build_part_factors(P);
}

MatDev mat_dev(pxg_plan P)
{
  MatDev m{};
  m.mode = P->mat_mode;
  m.shapes = P->shapes.as<DShape>();
  m.parts = P->parts.as<DPart>();
  m.boxes = P->boxes.as<DBox>();
  m.n_parts = P->n_parts;
  m.is_ref = P->is_ref;
  m.part_fac = P->part_fac.p;
  m.planes = P->planes.as<double>();
  m.n_planes = P->n_planes;
  m.first_plane = P->first_plane;
  m.z_invariant = P->z_invariant;
  m.plane_stride = 2 * P->cells;
  m.wg = P->wg;
  return m;
}

// ------------------------------------------------------------------ launches
// Profiling records: begin() before a launch, record() after it, on the
// stream the kernel is launched on (the plan's stream unless given).
void record(pxg_plan P, const char *name, double bytes, hipError_t e, int ev0, hipStream_t s = nullptr)
{
  hip_check(e, name);
  if (P->profiling) {
    const int ev1 = static_cast<int>(P->recs.size()) * 2 + 1;
    hip_check(hipEventRecord(P->ev[ev1], s ? s : P->stream), "event");
    P->recs.push_back({name, ev0, ev1, bytes});
  }
}

int begin(pxg_plan P, hipStream_t s = nullptr)
{
  if (!P->profiling) return -1;
  const int ev0 = static_cast<int>(P->recs.size()) * 2;
  if (static_cast<size_t>(ev0 + 2) > P->ev.size()) fail(PXG_ERR_STATE, "event pool exhausted");
  hip_check(hipEventRecord(P->ev[ev0], s ? s : P->stream), "event");
  return ev0;
}

template <class R>
FftArgs<R> fft_args(pxg_plan P)
{
  const pxg_problem &p = P->p;
  FftArgs<R> a{};
  a.field = P->field.as<cx<R>>();
  a.fstride = P->cells;
  a.tw = P->tw.as<cd>();
  a.kx = P->kx.as<cd>();
  a.ky = P->ky.as<cd>();
  a.absx = P->absx.as<double>();
  a.absy = P->absy.as<double>();
  a.kvec = P->kvec.as<double>();
  a.mat = mat_dev(P);
  a.xmin = p.xmin;
  a.dx = p.dx;
  a.ymin = p.ymin;
  a.dy = p.dy;
  a.zmin = p.zmin;
  a.dz = p.dz;
  a.inv_norm = 1.0 / static_cast<double>(P->cells);
  a.n_refr = p.n_refr ? p.n_refr : 10;
  a.quirks = p.quirks;
  a.mcode = P->mcode.as<uint16_t>();
  a.mprev = a.mcode;
  a.mxdb = P->mxdb.as<cd>();
  // the map is computed once per step for the batch: inside pass X when the
  // batch gives it many work-groups to spread over, else in its own launch
  // (always for the Bluestein passes)
  a.blue_m = P->blue_m;
  a.chirp = P->chirp.as<cd>();
  a.bf = P->blue_bf.as<cd>();
  a.bi = P->blue_bi.as<cd>();
  a.map_in_pass_x = p.batch >= MAP_IN_PASS_X_BATCH && !P->blue_m;
  return a;
}

bool store_step(pxg_plan P, unsigned st)
{
  const pxg_problem &p = P->p;
  return p.store_dsz > 0 && st % p.store_dsz == 0 &&
         static_cast<int>(st / p.store_dsz) < P->sol_dims[2];
}

template <class R>
void cube_store(pxg_plan P, unsigned st)
{
  const double bytes = static_cast<double>(P->cells) * P->p.batch * sizeof(cx<R>);
  const int ev0 = begin(P);
  record(P, "cube_store", bytes,
         launch_cube_store<R>(P->field.as<cx<R>>(), P->cells, P->p.nx, P->sol_dims[0],
                              P->sol.as<cx<double>>(), P->sol_fstride,
                              static_cast<int>(st / P->p.store_dsz), P->p.batch, P->stream),
         ev0);
}

// The plan's second and third streams (default priority) and their fork and
// join events, each made on first use; a failed creation leaves the handles
// made so far, and the next call makes the rest.
void make_aux_streams(pxg_plan P)
{
  for (hipStream_t *x : {&P->aux, &P->aux2})
    if (!*x) {
      hipStream_t made = nullptr;
      hip_check(hipStreamCreateWithFlags(&made, hipStreamNonBlocking), "hipStreamCreate");
      *x = made;
    }
  for (hipEvent_t *e : {&P->fork_ev, &P->join_ev, &P->join2_ev})
    if (!*e) {
      hipEvent_t made = nullptr;
      hip_check(hipEventCreateWithFlags(&made, hipEventDisableTiming), "event");
      *e = made;
    }
}

// The steps of run_fft3d for the two halves of the batch, each on a stream of
// its own with its own material map (built by its field-0 work-groups).  The
// halves' launches overlap: while one half's pass drains (its last work-groups
// leave CUs idle) the other half's next pass fills them.  Every field sees the
// same arithmetic as in one launch over the whole batch, so the results are
// bitwise the same.  The plan stream forks to the second stream at the start
// and joins it at the end, so work enqueued on the plan later stays ordered.
template <class R>
void run_fft3d_halves(pxg_plan P, unsigned first, unsigned nsteps)
{
  const pxg_problem &p = P->p;
  const int n = p.nx, nb = p.batch / 2;
  FftArgs<R> h[2] = {fft_args<R>(P), fft_args<R>(P)};
  const size_t b0 = static_cast<size_t>(nb);
  h[1].field += b0 * P->cells;
  h[1].kx += b0 * n;
  h[1].ky += b0 * n;
  h[1].kvec += b0;
  if (h[1].mat.part_fac)
    h[1].mat.part_fac = static_cast<const cd *>(h[1].mat.part_fac) + b0 * (P->n_parts + 1);
  h[1].mcode = P->mcode2.as<uint16_t>();
  h[1].mxdb = P->mxdb2.as<cd>();
  for (auto &a : h) a.map_in_pass_x = nb >= MAP_IN_PASS_X_BATCH;
  make_aux_streams(P);
  const hipStream_t s[2] = {P->stream, P->aux};
  hip_check(hipEventRecord(P->fork_ev, P->stream), "fork");
  hip_check(hipStreamWaitEvent(P->aux, P->fork_ev, 0), "fork");
  // the plan stream joins the second stream on every exit, a failed launch
  // included, so later work on the plan stays ordered behind what was enqueued
  struct Join {
    pxg_plan P;
    ~Join()
    {
      if (hipEventRecord(P->join_ev, P->aux) == hipSuccess) (void)hipStreamWaitEvent(P->stream, P->join_ev, 0);
    }
  } join{P};
  long ms[2] = {P->map_step, P->map_step2};
  P->map_step = P->map_step2 = -1;  // until the march has enqueued every step
  for (int i = 0; i < 2; ++i) {
    h[i].mode = Y_FWD;
    hip_check(launch_pass_y<R>(n, P->lines_y, h[i], nb, s[i]), "pass y");
  }
  for (unsigned st = first; st < first + nsteps; ++st) {
    const bool last = st + 1 == first + nsteps;
    for (int i = 0; i < 2; ++i) {
      h[i].step = st;
      h[i].map_prev_valid = ms[i] >= 0 && ms[i] + 1 == static_cast<long>(st);
      ms[i] = -1;
      if (P->mat_mode == MAT_CSG && !h[i].map_in_pass_x)
        hip_check(launch_material_map<R>(n, h[i], s[i]), "material map");
      hip_check(launch_pass_x<R>(n, P->lines_x, h[i], nb, s[i]), "pass x");
      if (P->mat_mode == MAT_CSG) ms[i] = st;
    }
    for (int i = 0; i < 2; ++i) {
      h[i].mode = Y_INV | (last ? 0 : Y_FWD);
      hip_check(launch_pass_y<R>(n, P->lines_y, h[i], nb, s[i]), "pass y");
    }
  }
  P->map_step = ms[0];
  P->map_step2 = ms[1];
}

template <class R>
void run_fft3d(pxg_plan P, unsigned first, unsigned nsteps)
{
  const pxg_problem &p = P->p;
  // batches in halves, except under profiling (one launch per pass is what it
  // times) and graph capture (one stream)
  if (fft3d_halves(p) && !P->one_launch && !P->profiling && !P->use_graphs && P->mcode2.p) {
    run_fft3d_halves<R>(P, first, nsteps);
    return;
  }
  P->map_step2 = -1;  // the second half's map is not kept up to date here
  FftArgs<R> a = fft_args<R>(P);
  const double pass_bytes = 2.0 * P->cells * p.batch * sizeof(cx<R>);
  const int n = p.nx;
  // Maps of their own launches, MAP_RING steps per launch: the parts are culled
  // once for the steps of a launch and each step's z0 plane is the previous
  // step's z1 plane in registers.  Step st reads slot (st - first) mod MAP_RING.
  const bool ring = P->mat_mode == MAT_CSG && !a.map_in_pass_x && !P->blue_m && map_slots(p) == MAP_RING;
  uint16_t *const mc0 = a.mcode;
  cd *const mx0 = a.mxdb;
  a.mode = Y_FWD;
  int ev0 = begin(P);
  record(P, "pass_y_fwd", pass_bytes, launch_pass_y<R>(n, P->lines_y, a, p.batch, P->stream), ev0);
  for (unsigned st = first; st < first + nsteps; ++st) {
    const bool last = st + 1 == first + nsteps;
    const bool store = store_step(P, st);
    a.step = st;  // the material map of this step (CSG) is built before pass Y
    a.map_prev_valid = P->map_step >= 0 && P->map_step + 1 == static_cast<long>(st);
    P->map_step = -1;  // until this step's map is launched
    if (ring) {
      const unsigned k = (st - first) % MAP_RING;
      if (k == 0) {
        const int nz = static_cast<int>(std::min<unsigned>(MAP_RING, first + nsteps - st));
        FftArgs<R> m = a;
        m.mcode = mc0;
        m.mxdb = mx0;
        // the previous step's map: slot 0 before this call (the single-map
        // paths), else the last slot of the previous launch (read by each lane
        // for its cells before any write of this launch)
        m.mprev = st == first ? mc0 : mc0 + (MAP_RING - 1) * P->cells;
        m.map_prev_valid = st == first ? a.map_prev_valid : 1;
        ev0 = begin(P);
        record(P, "material_map", static_cast<double>(P->cells) * nz * sizeof(uint16_t),
               launch_material_map_planes<R>(n, m, nz, P->stream), ev0);
      }
      a.mcode = mc0 + k * P->cells;
      a.mxdb = mx0 + k * P->cells;
    } else if (P->mat_mode == MAT_CSG && !a.map_in_pass_x) {
      ev0 = begin(P);
      record(P, "material_map", static_cast<double>(P->cells) * sizeof(uint16_t),
             launch_material_map<R>(n, a, P->stream), ev0);
    }
    ev0 = begin(P);
    record(P, "pass_x", pass_bytes, launch_pass_x<R>(n, P->lines_x, a, p.batch, P->stream), ev0);
    // slot 0 holds this step's map (the single-map paths continue from it)
    if (P->mat_mode == MAT_CSG && (!ring || a.mcode == mc0)) P->map_step = st;
    a.mode = Y_INV | ((last || store) ? 0 : Y_FWD);
    ev0 = begin(P);
    record(P, "pass_y", pass_bytes, launch_pass_y<R>(n, P->lines_y, a, p.batch, P->stream), ev0);
    if (store) {
      cube_store<R>(P, st);
      if (!last) {
        a.mode = Y_FWD;
        ev0 = begin(P);
        record(P, "pass_y_fwd", pass_bytes, launch_pass_y<R>(n, P->lines_y, a, p.batch, P->stream),
               ev0);
      }
    }
  }
}

// Steps [first[i], first[i] + nsteps) of batch-1 FFT3D plans pl[0..m) of one
// power-of-two line length and precision: the steps of run_fft3d (forward
// FFT_y, then per step the material map, pass X and pass Y, the last pass Y
// leaving real space), each launch covering every member.  pl[0]'s stream waits
// for every member's earlier work and runs the launches; every member's stream
// then waits for them.
template <class R>
void run_fft3d_group(pxg_plan *pl, int m, const unsigned *first, unsigned nsteps)
{
  pxg_plan L = pl[0];
  const int n = L->p.nx;
  hipStream_t s = L->stream;
  GroupArgs<R> g{};
  bool any_csg = false;
  for (int i = 0; i < m; ++i) {
    pxg_plan P = pl[i];
    if (i > 0) {
      hip_check(hipEventRecord(P->ready_ev, P->stream), "group join");
      hip_check(hipStreamWaitEvent(s, P->ready_ev, 0), "group join");
    }
    g.m[i] = fft_args<R>(P);
    g.m[i].map_in_pass_x = 0;  // batch 1: the map has its own launch
    g.m[i].step = first[i];
    g.m[i].map_prev_valid = P->map_step >= 0 && P->map_step + 1 == static_cast<long>(first[i]);
    any_csg |= P->mat_mode == MAT_CSG;
    P->map_step = -1;  // until the run's maps are launched
  }
  g.mode = Y_FWD;
  hip_check(launch_pass_y_group<R>(n, L->lines_y, g, m, s), "pass y");
  for (unsigned k = 0; k < nsteps; ++k) {
    for (int i = 0; i < m; ++i) {
      g.m[i].step = first[i] + k;
      if (k > 0) g.m[i].map_prev_valid = 1;  // the previous step's map is in mcode
    }
    if (any_csg) hip_check(launch_material_map_group<R>(n, g, m, s), "material map");
    hip_check(launch_pass_x_group<R>(n, L->lines_x, g, m, s), "pass x");
    g.mode = Y_INV | (k + 1 == nsteps ? 0 : Y_FWD);
    hip_check(launch_pass_y_group<R>(n, L->lines_y, g, m, s), "pass y");
  }
  for (int i = 0; i < m; ++i)
    if (pl[i]->mat_mode == MAT_CSG) pl[i]->map_step = static_cast<long>(first[i] + nsteps - 1);
  hip_check(hipEventRecord(L->group_done_ev, s), "group done");
  for (int i = 1; i < m; ++i) hip_check(hipStreamWaitEvent(pl[i]->stream, L->group_done_ev, 0), "group done");
}

// Lines per work-group the FFT3D passes are built for at size n: powers of two
// up to 16, at most 1024 lanes, FP64 line storage within the LDS of one CU.
bool lines_supported(int n, int l)
{
  return fft_size_supported(n) && l >= 1 && l <= 16 && (l & (l - 1)) == 0 && l * (n / 8) <= 1024 && n % l == 0 &&
         static_cast<size_t>(l) * (n + n / 8 + (l >= 8 ? 1 : 8 / l)) * sizeof(cd) <= 163840 - 1024;  // pass X's stride
}

// Mean duration of `reps` launches of one FFT3D pass with `lines` lines per
// work-group, on the plan's own field and material (mid-run z-step).
// First of `span` steps at which the autotuner runs the plan: the middle of its
// z range, or for host-sampled material planes that change with z the first
// steps they cover (-1: they cover fewer than span steps; the plan is not timed).
long tune_step(pxg_plan P, unsigned span)
{
  const long mid = std::max(1, P->p.nz / 2);
  if (P->mat_mode != MAT_PLANES || P->z_invariant) return mid;
  const long per = P->p.kind == PXG_ADI ? 2 : 1;  // planes per step (ADI: one per half-step)
  const long st = std::max<long>(1, (P->first_plane + per - 1) / per);
  if ((st + static_cast<long>(span)) * per - 1 >= static_cast<long>(P->first_plane) + P->n_planes) return -1;
  return st;
}

template <class R>
double time_pass(pxg_plan P, bool pass_x, int lines, int reps)
{
  const pxg_problem &p = P->p;
  FftArgs<R> a = fft_args<R>(P);
  a.step = static_cast<unsigned>(tune_step(P, 1));
  a.mode = Y_INV | Y_FWD;
  auto launch = [&] {
    return pass_x ? launch_pass_x<R>(p.nx, lines, a, p.batch, P->stream)
                  : launch_pass_y<R>(p.nx, lines, a, p.batch, P->stream);
  };
  // pass Y reads the material map of the step
  if (!pass_x && P->mat_mode == MAT_CSG) {
    if (a.map_in_pass_x)
      hip_check(launch_pass_x<R>(p.nx, P->lines_x, a, p.batch, P->stream), "pass x");
    else
      hip_check(launch_material_map<R>(p.nx, a, P->stream), "material map");
  }
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  float ms = 0.0f;
  hipError_t err = launch();
  if (err == hipSuccess) err = hipEventRecord(e0, P->stream);
  for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch();
  if (err == hipSuccess) err = hipEventRecord(e1, P->stream);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  hip_check(err, "autotune launch");
  return static_cast<double>(ms) / reps;
}

template <class R>
void autotune_fft3d(pxg_plan P, int reps)
{
  const int n = P->p.nx;
  double best_x = 1e30, best_y = 1e30;
  int lx = P->lines_x, ly = P->lines_y;
  for (int l = 1; l <= 16; l *= 2) {
    if (!lines_supported(n, l)) continue;
    const double tx = time_pass<R>(P, true, l, reps);
    if (tx < best_x) best_x = tx, lx = l;
    const double ty = time_pass<R>(P, false, l, reps);
    if (ty < best_y) best_y = ty, ly = l;
  }
  P->lines_x = lx;
  P->lines_y = ly;
}

// Batches that can march in two halves on two streams (run_fft3d_halves): time
// steps of the march both ways, each in two rounds keeping its faster one, and
// keep the faster way.  The results are the same bits either way.
template <class R>
void autotune_halves(pxg_plan P, int reps)
{
  P->one_launch = false;
  if (!fft3d_halves(P->p) || !P->mcode2.p || P->profiling || P->use_graphs) return;
  const unsigned steps = static_cast<unsigned>(std::max(2, reps));
  const long at = tune_step(P, steps);
  if (at < 0) return;
  const unsigned first = static_cast<unsigned>(at);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  double best[2] = {1e30, 1e30};
  try {
    for (int round = 0; round < 2; ++round)
      for (int way = 0; way < 2; ++way) {
        P->one_launch = way == 1;
        run_fft3d<R>(P, first, 1);  // warm: the way's streams and maps
        hip_check(hipEventRecord(e0, P->stream), "event");
        run_fft3d<R>(P, first, steps);
        hip_check(hipEventRecord(e1, P->stream), "event");
        hip_check(hipEventSynchronize(e1), "event");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
        best[way] = std::min(best[way], static_cast<double>(ms));
      }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P->one_launch = best[1] < best[0];
}

template <class R>
void run_fft2d(pxg_plan P, unsigned first, unsigned nsteps)
{
  const pxg_problem &p = P->p;
  FftArgs<R> a = fft_args<R>(P);
  const double pass_bytes = 2.0 * P->cells * p.batch * sizeof(cx<R>);
  for (unsigned st = first; st < first + nsteps; ++st) {
    a.step = st;
    a.sol = (p.store_dsz > 0 && static_cast<int>(st) < p.nz) ? P->sol.as<cx<R>>() : nullptr;
    a.sol_fstride = P->sol_fstride;
    const int ev0 = begin(P);
    record(P, "fft2d_step", pass_bytes, launch_fft2d_step<R>(p.nx, a, p.batch, P->stream), ev0);
  }
}

template <class R>
TriArgs<R> tri_args(pxg_plan P)
{
  const pxg_problem &p = P->p;
  TriArgs<R> a{};
  a.field = P->field.as<cx<R>>();
  a.n = p.nx;
  a.batch = p.batch;
  a.kvec = P->kvec.as<double>();
  a.mat = mat_dev(P);
  a.xmin = p.xmin;
  a.dx = p.dx;
  a.ymin = p.ymin;
  a.dy = p.dy;
  a.zmin = p.zmin;
  a.dz = p.dz;
  a.bc = p.bc;
  a.use_tbc = p.use_tbc;
  a.quirks = p.quirks;
  a.absorb_npix = p.absorb_npix_x;
  a.absorb_inv = p.absorb_inv_x;
  a.scratch = P->scratch.p;
  a.sol = p.store_dsz > 0 ? P->sol.as<cx<R>>() : nullptr;
  a.sol_fstride = P->sol_fstride;
  a.nz = p.nz;
  a.kmin = *std::min_element(P->k_host.begin(), P->k_host.end());
  a.khost = P->k_host.data();
  return a;
}

// Mean duration of `reps` ADI half-steps (x- or y-implicit) of a mid-run step
// with tile shape `tile`, on the plan's own field, scratch and material.
template <class R>
double time_adi_half(pxg_plan P, bool implicit_x, int tile, int reps)
{
  TriArgs<R> a = tri_args<R>(P);
  a.first_step = static_cast<unsigned>(tune_step(P, 1));
  a.adi_tile = tile;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  float ms = 0.0f;
  hipError_t err = launch_adi_half<R>(a, implicit_x, P->stream);
  if (err == hipSuccess) err = hipEventRecord(e0, P->stream);
  for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch_adi_half<R>(a, implicit_x, P->stream);
  if (err == hipSuccess) err = hipEventRecord(e1, P->stream);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  hip_check(err, "autotune launch");
  return static_cast<double>(ms) / reps;
}

template <class R>
void autotune_adi(pxg_plan P, int reps)
{
  for (int h = 0; h < 2; ++h) {
    double best = 1e30;
    int pick = P->adi_tile[h];
    for (int t = 0; t < adi_tile_count(); ++t) {
      const double ms = time_adi_half<R>(P, h == 0, t, reps);
      if (ms < best) best = ms, pick = t;
    }
    P->adi_tile[h] = pick;
  }
}

template <class R>
void build_cn_jobs(pxg_plan P)
{
  const int nb = P->is_jobs ? static_cast<int>(P->jp.size()) : P->p.batch;
  std::vector<CnJob> jobs(nb);
  size_t scoff = 0;
  P->max_P = 1;
  const MatDev shared = mat_dev(P);
  for (int b = 0; b < nb; ++b) {
    const pxg_problem &q = P->is_jobs ? P->jp[b] : P->p;
    const CnPartition cp = cn_partition(q.nx);
    CnJob &J = jobs[b];
    J.nx = q.nx;
    J.P = cp.P;
    J.m = cp.m;
    J.bc = q.bc;
    J.nz = q.nz;
    J.absorb_npix = q.absorb_npix_x;
    J.absorb_inv = q.absorb_inv_x;
    J.foff = P->foff[b];
    J.soff = q.store_dsz > 0 ? P->soff[b] : ~static_cast<size_t>(0);
    J.scoff = scoff;
    J.xmin = q.xmin;
    J.dx = q.dx;
    J.zmin = q.zmin;
    J.dz = q.dz;
    J.k = P->k_host[b];
    J.mat = shared;
    J.eq = P->eq_planes.p ? P->eq_planes.as<double>() : nullptr;
    J.eq_first = P->eq_first;
    J.eq_inv = P->eq_inv;
    if (P->is_jobs && P->job_mat[b] == MAT_CURVED) {
      J.mat.mode = MAT_CURVED;
      J.mat.wg = P->job_wg[b];
    }
    if (J.mat.mode == MAT_PLANES) J.mat.plane_stride = 2 * static_cast<size_t>(q.nx);
    scoff += cn_job_scratch(q.nx, sizeof(cx<R>));
    P->max_P = std::max(P->max_P, cp.P);
  }
  if (scoff > P->scratch.bytes) fail(PXG_ERR_STATE, "CN scratch too small (%zu < %zu)", P->scratch.bytes, scoff);
  // March kernels by chunk length: field and coefficients in registers (up
  // to cn_reg_rows() rows), everything in scratch for chunks of up to 8 rows
  // and beyond cn_big_rows(), field in registers with coefficients in scratch
  // in between (measured: 47 against 52 us/step at 16 rows, 34 against 30 at
  // 8 rows).
  std::vector<CnJob> reg, big, glob;
  P->max_P = P->max_P_reg = P->max_P_big = 1;
  for (const CnJob &J : jobs) {
    if (J.eq) {
      // a ParaxialEquation other than the default: the general march builds every row
      glob.push_back(J);
      P->max_P = std::max(P->max_P, J.P);
    } else if (J.m <= cn_reg_rows()) {
      reg.push_back(J);
      P->max_P_reg = std::max(P->max_P_reg, J.P);
    } else if (J.m > 8 && J.m <= cn_big_rows()) {
      big.push_back(J);
      P->max_P_big = std::max(P->max_P_big, J.P);
    } else {
      glob.push_back(J);
      P->max_P = std::max(P->max_P, J.P);
    }
  }
  P->njobs_reg = static_cast<int>(reg.size());
  // the fixed-matrix march when every register-class job has a z-invariant material
  P->reg_fixed = !reg.empty() && P->max_P_reg <= cn_fixed_max_p();
  for (const CnJob &J : reg)
    if (!(J.mat.mode == MAT_NONE || (J.mat.mode == MAT_PLANES && J.mat.z_invariant))) P->reg_fixed = false;
  P->njobs_big = static_cast<int>(big.size());
  P->njobs_glob = static_cast<int>(glob.size());
  if (!big.empty()) {
    P->jobs_big_dev.alloc(big.size() * sizeof(CnJob));
    hip_check(hipMemcpyAsync(P->jobs_big_dev.p, big.data(), P->jobs_big_dev.bytes, hipMemcpyHostToDevice,
                             P->stream),
              "upload CN jobs");
  }
  if (!glob.empty()) {
    P->jobs_dev.alloc(glob.size() * sizeof(CnJob));
    hip_check(hipMemcpyAsync(P->jobs_dev.p, glob.data(), P->jobs_dev.bytes, hipMemcpyHostToDevice, P->stream),
              "upload CN jobs");
  }
  if (!reg.empty()) {
    P->jobs_reg_dev.alloc(reg.size() * sizeof(CnJob));
    hip_check(hipMemcpyAsync(P->jobs_reg_dev.p, reg.data(), P->jobs_reg_dev.bytes, hipMemcpyHostToDevice,
                             P->stream),
              "upload CN jobs");
  }
  hip_check(hipStreamSynchronize(P->stream), "sync");
  P->jobs_dirty = false;
}

template <class R>
void run_cn(pxg_plan P, unsigned first, unsigned nsteps)
{
  if (P->jobs_dirty) build_cn_jobs<R>(P);
  size_t cells = 0;
  for (size_t c : P->fcells) cells += c;
  const double bytes = 2.0 * cells * sizeof(cx<R>) * nsteps;
  // Each job set with a kernel of its own marches on its own stream,
  // concurrently: k_cn_jobs on aux, k_cn_big on aux2, k_cn_reg on the plan's
  // stream (a set alone runs on the plan's stream).
  const int sets = (P->njobs_glob > 0) + (P->njobs_big > 0) + (P->njobs_reg > 0);
  hipStream_t gs = P->stream, bs = P->stream;
  if (sets > 1) {
    make_aux_streams(P);
    hip_check(hipEventRecord(P->fork_ev, P->stream), "fork");
    if (P->njobs_glob) {
      hip_check(hipStreamWaitEvent(P->aux, P->fork_ev, 0), "fork");
      gs = P->aux;
    }
    if (P->njobs_big && (P->njobs_glob || P->njobs_reg)) {
      hip_check(hipStreamWaitEvent(P->aux2, P->fork_ev, 0), "fork");
      bs = P->aux2;
    }
  }
  if (P->njobs_glob) {
    const int ev0 = begin(P, gs);
    record(P, "cn_march", bytes,
           launch_cn_jobs<R>(P->jobs_dev.as<CnJob>(), P->njobs_glob, P->max_P, P->field.as<cx<R>>(),
                             P->scratch.p, P->sol.as<cx<R>>(), first, nsteps, gs),
           ev0, gs);
  }
  if (P->njobs_big) {
    const int ev0 = begin(P, bs);
    record(P, "cn_march_big", bytes,
           launch_cn_big<R>(P->jobs_big_dev.as<CnJob>(), P->njobs_big, P->max_P_big, P->field.as<cx<R>>(),
                            P->scratch.p, P->sol.as<cx<R>>(), first, nsteps, bs),
           ev0, bs);
  }
  if (P->njobs_reg) {
    const int ev0 = begin(P);
    // the factorization is paid once per call: short calls keep the plain march
    if (P->reg_fixed && nsteps >= 8)
      record(P, "cn_march_fixed", bytes,
             launch_cn_fixed<R>(P->jobs_reg_dev.as<CnJob>(), P->njobs_reg, P->max_P_reg, P->field.as<cx<R>>(),
                                P->sol.as<cx<R>>(), first, nsteps, P->stream),
             ev0);
    else
      record(P, "cn_march_reg", bytes,
             launch_cn_reg<R>(P->jobs_reg_dev.as<CnJob>(), P->njobs_reg, P->max_P_reg, P->field.as<cx<R>>(),
                              P->sol.as<cx<R>>(), first, nsteps, P->stream),
             ev0);
  }
  if (gs != P->stream) {
    hip_check(hipEventRecord(P->join_ev, gs), "join");
    hip_check(hipStreamWaitEvent(P->stream, P->join_ev, 0), "join");
  }
  if (bs != P->stream) {
    hip_check(hipEventRecord(P->join2_ev, bs), "join");
    hip_check(hipStreamWaitEvent(P->stream, P->join2_ev, 0), "join");
  }
}

template <class R>
void run_adi(pxg_plan P, unsigned first, unsigned nsteps)
{
  TriArgs<R> a = tri_args<R>(P);
  const double pass_bytes = 2.0 * P->cells * P->p.batch * sizeof(cx<R>);
  for (unsigned st = first; st < first + nsteps; ++st) {
    a.first_step = st;
    a.adi_tile = P->adi_tile[0];
    int ev0 = begin(P);
    record(P, "adi_x", pass_bytes, launch_adi_half<R>(a, true, P->stream), ev0);
    a.adi_tile = P->adi_tile[1];
    ev0 = begin(P);
    record(P, "adi_y", pass_bytes, launch_adi_half<R>(a, false, P->stream), ev0);
    if (store_step(P, st)) cube_store<R>(P, st);
  }
}

template <class T>
void key_put(std::vector<char> &k, const T &v)
{
  const char *b = reinterpret_cast<const char *>(&v);
  k.insert(k.end(), b, b + sizeof(T));
}

// An FFT3D march as one HIP graph: everything run_fft3d enqueues is kernel
// launches on the plan stream (no memset, copy or host synchronisation), so the
// loop is captured once and replayed.  The key holds the call's steps, the
// work-group shapes, the map state the first step reads, the problem and the full
// argument block of the passes (every pointer and scalar the kernels receive): a
// replay therefore launches exactly what a fresh enqueue would.
template <class R>
void run_fft3d_graph(pxg_plan P, unsigned first, unsigned nsteps)
{
  std::vector<char> key;
  const FftArgs<R> a = fft_args<R>(P);
  key_put(key, first);
  key_put(key, nsteps);
  key_put(key, P->lines_x);
  key_put(key, P->lines_y);
  key_put(key, P->map_step);
  key_put(key, P->mat_mode);
  key_put(key, P->p);
  key_put(key, a);
  for (auto &g : P->graphs)
    if (g.key == key) {
      hip_check(hipGraphLaunch(g.exec, P->stream), "hipGraphLaunch");
      P->map_step = g.map_step_out;
      return;
    }
  hip_check(hipStreamBeginCapture(P->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  hipGraph_t graph = nullptr;
  try {
    run_fft3d<R>(P, first, nsteps);
  } catch (...) {
    (void)hipStreamEndCapture(P->stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  hip_check(hipStreamEndCapture(P->stream, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  hip_check(e, "hipGraphInstantiate");
  if (P->graphs.size() >= 8) {  // keep the 8 most recent shapes
    // the evicted graph may still be running on the stream
    hip_check(hipStreamSynchronize(P->stream), "sync");
    (void)hipGraphExecDestroy(P->graphs.front().exec);
    P->graphs.erase(P->graphs.begin());
  }
  P->graphs.push_back({std::move(key), P->map_step, exec});
  hip_check(hipGraphLaunch(exec, P->stream), "hipGraphLaunch");
}

void run(pxg_plan P, unsigned first, unsigned nsteps)
{
  if (!P->field_set) fail(PXG_ERR_STATE, "initial conditions not set (pxg_set_field)");
  if (first < 1) fail(PXG_ERR_INVALID, "steps start at 1 (step 0 is the initial condition)");
  if (nsteps == 0) return;
  const int k = P->p.kind;
  if (k == PXG_FFT3D && P->use_graphs && !P->profiling) {
    if (P->mixed)
      run_fft3d_graph<float>(P, first, nsteps);
    else
      run_fft3d_graph<double>(P, first, nsteps);
    return;
  }
  if (P->mixed) {
    if (k == PXG_FFT3D) run_fft3d<float>(P, first, nsteps);
    else if (k == PXG_FFT2D) run_fft2d<float>(P, first, nsteps);
    else if (k == PXG_CN) run_cn<float>(P, first, nsteps);
    else run_adi<float>(P, first, nsteps);
  } else {
    if (k == PXG_FFT3D) run_fft3d<double>(P, first, nsteps);
    else if (k == PXG_FFT2D) run_fft2d<double>(P, first, nsteps);
    else if (k == PXG_CN) run_cn<double>(P, first, nsteps);
    else run_adi<double>(P, first, nsteps);
  }
}

void check_plan(pxg_plan P)
{
  if (!P) fail(PXG_ERR_INVALID, "null plan");
}


void check_b(pxg_plan P, int b)
{
  check_plan(P);
  if (b < 0 || b >= P->p.batch) fail(PXG_ERR_INVALID, "batch index %d out of range [0,%d)", b, P->p.batch);
}

// Shape AABB in world coordinates from p = A^-1 (q - t) over the shape-space box.
DBox shape_box(const pxg_shape &s)
{
  const double *T = s.T;
  double A[3][3] = {{T[0], T[1], T[2]}, {T[4], T[5], T[6]}, {T[8], T[9], T[10]}};
  double t[3] = {T[3], T[7], T[11]};
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  DBox bx;
  if (!(std::fabs(det) > 1e-300)) {
    for (int i = 0; i < 3; ++i) {
      bx.lo[i] = -INFINITY;
      bx.hi[i] = INFINITY;
    }
    return bx;
  }
  double inv[3][3];
  inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
  inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
  inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  double e[3];
  if (s.type == PXG_SHAPE_SPHERE) {
    e[0] = e[1] = e[2] = std::fabs(s.p[0]);
  } else if (s.type == PXG_SHAPE_BOX) {
    for (int i = 0; i < 3; ++i) e[i] = std::fabs(s.p[i]) / 2.0;
  } else {
    const double r = std::max(std::fabs(s.p[0]), std::fabs(s.p[1]));
    e[0] = e[1] = r;
    e[2] = std::fabs(s.p[2]) / 2.0;
  }
  for (int i = 0; i < 3; ++i) {
    double c = -(inv[i][0] * t[0] + inv[i][1] * t[1] + inv[i][2] * t[2]);
    double h = std::fabs(inv[i][0]) * e[0] + std::fabs(inv[i][1]) * e[1] + std::fabs(inv[i][2]) * e[2];
    const double pad = 1e-9 * (std::fabs(c) + h) + 1e-9;
    bx.lo[i] = c - h - pad;
    bx.hi[i] = c + h + pad;
  }
  return bx;
}

// Host <-> device field transfers through the plan's pinned double buffer,
// in chunks of STAGE_ELEMS complex values; complex64 fields are converted on
// the host side of the copy.  Pageable copies ran at about 4.4 GB/s.
constexpr size_t STAGE_ELEMS = size_t(1) << 18;  // 4 MiB of complex128

void ensure_stage(pxg_plan P)
{
  for (int i = 0; i < 2; ++i) {
    if (!P->stage[i]) hip_check(hipHostMalloc(&P->stage[i], STAGE_ELEMS * 16, hipHostMallocDefault), "hipHostMalloc");
    if (!P->stage_ev[i]) {
      hip_check(hipEventCreateWithFlags(&P->stage_ev[i], hipEventDisableTiming), "event");
      hip_check(hipEventRecord(P->stage_ev[i], P->stream), "stage event");
    }
  }
}

void upload_field(pxg_plan P, char *dst, const double *c128, size_t n)
{
  ensure_stage(P);
  int k = 0;
  for (size_t i0 = 0; i0 < n; i0 += STAGE_ELEMS, k ^= 1) {
    const size_t c = std::min(STAGE_ELEMS, n - i0);
    hip_check(hipEventSynchronize(P->stage_ev[k]), "stage wait");  // the buffer's previous DMA is done
    if (P->mixed) {
      auto *buf = static_cast<cx<float> *>(P->stage[k]);
      for (size_t i = 0; i < c; ++i)
        buf[i] = {static_cast<float>(c128[2 * (i0 + i)]), static_cast<float>(c128[2 * (i0 + i) + 1])};
    } else {
      std::memcpy(P->stage[k], c128 + 2 * i0, c * 16);
    }
    hip_check(hipMemcpyAsync(dst + i0 * P->esz, P->stage[k], c * P->esz, hipMemcpyHostToDevice, P->stream),
              "set_field");
    hip_check(hipEventRecord(P->stage_ev[k], P->stream), "stage event");
  }
  hip_check(hipStreamSynchronize(P->stream), "sync");
}

void download_field(pxg_plan P, const char *src, double *c128, size_t n)
{
  ensure_stage(P);
  hip_check(hipStreamSynchronize(P->stream), "sync");
  const size_t nch = (n + STAGE_ELEMS - 1) / STAGE_ELEMS;
  auto issue = [&](size_t ch) {
    const size_t i0 = ch * STAGE_ELEMS, c = std::min(STAGE_ELEMS, n - i0);
    hip_check(hipMemcpyAsync(P->stage[ch & 1], src + i0 * P->esz, c * P->esz, hipMemcpyDeviceToHost, P->stream),
              "get_field");
    hip_check(hipEventRecord(P->stage_ev[ch & 1], P->stream), "stage event");
  };
  if (nch) issue(0);
  for (size_t ch = 0; ch < nch; ++ch) {
    hip_check(hipEventSynchronize(P->stage_ev[ch & 1]), "stage wait");
    const size_t i0 = ch * STAGE_ELEMS, c = std::min(STAGE_ELEMS, n - i0);
    // the other buffer is free (its chunk was consumed): start the next DMA
    // so that it overlaps this chunk's host copy
    if (ch + 1 < nch) issue(ch + 1);
    if (P->mixed) {
      const auto *buf = static_cast<const cx<float> *>(P->stage[ch & 1]);
      for (size_t i = 0; i < c; ++i) {
        c128[2 * (i0 + i)] = buf[i].x;
        c128[2 * (i0 + i) + 1] = buf[i].y;
      }
    } else {
      std::memcpy(c128 + 2 * i0, P->stage[ch & 1], c * 16);
    }
  }
}

}  // namespace

namespace pxg {
hipStream_t plan_stream(pxg_plan plan) { return plan ? plan->stream : nullptr; }
int ctx_device(pxg_ctx ctx) { return ctx ? ctx->device : 0; }
}  // namespace pxg

// =================================================================== C ABI
extern "C" {

const char *pxg_last_error(void) { return g_err.c_str(); }

const char *pxg_version(void)
{
  return "paxgpu 0.1 gfx950: stockham-fft(16..4096) fft3d-2pass fft2d cn-partitioned adi";
}

int pxg_ctx_create(int device, size_t hbm_budget_bytes, pxg_ctx *out)
{
  return guarded([&] {
    if (!out) fail(PXG_ERR_INVALID, "null output pointer");
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) fail(PXG_ERR_INVALID, "device %d not present (%d devices)", device, count);
    hip_check(hipSetDevice(device), "hipSetDevice");
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    {
      // each device holds its own copy of the no-alloc scratch at its own address
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      if (g_scratch_dev.size() < static_cast<size_t>(count)) g_scratch_dev.resize(count, nullptr);
      if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail(PXG_ERR_DEVICE, "no device address for the scratch array");
        g_scratch_dev[device] = ptr;
      }
    }
    auto *c = new pxg_ctx_s;
    c->device = device;
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : static_cast<size_t>(free_b * 0.9);
    *out = c;
  });
}

int pxg_ctx_destroy(pxg_ctx ctx)
{
  return guarded([&] {
    if (!ctx) fail(PXG_ERR_INVALID, "null context");
    bool last;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (ctx->closed) fail(PXG_ERR_STATE, "context already destroyed");
      ctx->closed = true;
      last = ctx->plans == 0;
    }
    if (last) delete ctx;
  });
}

size_t pxg_ctx_reserved(pxg_ctx ctx) { return ctx ? ctx->reserved.load() : 0; }

int pxg_plan_bytes(const pxg_problem *prob, size_t *bytes)
{
  return guarded([&] {
    if (!prob || !bytes) fail(PXG_ERR_INVALID, "null argument");
    validate(*prob);
    *bytes = plan_bytes(*prob);
  });
}

int pxg_plan_create(pxg_ctx ctx, const pxg_problem *prob, pxg_plan *out)
{
  pxg_plan P = nullptr;
  int rc = guarded([&] {
    if (!ctx || !prob || !out) fail(PXG_ERR_INVALID, "null argument");
    validate(*prob);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t need = plan_bytes(*prob);
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (ctx->closed) fail(PXG_ERR_STATE, "context destroyed");
      if (ctx->reserved.load() + need > ctx->budget)
        fail(PXG_ERR_BUDGET, "plan needs %zu bytes, budget %zu has %zu reserved", need, ctx->budget,
             ctx->reserved.load());
      ctx->reserved += need;
      ++ctx->plans;
    }
    P = new pxg_plan_s;
    P->ctx = ctx;
    P->reserved = need;
    P->p = *prob;
    if (P->p.n_refr == 0) P->p.n_refr = 10;
    P->mixed = field_c64(*prob);
    P->esz = P->mixed ? 8 : 16;
    const bool is3d = prob->kind == PXG_FFT3D || prob->kind == PXG_ADI;
    P->cells = is3d ? static_cast<size_t>(prob->nx) * prob->nx : prob->nx;
    hip_check(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate");
    P->field.alloc(P->cells * prob->batch * P->esz);
    hip_check(hipMemsetAsync(P->field.p, 0, P->field.bytes, P->stream), "memset");
    P->k_host.assign(prob->batch, prob->k);
    if (prob->kind == PXG_FFT3D) {
      P->mcode.alloc(P->cells * sizeof(uint16_t) * map_slots(*prob));
      P->mxdb.alloc(P->cells * sizeof(cd) * map_slots(*prob));
      if (fft3d_halves(*prob)) {
        P->mcode2.alloc(P->cells * sizeof(uint16_t));
        P->mxdb2.alloc(P->cells * sizeof(cd));
      }
    }
    if (prob->kind == PXG_FFT3D && fft_size_supported(prob->nx)) {
      // untuned defaults (pxg_plan_autotune measures): pass Y one wavefront of
      // lines per work-group, pass X wide work-groups for coalescing; the
      // Bluestein passes of other lengths run one line per work-group
      P->lines_y = std::min(max_lines(prob->nx, 81920), std::max(1, 512 / prob->nx));
      P->lines_x = max_lines(prob->nx, 162816);
      // keep enough work-groups in flight for small batches
      while (P->lines_y > 1 && (prob->nx / P->lines_y) * prob->batch < 512) P->lines_y /= 2;
      while (P->lines_x > 4 && (prob->nx / P->lines_x) * prob->batch < 512) P->lines_x /= 2;
    }
    // untuned x-implicit ADI tile: 128-byte rows of a line group, 16 lines of
    // complex64 (shape 0) or 8 of complex128 (shape 5: 8 x 128)
    if (prob->kind == PXG_ADI && prob->precision == PXG_PREC_FP64) P->adi_tile[0] = 5;
    if (prob->kind == PXG_FFT3D && prob->batch == 1) {
      hip_check(hipEventCreateWithFlags(&P->ready_ev, hipEventDisableTiming), "event");
      hip_check(hipEventCreateWithFlags(&P->group_done_ev, hipEventDisableTiming), "event");
    }
    if (prob->store_dsz > 0) {
      if (is3d) {
        const int cn = prob->nx / prob->store_dsx;
        P->sol_dims[0] = P->sol_dims[1] = cn;
        P->sol_dims[2] = prob->nz / prob->store_dsz + 1;
        P->sol_fstride = static_cast<size_t>(cn) * cn * P->sol_dims[2];
        P->sol.alloc(P->sol_fstride * prob->batch * sizeof(cx<double>));
      } else {
        P->sol_dims[0] = prob->nx;
        P->sol_dims[1] = prob->nz;
        P->sol_dims[2] = 1;
        P->sol_fstride = static_cast<size_t>(prob->nx) * prob->nz;
        P->sol.alloc(P->sol_fstride * prob->batch * P->esz);
      }
      hip_check(hipMemsetAsync(P->sol.p, 0, P->sol.bytes, P->stream), "memset");
    }
    if (prob->kind == PXG_CN || prob->kind == PXG_ADI) {
      const size_t sb = tridiag_scratch_bytes(*prob, P->mixed);
      if (sb) P->scratch.alloc(sb);
    }
    for (int b = 0; b < prob->batch; ++b) {
      P->foff.push_back(static_cast<size_t>(b) * P->cells);
      P->soff.push_back(static_cast<size_t>(b) * P->sol_fstride);
      P->fcells.push_back(P->cells);
    }
    rebuild_tables(P);
    hip_check(hipStreamSynchronize(P->stream), "sync");
    *out = P;
  });
  if (rc != PXG_OK && P) {
    std::string keep = g_err;
    pxg_plan_destroy(P);
    g_err = keep;
  } else if (rc != PXG_OK && ctx && prob && rc != PXG_ERR_BUDGET && rc != PXG_ERR_INVALID &&
             rc != PXG_ERR_UNSUPPORTED) {
    // reservation already released with the plan
  }
  return rc;
}

int pxg_plan_create_jobs(pxg_ctx ctx, const pxg_problem *probs, int n, pxg_plan *out)
{
  pxg_plan P = nullptr;
  int rc = guarded([&] {
    if (!ctx || !probs || !out || n < 1) fail(PXG_ERR_INVALID, "null argument or no jobs");
    size_t need = 0, cells = 0, sol = 0, scratch = 0;
    const int esz = field_c64(probs[0]) ? 8 : 16;
    for (int i = 0; i < n; ++i) {
      const pxg_problem &q = probs[i];
      validate(q);
      if (q.kind != PXG_CN) fail(PXG_ERR_UNSUPPORTED, "job plans support PXG_CN (job %d)", i);
      if (q.batch != 1) fail(PXG_ERR_INVALID, "job %d: batch must be 1", i);
      if (q.precision != probs[0].precision) fail(PXG_ERR_INVALID, "job %d: mixed precisions", i);
      cells += q.nx;
      if (q.store_dsz > 0) sol += static_cast<size_t>(q.nx) * q.nz;
      scratch += cn_job_scratch(q.nx, esz);
    }
    need = cells * esz + sol * esz + scratch + n * (sizeof(CnJob) + 64) + 4096;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (ctx->closed) fail(PXG_ERR_STATE, "context destroyed");
      if (ctx->reserved.load() + need > ctx->budget)
        fail(PXG_ERR_BUDGET, "job plan needs %zu bytes, budget %zu has %zu reserved", need, ctx->budget,
             ctx->reserved.load());
      ctx->reserved += need;
      ++ctx->plans;
    }
    P = new pxg_plan_s;
    P->ctx = ctx;
    P->reserved = need;
    P->p = probs[0];
    P->p.batch = n;
    P->is_jobs = true;
    P->jp.assign(probs, probs + n);
    P->mixed = field_c64(probs[0]);
    P->esz = esz;
    P->cells = 0;
    hip_check(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate");
    size_t fo = 0, so = 0;
    for (int i = 0; i < n; ++i) {
      P->foff.push_back(fo);
      P->soff.push_back(so);
      P->fcells.push_back(probs[i].nx);
      P->k_host.push_back(probs[i].k);
      fo += probs[i].nx;
      if (probs[i].store_dsz > 0) so += static_cast<size_t>(probs[i].nx) * probs[i].nz;
    }
    P->job_mat.assign(n, MAT_NONE);
    P->job_wg.assign(n, DCurved{});
    P->field.alloc(fo * esz);
    hip_check(hipMemsetAsync(P->field.p, 0, P->field.bytes, P->stream), "memset");
    if (so) {
      P->sol.alloc(so * esz);
      hip_check(hipMemsetAsync(P->sol.p, 0, P->sol.bytes, P->stream), "memset");
    }
    P->sol_total = so;
    P->scratch.alloc(scratch);
    P->kvec.alloc(sizeof(double) * n);
    hip_check(hipMemcpyAsync(P->kvec.p, P->k_host.data(), P->kvec.bytes, hipMemcpyHostToDevice, P->stream), "k");
    hip_check(hipStreamSynchronize(P->stream), "sync");
    *out = P;
  });
  if (rc != PXG_OK && P) {
    std::string keep = g_err;
    pxg_plan_destroy(P);
    g_err = keep;
  }
  return rc;
}

int pxg_plan_reset(pxg_plan P, const pxg_problem *prob)
{
  return guarded([&] {
    check_plan(P);
    if (!prob) fail(PXG_ERR_INVALID, "null problem");
    validate(*prob);
    const pxg_problem &o = P->p;
    if (P->is_jobs || prob->kind != o.kind || prob->precision != o.precision || prob->nx != o.nx ||
        prob->batch != o.batch || prob->store_dsx != o.store_dsx || prob->store_dsz != o.store_dsz ||
        (prob->store_dsz > 0 && prob->nz != o.nz))
      fail(PXG_ERR_INVALID, "pxg_plan_reset needs the same kind, precision, nx, batch and storage shape");
    hip_check(hipStreamSynchronize(P->stream), "sync");
    P->p = *prob;
    if (P->p.n_refr == 0) P->p.n_refr = 10;
    P->k_host.assign(prob->batch, prob->k);
    P->mat_mode = MAT_NONE;
    P->eq_planes.release();
    P->eq_n = P->eq_first = P->eq_inv = 0;
    P->map_step = -1;
    P->map_step2 = -1;
    P->field_set = false;
    P->jobs_dirty = true;
    rebuild_tables(P);
    if (P->sol.p) hip_check(hipMemsetAsync(P->sol.p, 0, P->sol.bytes, P->stream), "memset");
    hip_check(hipStreamSynchronize(P->stream), "sync");
  });
}

int pxg_set_job_material_curved_wg(pxg_plan P, int job, const pxg_curved_wg *wg)
{
  return guarded([&] {
    check_b(P, job);
    if (!P->is_jobs) fail(PXG_ERR_INVALID, "not a job plan");
    if (!wg) fail(PXG_ERR_INVALID, "null waveguide");
    P->job_wg[job] = DCurved{wg->R, wg->width, wg->clad_delta, wg->clad_beta, wg->in_delta, wg->in_beta,
                             wg->wglength, wg->dx, wg->dz, wg->use_smoothed};
    P->job_mat[job] = MAT_CURVED;
    P->jobs_dirty = true;
  });
}

int pxg_job_solution_dims(pxg_plan P, int job, int dims[3])
{
  return guarded([&] {
    check_b(P, job);
    if (!dims) fail(PXG_ERR_INVALID, "null output");
    const pxg_problem &q = P->is_jobs ? P->jp[job] : P->p;
    if (P->is_jobs) {
      dims[0] = q.store_dsz > 0 ? q.nx : 0;
      dims[1] = q.store_dsz > 0 ? q.nz : 0;
      dims[2] = q.store_dsz > 0 ? 1 : 0;
    } else {
      for (int i = 0; i < 3; ++i) dims[i] = P->sol_dims[i];
    }
  });
}

int pxg_plan_destroy(pxg_plan P)
{
  return guarded([&] {
    if (!P) fail(PXG_ERR_INVALID, "null plan");
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    for (auto e : P->ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
      if (P->stage[i]) (void)hipHostFree(P->stage[i]);
      if (P->stage_ev[i]) (void)hipEventDestroy(P->stage_ev[i]);
    }
    for (hipStream_t x : {P->aux, P->aux2})
      if (x) {
        (void)hipStreamSynchronize(x);
        (void)hipStreamDestroy(x);
      }
    for (hipEvent_t e : {P->fork_ev, P->join_ev, P->join2_ev, P->ready_ev, P->group_done_ev})
      if (e) (void)hipEventDestroy(e);
    for (auto &g : P->graphs) (void)hipGraphExecDestroy(g.exec);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    pxg_ctx c = P->ctx;
    bool last = false;
    if (c) {
      std::lock_guard<std::mutex> lk(c->mu);
      c->reserved -= P->reserved;
      last = --c->plans == 0 && c->closed;
    }
    delete P;
    if (last) delete c;
  });
}

int pxg_set_wavenumbers(pxg_plan P, const double *k, int n)
{
  return guarded([&] {
    check_plan(P);
    if (!k || n != P->p.batch) fail(PXG_ERR_INVALID, "need %d wavenumbers", P->p.batch);
    for (int i = 0; i < n; ++i)
      if (!(k[i] > 0)) fail(PXG_ERR_INVALID, "wavenumber %d must be > 0", i);
    P->k_host.assign(k, k + n);
    if (!P->is_jobs) rebuild_tables(P);
    P->jobs_dirty = true;
  });
}

int pxg_set_field(pxg_plan P, int b, const double *c128)
{
  return guarded([&] {
    check_b(P, b);
    if (!c128) fail(PXG_ERR_INVALID, "null field");
    const size_t n = P->fcells[b];
    char *dst = P->field.as<char>() + P->foff[b] * P->esz;
    upload_field(P, dst, c128, n);
    P->field_set = true;
    // copyCurrentSolution(0): slice / column 0 of the stored solution
    const pxg_problem &p = P->is_jobs ? P->jp[b] : P->p;
    if (p.store_dsz > 0) {
      const bool is3d = p.kind == PXG_FFT3D || p.kind == PXG_ADI;
      if (is3d) {
        // store slice 0 for this field only: temporarily view the batch entry
        const size_t cn = P->sol_dims[0];
        std::vector<double> host(2 * n);
        std::memcpy(host.data(), c128, sizeof(double) * 2 * n);
        if (P->mixed)
          for (size_t i = 0; i < 2 * n; ++i) host[i] = static_cast<float>(host[i]);
        const double d = static_cast<double>(p.nx) / cn;
        std::vector<cx<double>> slice(cn * cn);
        for (size_t i = 0; i < cn * cn; ++i) {
          const unsigned row = i % cn, col = i / cn;
          const unsigned r0 = row * d, r1 = row * d + d - 1, c0 = col * d, c1 = col * d + d - 1;
          double tx = 0, ty = 0;
          for (unsigned cc = c0; cc <= c1; ++cc) {
            double sx = 0, sy = 0;
            for (unsigned rr = r0; rr <= r1; ++rr) {
              sx += host[2 * (static_cast<size_t>(cc) * p.nx + rr)];
              sy += host[2 * (static_cast<size_t>(cc) * p.nx + rr) + 1];
            }
            tx += sx;
            ty += sy;
          }
          slice[i] = {tx / (d * d), ty / (d * d)};
        }
        char *sd = P->sol.as<char>() + P->soff[b] * sizeof(cx<double>);
        hip_check(hipMemcpyAsync(sd, slice.data(), slice.size() * sizeof(cx<double>),
                                 hipMemcpyHostToDevice, P->stream), "store slice 0");
      } else {
        char *sd = P->sol.as<char>() + P->soff[b] * P->esz;
        hip_check(hipMemcpyAsync(sd, dst, n * P->esz, hipMemcpyDeviceToDevice, P->stream),
                  "store column 0");
      }
      hip_check(hipStreamSynchronize(P->stream), "sync");
    }
  });
}

int pxg_get_field(pxg_plan P, int b, double *c128)
{
  return guarded([&] {
    check_b(P, b);
    if (!c128) fail(PXG_ERR_INVALID, "null output");
    const size_t n = P->fcells[b];
    const char *src = P->field.as<char>() + P->foff[b] * P->esz;
    download_field(P, src, c128, n);
  });
}

int pxg_solution_dims(pxg_plan P, int dims[3])
{
  return guarded([&] {
    check_plan(P);
    if (!dims) fail(PXG_ERR_INVALID, "null output");
    for (int i = 0; i < 3; ++i) dims[i] = P->sol_dims[i];
    if (P->is_jobs) fail(PXG_ERR_INVALID, "job plans: use pxg_job_solution_dims");
  });
}

int pxg_get_solution(pxg_plan P, int b, double *c128, size_t n_complex)
{
  return guarded([&] {
    check_b(P, b);
    const pxg_problem &q = P->is_jobs ? P->jp[b] : P->p;
    if (q.store_dsz <= 0) fail(PXG_ERR_STATE, "plan stores no solution (store_dsz = 0)");
    const bool is3d = q.kind == PXG_FFT3D || q.kind == PXG_ADI;
    const size_t n = P->is_jobs ? static_cast<size_t>(q.nx) * q.nz : P->sol_fstride;
    if (!c128 || n_complex < n) fail(PXG_ERR_INVALID, "output needs %zu complex values", n);
    hip_check(hipStreamSynchronize(P->stream), "sync");
    if (is3d || !P->mixed) {
      const char *src = P->sol.as<char>() + P->soff[b] * 16;
      hip_check(hipMemcpy(c128, src, n * 16, hipMemcpyDeviceToHost), "get_solution");
    } else {
      std::vector<cx<float>> tmp(n);
      const char *src = P->sol.as<char>() + P->soff[b] * 8;
      hip_check(hipMemcpy(tmp.data(), src, n * 8, hipMemcpyDeviceToHost), "get_solution");
      for (size_t i = 0; i < n; ++i) {
        c128[2 * i] = tmp[i].x;
        c128[2 * i + 1] = tmp[i].y;
      }
    }
  });
}

int pxg_set_equation_planes(pxg_plan P, const double *coef, int n_planes, int first_plane, int z_invariant)
{
  return guarded([&] {
    check_plan(P);
    if (P->p.kind != PXG_CN || P->is_jobs)
      fail(PXG_ERR_UNSUPPORTED, "a ParaxialEquation applies to single CN plans (CrankNicholson)");
    if (!coef) {
      P->eq_planes.release();
      P->eq_n = P->eq_first = P->eq_inv = 0;
    } else {
      if (n_planes < 1) fail(PXG_ERR_INVALID, "need at least one plane");
      const size_t per = 5 * static_cast<size_t>(P->p.nx);
      P->eq_planes.alloc(per * n_planes * sizeof(double));
      hip_check(hipMemcpy(P->eq_planes.p, coef, P->eq_planes.bytes, hipMemcpyHostToDevice), "equation planes");
      P->eq_n = n_planes;
      P->eq_first = first_plane;
      P->eq_inv = z_invariant != 0;
    }
    P->jobs_dirty = true;
  });
}

int pxg_set_material_none(pxg_plan P)
{
  return guarded([&] {
    check_plan(P);
    P->mat_mode = MAT_NONE;
    P->map_step = -1;
    P->map_step2 = -1;
    P->jobs_dirty = true;
  });
}

int pxg_set_material_csg(pxg_plan P, const pxg_shape *shapes, int n_shapes, const pxg_part *parts,
                         int n_parts, const pxg_module *modules, int n_modules, int is_reference_run)
{
  return guarded([&] {
    check_plan(P);
    if (n_shapes < 0 || n_parts < 0 || n_modules < 0) fail(PXG_ERR_INVALID, "negative count");
    if ((n_shapes && !shapes) || (n_parts && !parts) || (n_modules && !modules))
      fail(PXG_ERR_INVALID, "null CSG array");
    std::vector<DShape> ds(n_shapes);
    for (int i = 0; i < n_shapes; ++i) {
      if (shapes[i].type < 0 || shapes[i].type > 2) fail(PXG_ERR_INVALID, "shape %d: bad type", i);
      ds[i].type = shapes[i].type;
      ds[i].op = shapes[i].op;
      std::memcpy(ds[i].T, shapes[i].T, sizeof(double) * 12);
      std::memcpy(ds[i].p, shapes[i].p, sizeof(double) * 3);
      const double *T = shapes[i].T;
      ds[i].translate = T[0] == 1.0 && T[1] == 0.0 && T[2] == 0.0 && T[4] == 0.0 && T[5] == 1.0 &&
                        T[6] == 0.0 && T[8] == 0.0 && T[9] == 0.0 && T[10] == 1.0;
      const double r = shapes[i].p[0], r2 = r * r;
      ds[i].r2lo = r2 * (1.0 - 1e-14);
      ds[i].r2hi = r2 * (1.0 + 1e-14);
    }
    std::vector<DPart> order;
    std::vector<DBox> boxes;
    for (int m = 0; m < n_modules; ++m) {
      const pxg_module &mod = modules[m];
      if (mod.first_part < 0 || mod.n_parts < 0 || mod.first_part + mod.n_parts > n_parts)
        fail(PXG_ERR_INVALID, "module %d: part range out of bounds", m);
      for (int i = mod.n_parts - 1; i >= 0; --i) {
        const pxg_part &q = parts[mod.first_part + i];
        if (q.first_shape < 0 || q.n_shapes < 0 || q.first_shape + q.n_shapes > n_shapes)
          fail(PXG_ERR_INVALID, "part %d: shape range out of bounds", mod.first_part + i);
        order.push_back(DPart{q.first_shape, q.n_shapes, q.op, 0, q.delta, q.beta});
        DBox bx;
        for (int a = 0; a < 3; ++a) {
          bx.lo[a] = INFINITY;
          bx.hi[a] = -INFINITY;
        }
        for (int s = 0; s < q.n_shapes; ++s) {
          const pxg_shape &sh = shapes[q.first_shape + s];
          if (sh.op != PXG_OP_UNION) continue;
          DBox sb = shape_box(sh);
          for (int a = 0; a < 3; ++a) {
            bx.lo[a] = std::min(bx.lo[a], sb.lo[a]);
            bx.hi[a] = std::max(bx.hi[a], sb.hi[a]);
          }
        }
        boxes.push_back(bx);
      }
    }
    if (order.size() > 1024) fail(PXG_ERR_UNSUPPORTED, "at most 1024 CSG parts (got %zu)", order.size());
    P->shapes.alloc(std::max<size_t>(1, ds.size()) * sizeof(DShape));
    P->parts.alloc(std::max<size_t>(1, order.size()) * sizeof(DPart));
    P->boxes.alloc(std::max<size_t>(1, boxes.size()) * sizeof(DBox));
    if (!ds.empty())
      hip_check(hipMemcpy(P->shapes.p, ds.data(), ds.size() * sizeof(DShape), hipMemcpyHostToDevice), "csg");
    if (!order.empty()) {
      hip_check(hipMemcpy(P->parts.p, order.data(), order.size() * sizeof(DPart), hipMemcpyHostToDevice), "csg");
      hip_check(hipMemcpy(P->boxes.p, boxes.data(), boxes.size() * sizeof(DBox), hipMemcpyHostToDevice), "csg");
    }
    P->parts_host = order;
    P->n_parts = static_cast<int>(order.size());
    P->is_ref = is_reference_run;
    P->mat_mode = MAT_CSG;
    P->map_step = -1;
    P->map_step2 = -1;
    P->jobs_dirty = true;
    if (P->p.kind == PXG_FFT3D) build_part_factors(P);
  });
}

int pxg_set_material_curved_wg(pxg_plan P, const pxg_curved_wg *wg)
{
  return guarded([&] {
    check_plan(P);
    if (!wg) fail(PXG_ERR_INVALID, "null waveguide");
    if (P->p.kind == PXG_FFT3D || P->p.kind == PXG_ADI)
      fail(PXG_ERR_INVALID, "the curved waveguide is a 2D material");
    P->wg = DCurved{wg->R, wg->width, wg->clad_delta, wg->clad_beta, wg->in_delta, wg->in_beta,
                    wg->wglength, wg->dx, wg->dz, wg->use_smoothed};
    P->mat_mode = MAT_CURVED;
    P->map_step = -1;
    P->map_step2 = -1;
    P->jobs_dirty = true;
  });
}

int pxg_set_material_planes(pxg_plan P, const double *db, int n_planes, int first_plane, int z_invariant)
{
  return guarded([&] {
    check_plan(P);
    if (P->is_jobs) fail(PXG_ERR_UNSUPPORTED, "job plans take device materials only");
    if (!db || n_planes < 1) fail(PXG_ERR_INVALID, "need at least one plane");
    const size_t per = 2 * P->cells;
    P->planes.alloc(per * n_planes * sizeof(double));
    hip_check(hipMemcpy(P->planes.p, db, P->planes.bytes, hipMemcpyHostToDevice), "planes");
    P->n_planes = n_planes;
    P->first_plane = first_plane;
    P->z_invariant = z_invariant;
    P->mat_mode = MAT_PLANES;
    P->map_step = -1;
    P->map_step2 = -1;
    P->jobs_dirty = true;
  });
}

int pxg_plan_set_graphs(pxg_plan P, int enable)
{
  return guarded([&] {
    check_plan(P);
    P->use_graphs = enable != 0;
    if (!P->use_graphs) {
      hip_check(hipStreamSynchronize(P->stream), "sync");
      for (auto &g : P->graphs) (void)hipGraphExecDestroy(g.exec);
      P->graphs.clear();
    }
  });
}

int pxg_plan_set_lines(pxg_plan P, int lines_x, int lines_y)
{
  return guarded([&] {
    check_plan(P);
    if (P->p.kind != PXG_FFT3D) fail(PXG_ERR_UNSUPPORTED, "work-group shapes apply to FFT3D plans only");
    const int n = P->p.nx;
    if (P->blue_m && (lines_x > 1 || lines_y > 1))
      fail(PXG_ERR_UNSUPPORTED, "nx=%d (Bluestein passes) runs one line per work-group", n);
    if (P->blue_m) return;
    if ((lines_x && !lines_supported(n, lines_x)) || (lines_y && !lines_supported(n, lines_y)))
      fail(PXG_ERR_INVALID, "lines per work-group (%d, %d) not supported at nx=%d", lines_x, lines_y, n);
    if (lines_x) P->lines_x = lines_x;
    if (lines_y) P->lines_y = lines_y;
  });
}

int pxg_plan_get_lines(pxg_plan P, int *lines_x, int *lines_y)
{
  return guarded([&] {
    check_plan(P);
    if (!lines_x || !lines_y) fail(PXG_ERR_INVALID, "null output");
    *lines_x = P->lines_x;
    *lines_y = P->lines_y;
  });
}

int pxg_plan_set_adi_tiles(pxg_plan P, int x_tile, int y_tile)
{
  return guarded([&] {
    check_plan(P);
    if (P->p.kind != PXG_ADI) fail(PXG_ERR_UNSUPPORTED, "tile shapes apply to ADI plans only");
    for (int t : {x_tile, y_tile})
      if (t < -1 || t >= adi_tile_count()) fail(PXG_ERR_INVALID, "tile index %d outside [0, %d)", t, adi_tile_count());
    if (x_tile >= 0) P->adi_tile[0] = x_tile;
    if (y_tile >= 0) P->adi_tile[1] = y_tile;
  });
}

int pxg_plan_get_adi_tiles(pxg_plan P, int *x_tile, int *y_tile)
{
  return guarded([&] {
    check_plan(P);
    if (!x_tile || !y_tile) fail(PXG_ERR_INVALID, "null argument");
    *x_tile = P->adi_tile[0];
    *y_tile = P->adi_tile[1];
  });
}

int pxg_plan_autotune(pxg_plan P, int reps)
{
  return guarded([&] {
    check_plan(P);
    if ((P->p.kind == PXG_ADI || P->p.kind == PXG_FFT3D) && tune_step(P, 1) < 0) return;  // planes cover no step
    if (P->p.kind == PXG_ADI) {
      const std::array<int, 4> key = {P->p.nx, P->p.batch, P->p.precision, P->mat_mode};
      {
        std::lock_guard<std::mutex> lk(P->ctx->mu);
        auto it = P->ctx->tuned_adi.find(key);
        if (it != P->ctx->tuned_adi.end()) {
          P->adi_tile[0] = it->second.first;
          P->adi_tile[1] = it->second.second;
          return;
        }
      }
      if (reps <= 0) reps = 5;
      hip_check(hipStreamSynchronize(P->stream), "sync");
      if (P->mixed)
        autotune_adi<float>(P, reps);
      else
        autotune_adi<double>(P, reps);
      hip_check(hipMemsetAsync(P->field.p, 0, P->field.bytes, P->stream), "memset");
      hip_check(hipStreamSynchronize(P->stream), "sync");
      P->field_set = false;
      std::lock_guard<std::mutex> lk(P->ctx->mu);
      P->ctx->tuned_adi[key] = {P->adi_tile[0], P->adi_tile[1]};
      return;
    }
    if (P->p.kind != PXG_FFT3D || P->blue_m) return;  // the other kinds have one shape
    const std::array<int, 4> key = {P->p.nx, P->p.batch, P->p.precision, P->mat_mode};
    {
      std::lock_guard<std::mutex> lk(P->ctx->mu);
      auto it = P->ctx->tuned.find(key);
      if (it != P->ctx->tuned.end()) {
        P->lines_x = it->second[0];
        P->lines_y = it->second[1];
        P->one_launch = it->second[2] != 0;
        return;
      }
    }
    if (reps <= 0) reps = 5;
    hip_check(hipStreamSynchronize(P->stream), "sync");
    if (P->mixed) {
      autotune_fft3d<float>(P, reps);
      autotune_halves<float>(P, reps);
    } else {
      autotune_fft3d<double>(P, reps);
      autotune_halves<double>(P, reps);
    }
    hip_check(hipMemsetAsync(P->field.p, 0, P->field.bytes, P->stream), "memset");
    hip_check(hipStreamSynchronize(P->stream), "sync");
    P->field_set = false;
    P->map_step = -1;  // the timing runs wrote maps of another step
    P->map_step2 = -1;
    std::lock_guard<std::mutex> lk(P->ctx->mu);
    P->ctx->tuned[key] = {P->lines_x, P->lines_y, P->one_launch ? 1 : 0};
  });
}

// Host-sampled material and equation planes cover the steps of a call.
static void check_planes_cover(pxg_plan P, unsigned first_step, unsigned nsteps)
{
    if (P->mat_mode == MAT_PLANES && !P->z_invariant) {
      const pxg_problem &p = P->p;
      const int per_step = p.kind == PXG_ADI ? 2 : 1;
      const long lo = (p.kind == PXG_CN ? static_cast<long>(first_step) - 1 : static_cast<long>(first_step)) * per_step;
      const long hi = static_cast<long>(first_step + nsteps) * per_step - 1;
      if (lo < P->first_plane || hi >= static_cast<long>(P->first_plane) + P->n_planes)
        fail(PXG_ERR_INVALID, "material planes [%d, %d) do not cover steps %u..%u", P->first_plane,
             P->first_plane + P->n_planes, first_step, first_step + nsteps - 1);
    }
    if (P->eq_planes.p && !P->eq_inv &&
        (static_cast<long>(first_step) - 1 < P->eq_first ||
         static_cast<long>(first_step + nsteps) - 1 >= static_cast<long>(P->eq_first) + P->eq_n))
      fail(PXG_ERR_INVALID, "equation planes [%d, %d) do not cover steps %u..%u", P->eq_first,
           P->eq_first + P->eq_n, first_step, first_step + nsteps - 1);
}

int pxg_step(pxg_plan P, unsigned first_step, unsigned nsteps)
{
  return guarded([&] {
    check_plan(P);
    check_planes_cover(P, first_step, nsteps);
    P->profiling = false;
    run(P, first_step, nsteps);
  });
}

int pxg_step_group(pxg_plan *plans, int n_plans, const unsigned *first_steps, unsigned nsteps)
{
  return guarded([&] {
    if (!plans || !first_steps) fail(PXG_ERR_INVALID, "null argument");
    if (n_plans < 1 || n_plans > MAX_GROUP)
      fail(PXG_ERR_INVALID, "a group holds 1 to %d plans (got %d)", MAX_GROUP, n_plans);
    const pxg_plan L = plans[0];
    check_plan(L);
    for (int i = 0; i < n_plans; ++i) {
      const pxg_plan P = plans[i];
      check_plan(P);
      const pxg_problem &p = P->p;
      if (p.kind != PXG_FFT3D || p.batch != 1 || p.store_dsz != 0 || P->is_jobs)
        fail(PXG_ERR_UNSUPPORTED, "group member %d: FFT3D plans with batch 1 and no solution storage only", i);
      if (!fft_size_supported(p.nx))
        fail(PXG_ERR_UNSUPPORTED, "group member %d: power-of-two line lengths only (nx = %d)", i, p.nx);
      if (P->ctx != L->ctx || p.nx != L->p.nx || P->mixed != L->mixed)
        fail(PXG_ERR_INVALID, "group member %d: context, nx or precision differs from member 0", i);
      for (int j = 0; j < i; ++j)
        if (plans[j] == P) fail(PXG_ERR_INVALID, "plan given twice (members %d and %d)", j, i);
      if (!P->field_set) fail(PXG_ERR_STATE, "group member %d: initial conditions not set", i);
      if (first_steps[i] < 1) fail(PXG_ERR_INVALID, "steps start at 1 (member %d)", i);
      check_planes_cover(P, first_steps[i], nsteps);
    }
    if (nsteps == 0) return;
    if (L->mixed)
      run_fft3d_group<float>(plans, n_plans, first_steps, nsteps);
    else
      run_fft3d_group<double>(plans, n_plans, first_steps, nsteps);
  });
}

int pxg_sync(pxg_plan P)
{
  return guarded([&] {
    check_plan(P);
    hip_check(hipStreamSynchronize(P->stream), "hipStreamSynchronize");
  });
}

int pxg_step_profiled(pxg_plan P, unsigned first_step, unsigned nsteps, pxg_kernel_stat *stats,
                      int max_stats, int *n_stats, double *total_ms)
{
  return guarded([&] {
    check_plan(P);
    const size_t need = 2 * (static_cast<size_t>(nsteps) * 5 + 8);
    while (P->ev.size() < need) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      P->ev.push_back(e);
    }
    P->recs.clear();
    P->profiling = true;
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    hip_check(hipEventRecord(t0, P->stream), "event");
    try {
      run(P, first_step, nsteps);
    } catch (...) {
      P->profiling = false;
      (void)hipEventDestroy(t0);
      (void)hipEventDestroy(t1);
      throw;
    }
    P->profiling = false;
    hip_check(hipEventRecord(t1, P->stream), "event");
    hip_check(hipEventSynchronize(t1), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
    if (total_ms) *total_ms = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    struct Acc {
      const char *name;
      double sum;
      int n;
      double bytes;
    };
    std::vector<Acc> acc;
    for (const auto &r : P->recs) {
      float k = 0;
      hip_check(hipEventElapsedTime(&k, P->ev[r.e0], P->ev[r.e1]), "elapsed");
      auto it = std::find_if(acc.begin(), acc.end(), [&](const Acc &a) { return std::strcmp(a.name, r.name) == 0; });
      if (it == acc.end()) acc.push_back({r.name, k, 1, r.bytes});
      else {
        it->sum += k;
        it->n++;
      }
    }
    const int n = std::min<int>(max_stats, static_cast<int>(acc.size()));
    for (int i = 0; i < n && stats; ++i) stats[i] = {acc[i].name, acc[i].sum / acc[i].n, acc[i].n, acc[i].bytes};
    if (n_stats) *n_stats = n;
  });
}

int pxg_field_device_ptr(pxg_plan P, int b, void **dptr)
{
  return guarded([&] {
    check_b(P, b);
    if (!dptr) fail(PXG_ERR_INVALID, "null output");
    *dptr = P->field.as<char>() + P->foff[b] * P->esz;
  });
}

// ---- far field ----------------------------------------------------------------
namespace {

// post::FarField::farFieldAngleToIndx (src/farFieldPost.cpp:80-103), same
// operation order; |v| >= size/2 is tested before the conversion to int, which
// gives the reference's result without converting values beyond the int range.
unsigned ff_angle_index(double angle, unsigned size, double k, double step)
{
  const double v = size * k * step * std::sin(angle * PI / 180.0) / (2.0 * PI);
  if (!(std::fabs(v) < size / 2)) return angle > 0.0 ? size - 1 : 0;
  const int n = static_cast<int>(v);
  if (static_cast<unsigned>(std::abs(n)) >= size / 2) return angle > 0.0 ? size - 1 : 0;
  return static_cast<unsigned>(n + static_cast<int>(size / 2));
}

struct FfWindow {
  unsigned M = 0, ylo = 0, yhi = 0, xlo = 0, xhi = 0;
  int rows = 0, cols = 0;     // before the export resize
  bool resize = false;
  int out_rows = 0, out_cols = 0;
};

FfWindow ff_window(pxg_plan P, int b, unsigned pad, double phi_min, double phi_max, unsigned er,
                   unsigned ec)
{
  check_b(P, b);
  if (P->is_jobs || (P->p.kind != PXG_FFT3D && P->p.kind != PXG_ADI))
    fail(PXG_ERR_INVALID, "the far field needs a 3D plan (FFT3D or ADI)");
  FfWindow W;
  const unsigned n = static_cast<unsigned>(P->p.nx);
  // signalLength = max(pad, rows, cols) (src/farFieldPost.cpp:146-150)
  W.M = std::max(pad, n);
  if (W.M > (1u << 24)) fail(PXG_ERR_UNSUPPORTED, "far-field signal length %u above 2^24", W.M);
  const double k = P->k_host[b];
  W.ylo = ff_angle_index(phi_min, W.M, k, P->p.dy);
  W.yhi = ff_angle_index(phi_max, W.M, k, P->p.dy);
  W.xlo = ff_angle_index(phi_min, W.M, k, P->p.dx);
  W.xhi = ff_angle_index(phi_max, W.M, k, P->p.dx);
  if (W.yhi < W.ylo || W.xhi < W.xlo)
    fail(PXG_ERR_INVALID, "empty angular window [%g, %g] degrees", phi_min, phi_max);
  W.rows = static_cast<int>(W.yhi - W.ylo + 1);
  W.cols = static_cast<int>(W.xhi - W.xlo + 1);
  // resizeMatrix keeps the matrix when either export size reaches it
  W.resize = er > 0 && ec > 0 && !(er >= static_cast<unsigned>(W.rows) ||
                                   ec >= static_cast<unsigned>(W.cols));
  W.out_rows = W.resize ? static_cast<int>(er) : W.rows;
  W.out_cols = W.resize ? static_cast<int>(ec) : W.cols;
  return W;
}

std::vector<cx<double>> twiddle_table(unsigned n)
{
  std::vector<cx<double>> tw(n);
  for (unsigned m = 0; m < n; ++m) {
    const double a = 2.0 * PI * (static_cast<double>(m) / n);
    tw[m] = {std::cos(a), -std::sin(a)};
  }
  return tw;
}

}  // namespace

int pxg_far_field_dims(pxg_plan P, int b, unsigned pad, double phi_min, double phi_max,
                       unsigned export_rows, unsigned export_cols, int dims[2])
{
  return guarded([&] {
    if (!dims) fail(PXG_ERR_INVALID, "null output");
    const FfWindow W = ff_window(P, b, pad, phi_min, phi_max, export_rows, export_cols);
    dims[0] = W.out_rows;
    dims[1] = W.out_cols;
  });
}

int pxg_far_field(pxg_plan P, int b, pxg_plan R, int rb, unsigned pad, double phi_min,
                  double phi_max, unsigned export_rows, unsigned export_cols, double *out)
{
  return pxg_far_field_padded(P, b, R, rb, pad, PXG_PAD_ZERO, phi_min, phi_max, export_rows,
                              export_cols, out);
}

int pxg_far_field_padded(pxg_plan P, int b, pxg_plan R, int rb, unsigned pad, int padding,
                         double phi_min, double phi_max, unsigned export_rows, unsigned export_cols,
                         double *out)
{
  return guarded([&] {
    if (padding != PXG_PAD_ZERO && padding != PXG_PAD_CONSTANT && padding != PXG_PAD_LINEAR_FIT)
      fail(PXG_ERR_INVALID, "unknown padding %d", padding);
    if (padding == PXG_PAD_LINEAR_FIT && P->p.nx < 10)
      fail(PXG_ERR_UNSUPPORTED, "linear-fit padding fits 10 entries: needs at least 10 points per side");
    const FfWindow W = ff_window(P, b, pad, phi_min, phi_max, export_rows, export_cols);
    if (!out) fail(PXG_ERR_INVALID, "null output");
    if (!P->field_set) fail(PXG_ERR_STATE, "field not set (pxg_set_field)");
    const int n = P->p.nx;
    if (R) {
      check_b(R, rb);
      if (R->is_jobs || (R->p.kind != PXG_FFT3D && R->p.kind != PXG_ADI) || R->p.nx != n)
        fail(PXG_ERR_INVALID, "the reference must be a 3D field of the same size");
      if (R->ctx->device != P->ctx->device) fail(PXG_ERR_INVALID, "reference plan on another device");
      if (!R->field_set) fail(PXG_ERR_STATE, "reference field not set");
    }
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    if (R && R != P) hip_check(hipStreamSynchronize(R->stream), "reference sync");
    int np = 8;
    while (np < n) np *= 2;
    if (np > 4096) fail(PXG_ERR_UNSUPPORTED, "far field of fields above 4096 points per side");
    const size_t need = static_cast<size_t>(W.M) * 16 + static_cast<size_t>(np) * 16 +
                        static_cast<size_t>(W.rows) * n * 16 + static_cast<size_t>(W.rows) * W.cols * 8 +
                        (W.resize ? static_cast<size_t>(W.out_rows) * W.out_cols * 8 : 0);
    if (P->ctx->reserved.load() + need > P->ctx->budget)
      fail(PXG_ERR_BUDGET, "far field needs %zu bytes of scratch", need);
    hipStream_t s = P->stream;
    // the M- and NP-entry twiddle tables stay with the plan for later calls of the same lengths
    if (P->ff_twm_len != W.M) {
      P->ff_twm_len = 0;
      upload_cx<double>(P->ff_twm, twiddle_table(W.M), s);
      P->ff_twm_len = W.M;
    }
    if (P->ff_twn_len != static_cast<unsigned>(np)) {
      P->ff_twn_len = 0;
      upload_cx<double>(P->ff_twn, twiddle_table(static_cast<unsigned>(np)), s);
      P->ff_twn_len = static_cast<unsigned>(np);
    }
    DBuf tmp, res, rsz;
    tmp.alloc(static_cast<size_t>(W.rows) * n * sizeof(cd));
    res.alloc(static_cast<size_t>(W.rows) * W.cols * sizeof(double));
    FfArgs a{};
    a.field = P->field.as<char>() + P->foff[b] * P->esz;
    a.field_c64 = P->esz == 8;
    if (R) {
      a.ref = R->field.as<char>() + R->foff[rb] * R->esz;
      a.ref_c64 = R->esz == 8;
    }
    a.n = n;
    a.M = static_cast<int>(W.M);
    a.D = static_cast<int>(W.M) / np;
    a.tw_m = P->ff_twm.as<cd>();
    a.tw_np = P->ff_twn.as<cd>();
    a.ylo = static_cast<int>(W.ylo);
    a.ny = W.rows;
    a.xlo = static_cast<int>(W.xlo);
    a.nx = W.cols;
    a.tmp = tmp.as<cd>();
    a.out = res.as<double>();
    a.inv_len = 1.0 / W.M;
    a.pad_const = padding == PXG_PAD_CONSTANT;
    a.lin_L = static_cast<int>((W.M - static_cast<unsigned>(n)) / 2);
    a.pad_linear = padding == PXG_PAD_LINEAR_FIT && a.lin_L > 0;
    if (a.pad_linear) {
      // the fit grids of post::FarField::linearPadding (src/farFieldPost.cpp:271-273,
      // 296-298): getX(i) = xmin + i dx, Armadillo linspace (last point = end)
      auto getx = [&](int i) { return P->p.xmin + i * P->p.dx; };
      auto grid = [](double start, double end, double *x) {
        const double delta = (end - start) / 9.0;
        for (int k = 0; k < 9; ++k) x[k] = start + k * delta;
        x[9] = end;
      };
      // gsl_fit_linear's x-only running means: mean(x), then deviations and the mean square
      auto stats = [](const double *x, double *dxk, double &mx, double &mdx2) {
        mx = 0.0;
        for (int k = 0; k < 10; ++k) mx += (x[k] - mx) / (k + 1.0);
        mdx2 = 0.0;
        for (int k = 0; k < 10; ++k) {
          dxk[k] = x[k] - mx;
          mdx2 += (dxk[k] * dxk[k] - mdx2) / (k + 1.0);
        }
      };
      double x1[10], x2[10];
      grid(0.0, getx(10) - getx(0), x1);
      grid(getx(n - 10), getx(n), x2);
      stats(x1, a.fit1_dx, a.fit1_mx, a.fit1_mdx2);
      stats(x2, a.fit2_dx, a.fit2_mx, a.fit2_mdx2);
      a.lin_dxf = x1[1] - x1[0];
      a.lin_x2end = x2[9];
      a.lin_wrap = (P->p.quirks & PXG_FIX_LINEAR_PAD) == 0;
    }
    const bool pow2 = (W.M & (W.M - 1)) == 0 && W.M >= static_cast<unsigned>(np);
    if (pow2)
      hip_check(launch_far_field(a, np, s), "far field");
    else
      hip_check(launch_far_field_direct(a, s), "far field");
    const double *src = res.as<double>();
    if (W.resize) {
      rsz.alloc(static_cast<size_t>(W.out_rows) * W.out_cols * sizeof(double));
      hip_check(launch_ff_resize(res.as<double>(), W.rows, W.cols, rsz.as<double>(), W.out_rows,
                                 W.out_cols, s),
                "far field resize");
      src = rsz.as<double>();
    }
    hip_check(hipMemcpyAsync(out, src, static_cast<size_t>(W.out_rows) * W.out_cols * sizeof(double),
                             hipMemcpyDeviceToHost, s),
              "far field download");
    hip_check(hipStreamSynchronize(s), "far field sync");
  });
}

int pxg_intensity_trapezoid(pxg_plan P, int b, int n, const int *steps, const int *ix_start,
                            const int *ix_end, double *sums)
{
  return guarded([&] {
    check_b(P, b);
    const pxg_problem &q = P->is_jobs ? P->jp[b] : P->p;
    if (q.kind != PXG_CN && q.kind != PXG_FFT2D) fail(PXG_ERR_INVALID, "needs a 2D plan (CN or FFT2D)");
    if (q.store_dsz != 1 || q.store_dsx != 1)
      fail(PXG_ERR_STATE, "needs the full stored solution (store_dsz = store_dsx = 1)");
    if (n < 0 || (n > 0 && (!steps || !ix_start || !ix_end || !sums))) fail(PXG_ERR_INVALID, "null array");
    for (int i = 0; i < n; ++i) {
      if (steps[i] < 0 || steps[i] >= q.nz) fail(PXG_ERR_INVALID, "step %d out of [0,%d)", steps[i], q.nz);
      if (ix_start[i] < 0 || ix_start[i] >= q.nx || ix_end[i] < 0 || ix_end[i] >= q.nx)
        fail(PXG_ERR_INVALID, "index range [%d,%d] outside [0,%d)", ix_start[i], ix_end[i], q.nx);
    }
    if (n == 0) return;
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->stream;
    std::vector<int> idx(3 * static_cast<size_t>(n));
    std::copy(steps, steps + n, idx.begin());
    std::copy(ix_start, ix_start + n, idx.begin() + n);
    std::copy(ix_end, ix_end + n, idx.begin() + 2 * n);
    DBuf di, ds;
    di.alloc(idx.size() * sizeof(int));
    ds.alloc(static_cast<size_t>(n) * sizeof(double));
    hip_check(hipMemcpyAsync(di.p, idx.data(), di.bytes, hipMemcpyHostToDevice, s), "upload");
    const char *sol = P->sol.as<char>() + P->soff[b] * P->esz;
    const int *d = di.as<int>();
    hip_check(launch_trapz_intensity(sol, P->esz == 8, q.nx, n, d, d + n, d + 2 * n, ds.as<double>(), s),
              "trapezoid");
    hip_check(hipMemcpyAsync(sums, ds.p, ds.bytes, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

namespace {

// What a post-processing module reads: the current field or the stored solution
// of field/job b, as (rows, cols, slices) of complex cells on the device.
struct QuantitySource {
  const void *p = nullptr;
  int c64 = 0;
  int dims[3] = {0, 0, 0};
  bool matrix = false;  // a 2D result, which resizeMatrix applies to
  size_t cells() const { return static_cast<size_t>(dims[0]) * dims[1] * dims[2]; }
};

QuantitySource quantity_source(pxg_plan P, int b, int source)
{
  check_b(P, b);
  const pxg_problem &q = P->is_jobs ? P->jp[b] : P->p;
  const bool is3d = q.kind == PXG_FFT3D || q.kind == PXG_ADI;
  QuantitySource S;
  if (source == PXG_SRC_FIELD) {
    if (!P->field_set) fail(PXG_ERR_STATE, "field not set (pxg_set_field)");
    S.p = P->field.as<char>() + P->foff[b] * P->esz;
    S.c64 = P->esz == 8;
    S.dims[0] = q.nx;
    S.dims[1] = is3d ? q.nx : 1;
    S.dims[2] = 1;
    S.matrix = is3d;
  } else if (source == PXG_SRC_SOLUTION) {
    if (q.store_dsz <= 0) fail(PXG_ERR_STATE, "plan stores no solution (store_dsz = 0)");
    const int esz = (is3d || !P->mixed) ? 16 : 8;
    S.p = P->sol.as<char>() + P->soff[b] * esz;
    S.c64 = esz == 8;
    if (P->is_jobs) {
      S.dims[0] = q.nx;
      S.dims[1] = q.nz;
      S.dims[2] = 1;
    } else {
      for (int i = 0; i < 3; ++i) S.dims[i] = P->sol_dims[i];
    }
    S.matrix = !is3d;
  } else {
    fail(PXG_ERR_INVALID, "unknown source %d", source);
  }
  if (S.cells() == 0) fail(PXG_ERR_STATE, "empty %s", source == PXG_SRC_FIELD ? "field" : "solution");
  return S;
}

// resizeMatrix applies to matrices when both export sizes are set and both are
// below the matrix's (src/postProcessing.cpp:33-37 keeps the matrix otherwise)
bool quantity_resized(const QuantitySource &S, unsigned er, unsigned ec)
{
  return S.matrix && er > 0 && ec > 0 && er < static_cast<unsigned>(S.dims[0]) &&
         ec < static_cast<unsigned>(S.dims[1]);
}

}  // namespace

int pxg_field_quantity_dims(pxg_plan P, int b, int source, unsigned export_rows,
                            unsigned export_cols, int dims[3])
{
  return guarded([&] {
    if (!dims) fail(PXG_ERR_INVALID, "null output");
    const QuantitySource S = quantity_source(P, b, source);
    const bool rs = quantity_resized(S, export_rows, export_cols);
    dims[0] = rs ? static_cast<int>(export_rows) : S.dims[0];
    dims[1] = rs ? static_cast<int>(export_cols) : S.dims[1];
    dims[2] = S.dims[2];
  });
}

int pxg_field_quantity(pxg_plan P, int b, int source, int quantity, unsigned export_rows,
                       unsigned export_cols, double *out)
{
  return guarded([&] {
    const QuantitySource S = quantity_source(P, b, source);
    if (quantity < PXG_Q_ABS || quantity > PXG_Q_REAL) fail(PXG_ERR_INVALID, "unknown quantity %d", quantity);
    if (!out) fail(PXG_ERR_INVALID, "null output");
    const bool rs = quantity_resized(S, export_rows, export_cols);
    const size_t n = S.cells();
    const size_t n_out = rs ? static_cast<size_t>(export_rows) * export_cols : n;
    const size_t need = n * 8 + (rs ? n_out * 8 : 0);
    if (P->ctx->reserved.load() + need > P->ctx->budget)
      fail(PXG_ERR_BUDGET, "field quantity needs %zu bytes of scratch", need);
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->stream;
    DBuf res, rsz;
    res.alloc(n * sizeof(double));
    hip_check(launch_field_quantity(S.p, S.c64, n, quantity, res.as<double>(), s), "field quantity");
    const double *src = res.as<double>();
    if (rs) {
      rsz.alloc(n_out * sizeof(double));
      hip_check(launch_ff_resize(res.as<double>(), S.dims[0], S.dims[1], rsz.as<double>(),
                                 static_cast<int>(export_rows), static_cast<int>(export_cols), s),
                "field quantity resize");
      src = rsz.as<double>();
    }
    hip_check(hipMemcpyAsync(out, src, n_out * sizeof(double), hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

int pxg_field_uint8(pxg_plan P, int b, int source, int log_scale, const double *min_value,
                    const double *max_value, unsigned char *out)
{
  return guarded([&] {
    const QuantitySource S = quantity_source(P, b, source);
    if (log_scale != 0 && log_scale != 1) fail(PXG_ERR_INVALID, "log_scale must be 0 or 1");
    if (!out) fail(PXG_ERR_INVALID, "null output");
    const size_t n = S.cells();
    const int G = extrema_groups(n);
    const size_t need = n + static_cast<size_t>(G) * 16;
    if (P->ctx->reserved.load() + need > P->ctx->budget)
      fail(PXG_ERR_BUDGET, "8-bit image needs %zu bytes of scratch", need);
    // setMinValue / setMaxValue store the logarithm (src/postProcessMod.cpp:78-88)
    const double lo = min_value ? std::log(*min_value) : 0.0;
    const double hi = max_value ? std::log(*max_value) : 0.0;
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->stream;
    DBuf part, img;
    part.alloc(static_cast<size_t>(G) * 2 * sizeof(double));
    img.alloc(n);
    hip_check(launch_field_uint8(S.p, S.c64, n, part.as<double>(), log_scale, min_value != nullptr, lo,
                                 max_value != nullptr, hi, img.as<unsigned char>(), s),
              "8-bit image");
    hip_check(hipMemcpyAsync(out, img.p, n, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "sync");
  });
}

int pxg_thomas_batched(pxg_ctx ctx, int layout, int nsys, int n, double *diag, const double *sub,
                       double *rhs)
{
  return guarded([&] {
    if (!ctx) fail(PXG_ERR_INVALID, "null context");
    if (nsys < 1 || n < 1) fail(PXG_ERR_INVALID, "need nsys >= 1 and n >= 1");
    if (layout != PXG_LAYOUT_CONTIGUOUS && layout != PXG_LAYOUT_INTERLEAVED)
      fail(PXG_ERR_INVALID, "unknown layout %d", layout);
    if (!diag || !rhs || (n > 1 && !sub)) fail(PXG_ERR_INVALID, "null array");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t N = static_cast<size_t>(nsys) * n;
    const size_t need = N * 16 * 3 + N * 16 * 2;
    if (ctx->reserved.load() + need > ctx->budget) fail(PXG_ERR_BUDGET, "thomas needs %zu bytes", need);
    DBuf d, s, r, w;
    d.alloc(N * 16);
    s.alloc(std::max<size_t>(1, static_cast<size_t>(nsys) * (n - 1)) * 16);
    r.alloc(N * 16);
    w.alloc(N * 16 * 2);
    hip_check(hipMemcpy(d.p, diag, N * 16, hipMemcpyHostToDevice), "thomas upload");
    if (n > 1) hip_check(hipMemcpy(s.p, sub, static_cast<size_t>(nsys) * (n - 1) * 16, hipMemcpyHostToDevice), "thomas upload");
    hip_check(hipMemcpy(r.p, rhs, N * 16, hipMemcpyHostToDevice), "thomas upload");
    hip_check(launch_thomas_batched(layout, nsys, n, d.as<cx<double>>(), s.as<cx<double>>(),
                                    r.as<cx<double>>(), w.as<cx<double>>(), nullptr),
              "thomas kernel");
    hip_check(hipDeviceSynchronize(), "thomas sync");
    hip_check(hipMemcpy(diag, d.p, N * 16, hipMemcpyDeviceToHost), "thomas download");
    hip_check(hipMemcpy(rhs, r.p, N * 16, hipMemcpyDeviceToHost), "thomas download");
  });
}

int pxg_thomas_batched_noalloc(pxg_ctx ctx, int layout, int nsys, int n, double *diag, const double *sub,
                               double *rhs, void *stream)
{
  return guarded([&] {
    if (!ctx) fail(PXG_ERR_INVALID, "null context");
    if (nsys < 1 || n < 1) fail(PXG_ERR_INVALID, "need nsys >= 1 and n >= 1");
    if (layout != PXG_LAYOUT_CONTIGUOUS && layout != PXG_LAYOUT_INTERLEAVED)
      fail(PXG_ERR_INVALID, "unknown layout %d", layout);
    if (!diag || !rhs || (n > 1 && !sub)) fail(PXG_ERR_INVALID, "null array");
    const size_t rows = static_cast<size_t>(nsys) * static_cast<size_t>(n);
    if (rows > PXG_THOMAS_NOALLOC_MAX_ROWS)
      fail(PXG_ERR_UNSUPPORTED, "%zu rows (nsys x n) exceed the %u rows of the no-alloc scratch", rows,
           static_cast<unsigned>(PXG_THOMAS_NOALLOC_MAX_ROWS));
    void *work = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      if (static_cast<size_t>(ctx->device) < g_scratch_dev.size()) work = g_scratch_dev[ctx->device];
    }
    if (!work) fail(PXG_ERR_STATE, "scratch of device %d not set up", ctx->device);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    // the kernel writes every row's scratch entry before reading it: nothing to clear
    hip_check(launch_thomas_batched(layout, nsys, n, reinterpret_cast<cx<double> *>(diag),
                                    reinterpret_cast<const cx<double> *>(sub),
                                    reinterpret_cast<cx<double> *>(rhs), static_cast<cx<double> *>(work),
                                    static_cast<hipStream_t>(stream)),
              "thomas kernel");
  });
}

}  // extern "C"
