// Drop-in solvers: Solver, Solver2D (CrankNicholson, FFTSolver2D) and Solver3D
// (FFTSolver3D, ADI) on libpaxgpu.  The reference's plug-in point solveStep()
// advances the device field one step; solve() runs the whole z-march as one
// device call (in chunks when a host-sampled material has to be streamed).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <mutex>
#include <sstream>
#include <typeinfo>
#include <vector>

#include "paxpro/paxpro.hpp"

namespace paxpro {
namespace {
int g_default_precision = PXG_PREC_FP64;
}

pxg_ctx context()
{
  static std::once_flag once;
  static pxg_ctx ctx = nullptr;
  std::call_once(once, [] {
    const char *dev = std::getenv("PAXPRO_DEVICE");
    check(pxg_ctx_create(dev ? std::atoi(dev) : 0, 0, &ctx));
  });
  return ctx;
}

void setDefaultPrecision(int p) { g_default_precision = p; }
int defaultPrecision() { return g_default_precision; }
}  // namespace paxpro

using paxpro::check;

namespace {

// Device-evaluated material for the simulation, if the device can evaluate it
// exactly: a plain ParaxialSimulation with no material or a CSGMaterial, a
// GenericScattering run with a CSGMaterial or its reference (vacuum) run, or an
// unmodified CurvedWaveGuideFD (2D).  Anything else is host-sampled.
bool set_device_material(pxg_plan plan, const ParaxialSimulation &sim, bool is2d)
{
  if (is2d && typeid(sim) == typeid(CurvedWaveGuideFD)) {
    const auto &w = static_cast<const CurvedWaveGuideFD &>(sim);
    pxg_curved_wg wg{};
    wg.R = w.getRadiusOfCurvature();
    wg.width = w.getWidth();
    wg.clad_delta = w.getCladding().getDelta();
    wg.clad_beta = w.getCladding().getBeta();
    wg.in_delta = w.getInsideMaterial() ? w.getInsideMaterial()->getDelta() : 0.0;
    wg.in_beta = w.getInsideMaterial() ? w.getInsideMaterial()->getBeta() : 0.0;
    wg.wglength = w.getWaveguideLength();
    wg.dx = w.transverseDiscretization().step;
    wg.dz = w.longitudinalDiscretization().step;
    wg.use_smoothed = w.usesSmoothedWG();
    check(pxg_set_material_curved_wg(plan, &wg));
    return true;
  }
  if (!is2d && typeid(sim) == typeid(GenericScattering)) {
    // GenericScattering::getXrayMatProp: zero in the reference run, else the material
    const auto &gs = static_cast<const GenericScattering &>(sim);
    if (sim.material != nullptr && typeid(*sim.material) == typeid(CSGMaterial)) {
      const auto &csg = static_cast<const CSGMaterial &>(*sim.material);
      std::vector<pxg_shape> s;
      std::vector<pxg_part> p;
      std::vector<pxg_module> m;
      csg.flatten(s, p, m);
      check(pxg_set_material_csg(plan, s.data(), static_cast<int>(s.size()), p.data(),
                                 static_cast<int>(p.size()), m.data(), static_cast<int>(m.size()),
                                 gs.isReference() || csg.isReferenceRun));
      return true;
    }
    if (gs.isReference()) {
      check(pxg_set_material_none(plan));
      return true;
    }
    return false;
  }
  if (typeid(sim) != typeid(ParaxialSimulation)) return false;
  if (sim.material == nullptr) {
    check(pxg_set_material_none(plan));
    return true;
  }
  if (typeid(*sim.material) == typeid(CSGMaterial)) {
    const auto &csg = static_cast<const CSGMaterial &>(*sim.material);
    std::vector<pxg_shape> s;
    std::vector<pxg_part> p;
    std::vector<pxg_module> m;
    csg.flatten(s, p, m);
    check(pxg_set_material_csg(plan, s.data(), static_cast<int>(s.size()), p.data(),
                               static_cast<int>(p.size()), m.data(), static_cast<int>(m.size()),
                               csg.isReferenceRun));
    return true;
  }
  return false;
}

}  // namespace

// -------------------------------------------------------------------- Solver
Solver::Solver(const char *n, Dimension_t d) : dim(d), name(n), precision(paxpro::defaultPrecision()) {}
Solver::~Solver() { destroyPlan(); }
void Solver::reset()
{
  currentStep = 1;
  destroyPlan();
}
void Solver::destroyPlan()
{
  if (plan) pxg_plan_destroy(plan);
  plan = nullptr;
}
void Solver::setPrecision(int p)
{
  precision = p;
  destroyPlan();
}
const arma::cx_cube &Solver::getSolution3D() const
{
  throw std::runtime_error("The 3D version of getSolution() is not implemented!");
}
const arma::cx_mat &Solver::getSolution() const
{
  throw std::runtime_error("The 2D version of getSolution() is not implemented!");
}
const arma::cx_vec &Solver::getLastSolution() const
{
  throw std::runtime_error("The 2D version of getLastSolution() is not implemented!");
}
const arma::cx_mat &Solver::getLastSolution3D() const
{
  throw std::runtime_error("The 3D version of getLastSolution3D() is not implemented!");
}
arma::cx_mat &Solver::getLastSolution3D()
{
  throw std::runtime_error("The 3D version of non-const getLastSolution3D() is not implemented!");
}

// ------------------------------------------------------------------ Solver2D
Solver2D::~Solver2D()
{
  delete solution;
  delete prevSolution;
  delete currentSolution;
  if (ownsParaxialEquationObject) delete eq;
}

// src/solver2D.cpp:21-42
void Solver2D::setSimulator(ParaxialSimulation &wg)
{
  Solver::setSimulator(wg);
  destroyPlan();
  const unsigned nx = guide->nodeNumberTransverse();
  const unsigned nz = guide->nodeNumberLongitudinal();
  delete solution;
  delete prevSolution;
  delete currentSolution;
  solution = new arma::cx_mat(nx / guide->transverseDiscretization().downsamplingRatio, nz);
  prevSolution = new arma::cx_vec(nx);
  currentSolution = new arma::cx_vec(nx);
  if (ownsParaxialEquationObject) delete eq;
  eq = new ParaxialEquation();
  ownsParaxialEquationObject = true;
}

void Solver2D::setEquation(const ParaxialEquation &neweq)
{
  if (ownsParaxialEquationObject) delete eq;
  eq = &neweq;
  ownsParaxialEquationObject = false;
  destroyPlan();
}

void Solver2D::setInitialConditions(const arma::cx_vec &vec) { setLeftBC(vec.memptr()); }

void Solver2D::setLeftBC(const cdouble values[])
{
  if (solution == nullptr) throw std::runtime_error("A solver must be given before setting the boundary conditions!");
  for (unsigned i = 0; i < guide->nodeNumberTransverse(); ++i) (*currentSolution)(i) = values[i];
  copyCurrentSolution(0);
  fieldDirty = true;
  destroyPlan();
}

void Solver2D::setXBC(const cdouble top[], const cdouble bottom[])
{
  if (solution == nullptr) throw std::runtime_error("A solver must be given before setting the boundary conditions!");
  fetchSolution();
  const unsigned row = solution->n_rows - 1;
  for (unsigned i = 0; i < guide->nodeNumberLongitudinal(); ++i) {
    (*solution)(row, i) = top[i];
    (*solution)(0, i) = bottom[i];
  }
}

void Solver2D::realPart(double *r) const { realOrImagPart(r, Comp_t::REAL); }
void Solver2D::imagPart(double *r) const { realOrImagPart(r, Comp_t::IMAG); }

void Solver2D::realOrImagPart(double *out, Comp_t comp) const
{
  if (solution == nullptr) throw std::runtime_error("No equation system has been solved!");
  const arma::cx_mat &s = getSolution();
  const unsigned nx = guide->nodeNumberTransverse(), nz = guide->nodeNumberLongitudinal();
  for (unsigned ix = 0; ix < nx; ++ix)
    for (unsigned iz = 0; iz < nz; ++iz)
      out[ix * nz + iz] = comp == Comp_t::REAL ? s(ix, iz).real() : s(ix, iz).imag();
}

void Solver2D::getField(arma::mat &field) const
{
  const arma::cx_mat &s = getSolution();
  field.set_size(s.n_rows, s.n_cols);
  for (unsigned iz = 0; iz < s.n_cols; ++iz)
    for (unsigned ix = 0; ix < s.n_rows; ++ix) field(ix, iz) = s(ix, iz).real();
}

void Solver2D::getPhase(arma::mat &phase) const
{
  const arma::cx_mat &s = getSolution();
  phase.set_size(s.n_rows, s.n_cols);
  for (unsigned iz = 0; iz < s.n_cols; ++iz)
    for (unsigned ix = 0; ix < s.n_rows; ++ix) phase(ix, iz) = std::arg(s(ix, iz));
}

// src/solver2D.cpp:173-193 with transverse ratio 1: prev = cur; column `step`
// of the solution is the device's stored copy of cur.
void Solver2D::copyCurrentSolution(unsigned step)
{
  for (unsigned i = 0; i < currentSolution->n_elem; ++i) (*prevSolution)(i) = (*currentSolution)(i);
  if (step == 0 && solution) {
    for (unsigned i = 0; i < solution->n_rows && i < currentSolution->n_elem; ++i)
      (*solution)(i, 0) = (*currentSolution)(i);
  } else {
    solutionStale = true;
  }
}

void Solver2D::filterInLongitudinalDirection()
{
  if (guide->longitudinalDiscretization().downsamplingRatio != 1)
    throw std::runtime_error("longitudinal filtering needs the VISA low-pass filter (not in this build)");
}

// src/solver2D.cpp:267-282: keep column floor(iz*delta + delta/2).
void Solver2D::downSampleLongitudinalDirection()
{
  const arma::cx_mat &full = getSolution();
  const unsigned nz = guide->nodeNumberLongitudinal() / guide->longitudinalDiscretization().downsamplingRatio;
  auto *copy = new arma::cx_mat(full.n_rows, nz);
  const double delta = static_cast<double>(full.n_cols) / static_cast<double>(nz);
  for (unsigned iz = 0; iz < nz; ++iz) {
    const unsigned indx = iz * delta + delta / 2.0;
    for (unsigned ix = 0; ix < copy->n_rows; ++ix) (*copy)(ix, iz) = full(ix, indx);
  }
  delete solution;
  solution = copy;
}

const arma::cx_mat &Solver2D::getSolution() const
{
  fetchSolution();
  return *solution;
}

void Solver2D::fetchSolution() const
{
  if (!solutionStale || !plan) return;
  int dims[3];
  check(pxg_solution_dims(plan, dims));
  if (solution->n_rows != static_cast<arma::uword>(dims[0]) || solution->n_cols != static_cast<arma::uword>(dims[1]))
    solution->set_size(dims[0], dims[1]);
  check(pxg_get_solution(plan, 0, reinterpret_cast<double *>(solution->memptr()), solution->n_elem));
  solutionStale = false;
}

void Solver2D::solve()
{
  if (!solution || !guide || !eq) throw std::runtime_error("Solver2D::solve: solver not initialised");
  initValuesFromWaveGuide();
  const unsigned n = Nz > 1 ? Nz - 1 : 0;
  if (n) runOnDevice(currentStep, n);
  currentStep += n;
  if (!longitudinalFilterDisabled) {
    filterInLongitudinalDirection();
    downSampleLongitudinalDirection();
  }
}

void Solver2D::step()
{
  if (currentStep == 1) initValuesFromWaveGuide();
  solveStep(currentStep);
  copyCurrentSolution(currentStep);
  currentStep++;
}

void Solver2D::reset()
{
  Solver::reset();
  fieldDirty = true;
}

void Solver2D::initValuesFromWaveGuide()
{
  if (guide == nullptr) throw std::runtime_error("No waveguide specified!");
  Nx = guide->nodeNumberTransverse();
  Nz = guide->nodeNumberLongitudinal();
  stepX = guide->transverseDiscretization().step;
  xmin = guide->transverseDiscretization().min;
  stepZ = guide->longitudinalDiscretization().step;
  zmin = guide->longitudinalDiscretization().min;
  wavenumber = guide->getWavenumber();
}

void Solver2D::ensurePlan()
{
  if (plan && !fieldDirty) return;
  initValuesFromWaveGuide();
  if (!plan) {
    pxg_problem p{};
    p.kind = gpuKind();
    p.precision = precision;
    p.nx = Nx;
    p.nz = Nz;
    p.batch = 1;
    p.xmin = xmin;
    p.dx = stepX;
    p.ymin = 0.0;
    p.dy = stepX;
    p.zmin = zmin;
    p.dz = stepZ;
    p.k = wavenumber;
    p.bc = boundaryCondition == BC_t::TRANSPARENT ? PXG_BC_TRANSPARENT : PXG_BC_DIRICHLET;
    p.n_refr = 10;
    p.store_dsx = guide->transverseDiscretization().downsamplingRatio;
    p.store_dsz = 1;
    check(pxg_plan_create(paxpro::context(), &p, &plan));
  }
  check(pxg_set_field(plan, 0, reinterpret_cast<const double *>(prevSolution->memptr())));
  fieldDirty = false;
}

void Solver2D::runOnDevice(unsigned first, unsigned n)
{
  ensurePlan();
  if (!set_device_material(plan, *guide, true)) {
    // host sampling of the simulation's getXrayMatProp, as the reference does per cell
    const bool cn = gpuKind() == PXG_CN;
    const unsigned p0 = cn ? first - 1 : first, np = cn ? n + 1 : n;
    std::vector<double> db(static_cast<size_t>(np) * 2 * Nx);
    for (unsigned q = 0; q < np; ++q) {
      const double z = guide->getZ(p0 + q) + (cn ? 0.0 : 0.5 * stepZ);
      for (unsigned i = 0; i < Nx; ++i)
        guide->getXrayMatProp(guide->getX(i), z, db[(static_cast<size_t>(q) * Nx + i) * 2],
                              db[(static_cast<size_t>(q) * Nx + i) * 2 + 1]);
    }
    check(pxg_set_material_planes(plan, db.data(), np, p0, 0));
  }
  if (gpuKind() == PXG_CN && eq && typeid(*eq) != typeid(ParaxialEquation)) {
    // a ParaxialEquation of the user's: its coefficients sampled on the host where
    // CrankNicholson::solveStep evaluates them (src/crancNicholson.cpp:30-60):
    // F, G, J at x and H at x +- dx/2, on the planes z of steps first-1 .. first+n-1
    const unsigned p0 = first - 1, np = n + 1;
    std::vector<double> co(static_cast<size_t>(np) * 5 * Nx);
    for (unsigned q = 0; q < np; ++q) {
      const double z = guide->getZ(p0 + q);
      for (unsigned i = 0; i < Nx; ++i) {
        const double x = guide->getX(i);
        double *c = &co[(static_cast<size_t>(q) * Nx + i) * 5];
        c[0] = eq->F(x, z);
        c[1] = eq->G(x, z);
        c[2] = eq->H(x + 0.5 * stepX, z);
        c[3] = eq->H(x - 0.5 * stepX, z);
        c[4] = eq->J(x, z);
      }
    }
    check(pxg_set_equation_planes(plan, co.data(), np, p0, 0));
  }
  check(pxg_step(plan, first, n));
  check(pxg_get_field(plan, 0, reinterpret_cast<double *>(currentSolution->memptr())));
  for (unsigned i = 0; i < currentSolution->n_elem; ++i) (*prevSolution)(i) = (*currentSolution)(i);
  solutionStale = true;
}

void CrankNicholson::solveStep(unsigned iz) { runOnDevice(iz, 1); }
void FFTSolver2D::solveStep(unsigned step) { runOnDevice(step, 1); }

// ------------------------------------------------------------------ Solver3D
Solver3D::~Solver3D()
{
  delete solution;
  delete currentSolution;
  delete prevSolution;
}

// src/solver3D.cpp:19-44: Ny := Nx.
void Solver3D::setSimulator(ParaxialSimulation &sim)
{
  Solver::setSimulator(sim);
  destroyPlan();
  Nx = guide->nodeNumberTransverse();
  Nz = guide->nodeNumberLongitudinal();
  Ny = Nx;
  const unsigned dx = Nx / guide->transverseDiscretization().downsamplingRatio;
  const unsigned dz = Nz / guide->longitudinalDiscretization().downsamplingRatio;
  delete solution;
  delete prevSolution;
  delete currentSolution;
  prevSolution = new arma::cx_mat(Ny, Nx);
  currentSolution = new arma::cx_mat(Ny, Nx);
  solution = new arma::cx_cube(dx, dx, dz + 1);
}

void Solver3D::updateDimensionsOfArrays() { setSimulator(*guide); }

// src/solver3D.cpp:127-146
void Solver3D::setInitialConditions(const arma::cx_mat &values)
{
  if (!currentSolution || !prevSolution || !solution)
    throw std::runtime_error("The function setSimulator needs to be called before setInitialConditions!");
  if (values.n_rows != currentSolution->n_rows || values.n_cols != currentSolution->n_cols) {
    std::stringstream ss;
    ss << "Dimension of matrices does not match!\n";
    ss << "Given: Nrows: " << values.n_rows << " Ncols: " << values.n_cols;
    ss << "\nRequired: Nrows: " << currentSolution->n_rows << " Ncols: " << currentSolution->n_cols;
    throw std::runtime_error(ss.str());
  }
  *currentSolution = values;
  copyCurrentSolution(0);
  fieldDirty = true;
  destroyPlan();
}

void Solver3D::copyCurrentSolution(unsigned)
{
  *prevSolution = *currentSolution;
  solutionStale = true;
}

void Solver3D::reset()
{
  Solver::reset();
  fieldDirty = true;
}

void Solver3D::realTimeVisualization()
{
  std::clog << "Solver3D: realtime visualisation is not part of this build; ignored\n";
}
void Solver3D::setPlotLimits(double, double, double, double) {}
void Solver3D::setPlotLimits(double, double, double, double, bool) {}

const arma::cx_cube &Solver3D::getSolution3D() const
{
  if (solutionStale && plan) {
    int dims[3];
    check(pxg_solution_dims(plan, dims));
    if (solution->n_rows != static_cast<arma::uword>(dims[0]) || solution->n_slices != static_cast<arma::uword>(dims[2]))
      solution->set_size(dims[0], dims[1], dims[2]);
    check(pxg_get_solution(plan, 0, reinterpret_cast<double *>(solution->memptr()), solution->n_elem));
    solutionStale = false;
  }
  return *solution;
}

void Solver3D::ensurePlan()
{
  if (plan && !fieldDirty) return;
  if (!plan) {
    pxg_problem p{};
    p.kind = gpuKind();
    p.precision = precision;
    p.nx = Nx;
    p.nz = Nz;
    p.batch = 1;
    p.xmin = guide->transverseDiscretization().min;
    p.dx = guide->transverseDiscretization().step;
    p.ymin = guide->verticalDiscretization().min;
    p.dy = guide->verticalDiscretization().step;
    p.zmin = guide->longitudinalDiscretization().min;
    p.dz = guide->longitudinalDiscretization().step;
    p.k = guide->getWavenumber();
    p.use_tbc = 1;
    p.n_refr = 10;
    p.store_dsx = guide->transverseDiscretization().downsamplingRatio;
    p.store_dsz = guide->longitudinalDiscretization().downsamplingRatio;
    configure(p);
    check(pxg_plan_create(paxpro::context(), &p, &plan));
  }
  check(pxg_set_field(plan, 0, reinterpret_cast<const double *>(prevSolution->memptr())));
  fieldDirty = false;
}

void Solver3D::runOnDevice(unsigned first, unsigned n)
{
  ensurePlan();
  if (set_device_material(plan, *guide, false)) {
    check(pxg_step(plan, first, n));
  } else {
    // stream host-sampled planes in chunks of about 256 MiB
    const size_t per_step = static_cast<size_t>(planesPerStep()) * 2 * Nx * Ny * sizeof(double);
    const unsigned chunk = std::max<unsigned>(1, std::min<size_t>(n, (256u << 20) / per_step));
    std::vector<double> db;
    for (unsigned s = first; s < first + n; s += chunk) {
      const unsigned c = std::min(chunk, first + n - s);
      samplePlanes(s, c, db);
      check(pxg_set_material_planes(plan, db.data(), static_cast<int>(c * planesPerStep()),
                                    static_cast<int>(s * planesPerStep()), 0));
      check(pxg_step(plan, s, c));
    }
  }
  check(pxg_get_field(plan, 0, reinterpret_cast<double *>(currentSolution->memptr())));
}

void Solver3D::solveStep(unsigned step) { runOnDevice(step, 1); }

void Solver3D::step()
{
  solveStep(currentStep);
  copyCurrentSolution(currentStep++);
}

// src/solver3D.cpp:168-199: the whole march on the device, in chunks so that
// progress can be reported every secBetweenStatusMessage seconds.
void Solver3D::solve()
{
  auto last = std::chrono::steady_clock::now();
  const auto start = last;
  bool printEstimate = true;
  const unsigned total = Nz > 1 ? Nz - 1 : 0;
  unsigned done = 0;
  unsigned chunk = std::max(1u, std::min(total, 64u));
  while (done < total) {
    const unsigned c = std::min(chunk, total - done);
    solveStep(currentStep);  // first step of the chunk keeps solveStep overrides in the loop
    copyCurrentSolution(currentStep++);
    done += 1;
    if (c > 1) {
      runOnDevice(currentStep, c - 1);
      currentStep += c - 1;
      *prevSolution = *currentSolution;
      done += c - 1;
    }
    chunk = std::min(chunk * 2, 4096u);
    const auto now = std::chrono::steady_clock::now();
    const std::chrono::duration<double> el = now - last;
    if (el > std::chrono::duration<double>(secBetweenStatusMessage)) {
      std::clog << "Propagation step: " << done << " of " << Nz << std::endl;
      last = now;
      if (printEstimate) {
        const std::chrono::duration<double> tot = now - start;
        const int sec = static_cast<int>(tot.count() / done * Nz);
        std::clog << "Estimated time: " << sec / 3600 << "h " << (sec % 3600) / 60 << "min " << sec % 60 << "s\n";
        printEstimate = false;
      }
    }
  }
  solutionStale = true;
}

// -------------------------------------------------------------- FFTSolver3D
void FFTSolver3D::visualizeRealSpace()
{
  std::clog << "FFTSolver3D: realtime visualisation is not part of this build; ignored\n";
}
void FFTSolver3D::visualizeFourierSpace() { visualizeRealSpace(); }
void FFTSolver3D::setIntensityMinMax(double, double) {}
void FFTSolver3D::setPhaseMinMax(double, double) {}
void FFTSolver3D::storeImages(const char *) { visualizeRealSpace(); }

// src/fftSolver3D.cpp:327-349
void FFTSolver3D::absorbingBC(double width, double dampingLength)
{
  if (guide == nullptr)
    throw std::runtime_error("No waveguide set! The solver needs first to be passed to a waveguide object. Then call this function!");
  absorbX = width / guide->transverseDiscretization().step + 1;
  absorbY = width / guide->verticalDiscretization().step + 1;
  std::clog << "Absorbing BC: " << absorbX << " px in x-direction. " << absorbY << " px in y-direction.\n";
  invDampX = guide->transverseDiscretization().step / dampingLength;
  invDampY = guide->verticalDiscretization().step / dampingLength;
  destroyPlan();
  fieldDirty = true;
}

void FFTSolver3D::reset() { Solver3D::reset(); }

void FFTSolver3D::configure(pxg_problem &p) const
{
  p.absorb_npix_x = absorbX;
  p.absorb_npix_y = absorbY;
  p.absorb_inv_x = invDampX;
  p.absorb_inv_y = invDampY;
  p.n_refr = nStepsInRefrIntegral;
}

// Effective (delta, beta) per cell of FFTSolver3D::refraction
// (src/fftSolver3D.cpp:122-156,253-273) for a MaterialFunction the device
// cannot evaluate.
void FFTSolver3D::samplePlanes(unsigned first, unsigned n, std::vector<double> &db) const
{
  const size_t cells = static_cast<size_t>(Nx) * Ny;
  db.assign(cells * 2 * n, 0.0);
  const unsigned nr = nStepsInRefrIntegral;
  for (unsigned q = 0; q < n; ++q) {
    const unsigned st = first + q;
    const double z1 = guide->getZ(st), z0 = guide->getZ(st - 1);
    double *pl = db.data() + cells * 2 * q;
#pragma omp parallel for
    for (long i = 0; i < static_cast<long>(cells); ++i) {
      const unsigned row = i % Ny, col = i / Ny;
      const double x = guide->getX(col), y = guide->getY(row);
      double d, b, d0, b0;
      guide->getXrayMatProp(x, y, z1, d, b);
      guide->getXrayMatProp(x, y, z0, d0, b0);
      if (std::fabs(d - d0) > 1e-10 || std::fabs(b - b0) > 1e-10) {
        double dt, bt;
        guide->getXrayMatProp(x, y, z0, d, b);
        guide->getXrayMatProp(x, y, z1, dt, bt);
        d += dt;
        b += bt;
        const double h = (z1 - z0) / nr;
        for (unsigned s = 1; s < nr - 1; ++s) {
          guide->getXrayMatProp(x, y, z0 + s * h, dt, bt);
          d += 2.0 * dt;
          b += 2.0 * bt;
        }
        d /= (2.0 * nr);
        b /= (2.0 * nr);
      }
      pl[2 * i] = d;
      pl[2 * i + 1] = b;
    }
  }
}

// ----------------------------------------------------------------------- ADI
void ADI::configure(pxg_problem &p) const
{
  p.use_tbc = useTBC ? 1 : 0;
  p.quirks = quirks;
}

// Material of the x-implicit pass at z_step and the y-implicit pass at
// z_step + dz/2 (src/alternatingDirectionSolver.cpp:21,91).
void ADI::samplePlanes(unsigned first, unsigned n, std::vector<double> &db) const
{
  const size_t cells = static_cast<size_t>(Nx) * Ny;
  db.assign(cells * 2 * 2 * n, 0.0);
  const double h = guide->longitudinalDiscretization().step / 2.0;
  for (unsigned q = 0; q < 2 * n; ++q) {
    const unsigned st = first + q / 2;
    const double z = guide->getZ(st) + ((q & 1) ? h : 0.0);
    double *pl = db.data() + cells * 2 * q;
#pragma omp parallel for
    for (long i = 0; i < static_cast<long>(cells); ++i) {
      const unsigned row = i % Ny, col = i / Ny;
      guide->getXrayMatProp(guide->getX(col), guide->getY(row), z, pl[2 * i], pl[2 * i + 1]);
    }
  }
}
