// GenericScattering and post::FarField of the drop-in API.  The flow follows
// src/genericScattering.cpp:20-213 and src/farFieldPost.cpp:15-20,141-230,320-333
// of the reference; propagation and far field run on the device.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "paxpro/paxpro.hpp"

using paxpro::check;

// ------------------------------------------------------------------ post::FarField
void post::FarField::setAngleRange(double angMin, double angMax)
{
  phiMin = angMin;
  phiMax = angMax;
}

void post::FarField::addAttrib(std::vector<H5Attr> &attrs) const
{
  if (sim == nullptr) throw std::runtime_error("FarField: no ParaxialSimulation linked");
  const double pi = std::acos(-1.0);
  attrs.push_back(makeAttr("phiMin", phiMin));
  attrs.push_back(makeAttr("phiMax", phiMax));
  attrs.push_back(makeAttr("qmin", phiMin * sim->getWavenumber() * pi / 180.0));
  attrs.push_back(makeAttr("qmax", phiMax * sim->getWavenumber() * pi / 180.0));
}

void post::FarField::verifyConsistentAngles(double dx, double dy, double wavenumber) const
{
  const double d = dx > dy ? dy : dx;
  const double PI = std::acos(-1.0);
  const double qmax = PI / d;
  const double alphaMax = 2.0 * std::asin(qmax / wavenumber);
  if (alphaMax * 180.0 / PI < phiMax)
    throw std::runtime_error("The specified scattering angle is to large. Increase the spatial resolution!");
}

void post::FarField::result(const Solver &solver, arma::mat &res)
{
  if (sim == nullptr) throw std::runtime_error("FarField: no ParaxialSimulation linked");
  if (padding == Pad_t::LINEAR_FIT)
    std::clog << "Warning! Linear padding has not been tested properly and may not work!\n";
  const int pad_mode = padding == Pad_t::CONSTANT     ? PXG_PAD_CONSTANT
                       : padding == Pad_t::LINEAR_FIT ? PXG_PAD_LINEAR_FIT
                                                      : PXG_PAD_ZERO;
  const arma::cx_mat &solution = solver.getLastSolution3D();
  const pxg_plan plan = solver.devicePlan();
  if (plan == nullptr) throw std::runtime_error("FarField: the solver holds no field on the device");
  // the pad length grows to the size of the solution and of res (src/farFieldPost.cpp:146-153)
  signalLength = std::max<unsigned>(signalLength, solution.n_rows);
  signalLength = std::max<unsigned>(signalLength, solution.n_cols);
  const unsigned Nx = std::max<unsigned>(signalLength, res.n_rows);
  const unsigned Ny = std::max<unsigned>(signalLength, res.n_cols);
  if (Nx != Ny) throw std::runtime_error("FarField: only square matrices are supported");
  const unsigned er = resizeMatrices ? exportRows : 0, ec = resizeMatrices ? exportCols : 0;
  int dims[2] = {0, 0}, full[2] = {0, 0};
  check(pxg_far_field_dims(plan, 0, Nx, phiMin, phiMax, er, ec, dims));
  check(pxg_far_field_dims(plan, 0, Nx, phiMin, phiMax, 0, 0, full));
  if (static_cast<unsigned>(full[0]) == Nx)
    std::cout << "Warning! The requested scattering angle is beyond the maximum limit! Increase the resolution!\n";
  pxg_plan holder = nullptr;
  if (reference != nullptr) {
    if (reference->n_rows != solution.n_rows || reference->n_cols != solution.n_cols)
      throw std::runtime_error("FarField: the reference does not match the solution");
    // the reference field is held by a device plan of the same grid (no stepping)
    pxg_problem p{};
    p.kind = PXG_ADI;
    p.precision = PXG_PREC_FP64;
    p.nx = static_cast<int>(solution.n_cols);
    p.nz = 2;
    p.batch = 1;
    p.xmin = sim->transverseDiscretization().min;
    p.dx = sim->transverseDiscretization().step;
    p.ymin = sim->verticalDiscretization().min;
    p.dy = sim->verticalDiscretization().step;
    p.dz = 1.0;
    p.k = sim->getWavenumber();
    p.use_tbc = 1;
    check(pxg_plan_create(paxpro::context(), &p, &holder));
    const int rc = pxg_set_field(holder, 0, reinterpret_cast<const double *>(reference->memptr()));
    if (rc != PXG_OK) {
      pxg_plan_destroy(holder);
      check(rc);
    }
  }
  res.set_size(dims[0], dims[1]);
  const int rc = pxg_far_field_padded(plan, 0, holder, 0, Nx, pad_mode, phiMin, phiMax, er, ec, res.memptr());
  if (holder) pxg_plan_destroy(holder);
  check(rc);
}

// --------------------------------------------------------------- GenericScattering
GenericScattering::GenericScattering(const char *name) : ParaxialSimulation(name)
{
  gbeam.setDim(ParaxialSource::Dim_t::THREE_D);
  ff.linkParaxialSim(*this);
}

GenericScattering::~GenericScattering()
{
  delete reference;
  reference = nullptr;
}

void GenericScattering::setMaxScatteringAngle(double anglemax) { ff.setAngleRange(-anglemax, anglemax); }

// src/genericScattering.cpp:20-131
void GenericScattering::init()
{
  reset();
  isReferenceRun = true;
  ff.setPadLength(FFTPadLength);
  ff.setExportDimensions(exportNx, exportNy);
  setTransverseDiscretization(xmin, xmax, dx, downSampleX);
  setVerticalDiscretization(ymin, ymax, dy);
  setLongitudinalDiscretization(zmin, zmax, (zmax - zmin) / 3.0, 1);  // reference run
  setWaveLength(wavelength);
  gbeam.setWavelength(wavelength);
  if (realTimeVisualization || imgname != "")
    std::clog << "GenericScattering: visualisation and image export are not part of this build; ignored\n";
  switch (propagator) {
    case SolverType_t::ADI:
      setSolver(adisolver);
      break;
    case SolverType_t::FFT:
      setSolver(fft3Dsolver);
      break;
    case SolverType_t::PROJ:
      throw std::runtime_error("GenericScattering: the projection solver is not part of this build");
  }
  setBoundaryConditions(gbeam);
  ff.verifyConsistentAngles(dx, dy, wavenumber);
}

void GenericScattering::getXrayMatProp(double x, double y, double z, double &delta, double &beta) const
{
  if (isReferenceRun) {
    delta = 0.0;
    beta = 0.0;
    return;
  }
  material->getXrayMatProp(x, y, z, delta, beta);
}

// src/genericScattering.cpp:148-213
void GenericScattering::solve()
{
  if (material == nullptr) throw std::runtime_error("No material set!");
  init();
  printInfo();

  ParaxialSimulation::solve();  // reference run
  delete reference;
  reference = new arma::cx_mat(static_cast<const Solver &>(*solver).getLastSolution3D());
  if (subtract_reference) ff.setReference(*reference);
  std::clog << "Reference solution computed\n";

  reset();
  isReferenceRun = false;
  setLongitudinalDiscretization(zmin, zmax, dz, downSampleZ);
  solver->updateDimensionsOfArrays();
  setBoundaryConditions(gbeam);
  ParaxialSimulation::solve();
}

void GenericScattering::printInfo() const
{
  if (supressMessages) return;
  std::cout << "===================== SIMULATION INFO ===========================\n";
  std::cout << "Maximum size of matrices for export: Nx= " << exportNx << " Ny=" << exportNy << std::endl;
  std::cout << "FFT pad length: " << FFTPadLength << std::endl;
  std::cout << "Domain size: x=[" << xmin << "," << xmax << "], y=[" << ymin << "," << ymax << "], z=[" << zmin
            << "," << zmax << "]\n";
  std::cout << "Stepsize: dx=" << dx << ", dy=" << dy << ", dz=" << dz << std::endl;
  std::cout << "Downsample 3D solution by factor: downX=" << downSampleX << " downY=" << downSampleY
            << " downZ=" << downSampleZ << std::endl;
  std::cout << "Wavelength: " << wavelength << std::endl;
  std::cout << "Gaussian beam waist: " << gbeam.getWaist() << std::endl;
  std::cout << "Scattering angle: [" << ff.getMinScatteringAngle() << "," << ff.getMaxScatteringAngle() << "]\n";
  std::cout << "==================================================================\n";
}

void GenericScattering::getFarField(arma::mat &farF) { ff.result(*solver, farF); }

// ----------------------------------------------------------- closestIndex
// src/paraxialSimulation.cpp:215-219.  The reference converts a negative value to
// unsigned (undefined); here that throws.
void ParaxialSimulation::closestIndex(double x, double z, unsigned int &ix, unsigned int &iz) const
{
  const double vx = (x - xDisc.min) / xDisc.step, vz = (z - zDisc.min) / zDisc.step;
  if (!(vx >= 0.0) || !(vz >= 0.0) || vx >= 4294967296.0 || vz >= 4294967296.0)
    throw std::runtime_error("closestIndex: position outside the grid");
  ix = static_cast<unsigned int>(vx);
  iz = static_cast<unsigned int>(vz);
}

// ------------------------------------------------------- post::Transmittivity
// 0.5 (x_e - x_s)/(e - s) of trapezoidalIntegrateIntensityZ (src/transmittivity.cpp:30-42),
// with the reference's unsigned difference.
double post::Transmittivity::trapezoidScale(unsigned int ixStart, unsigned int ixEnd) const
{
  const double dx = (guide->getX(static_cast<int>(ixEnd)) - guide->getX(static_cast<int>(ixStart))) /
                    static_cast<double>(ixEnd - ixStart);
  return dx * 0.5;
}

double post::Transmittivity::trapezoidalIntegrateIntensityZ(unsigned int ixStart, unsigned int ixEnd) const
{
  const arma::cx_vec &u = guide->getSolver().getLastSolution();
  if (ixStart >= u.n_elem || ixEnd >= u.n_elem) throw std::runtime_error("Transmittivity: guide outside the grid");
  double integral = std::pow(std::abs(u(ixStart)), 2) + std::pow(std::abs(u(ixEnd)), 2);
  for (unsigned int ix = ixStart + 1; ix + 1 <= ixEnd; ++ix) integral += 2.0 * std::pow(std::abs(u(ix)), 2);
  return integral * trapezoidScale(ixStart, ixEnd);
}

// src/transmittivity.cpp:44-83 (no border tracker in this build)
void post::Transmittivity::compute(double z)
{
  if (guide == nullptr) throw std::runtime_error("Transmittivity: no waveguide linked");
  unsigned int wgStart = 0, wgEnd = 0, zIndx = 0;
  if (computeIntensityAtZero) {
    guide->closestIndex(0.0, 0.0, wgStart, zIndx);
    guide->closestIndex(guide->getWidth(), 0.0, wgEnd, zIndx);
    intensityAtZero = trapezoidalIntegrateIntensityZ(wgStart, wgEnd);
    computeIntensityAtZero = false;
  }
  guide->closestIndex(guide->waveGuideStartX(z), z, wgStart, zIndx);
  guide->closestIndex(guide->waveGuideEndX(z), z, wgEnd, zIndx);
  transmission.push_back(trapezoidalIntegrateIntensityZ(wgStart, wgEnd) / intensityAtZero);
}

void post::Transmittivity::computeSteps(unsigned int first, unsigned int n)
{
  if (guide == nullptr) throw std::runtime_error("Transmittivity: no waveguide linked");
  if (n == 0) return;
  const pxg_plan plan = guide->getSolver().devicePlan();
  if (plan == nullptr) throw std::runtime_error("Transmittivity: the solver holds no solution on the device");
  const bool zero = computeIntensityAtZero;
  const unsigned m = n + (zero ? 1 : 0);
  std::vector<int> steps(m), s(m), e(m);
  std::vector<unsigned> us(m), ue(m);
  unsigned int zIndx = 0;
  size_t k = 0;
  if (zero) {
    // the first compute() runs after step `first`, on that step's field
    guide->closestIndex(0.0, 0.0, us[0], zIndx);
    guide->closestIndex(guide->getWidth(), 0.0, ue[0], zIndx);
    steps[0] = static_cast<int>(first);
    k = 1;
  }
  for (unsigned i = 0; i < n; ++i, ++k) {
    const double z = guide->getZ(first + i);
    guide->closestIndex(guide->waveGuideStartX(z), z, us[k], zIndx);
    guide->closestIndex(guide->waveGuideEndX(z), z, ue[k], zIndx);
    steps[k] = static_cast<int>(first + i);
  }
  for (unsigned i = 0; i < m; ++i) {
    if (us[i] > 2147483647u || ue[i] > 2147483647u) throw std::runtime_error("Transmittivity: guide outside the grid");
    s[i] = static_cast<int>(us[i]);
    e[i] = static_cast<int>(ue[i]);
  }
  std::vector<double> sums(m);
  check(pxg_intensity_trapezoid(plan, 0, static_cast<int>(m), steps.data(), s.data(), e.data(), sums.data()));
  k = 0;
  if (zero) {
    intensityAtZero = sums[0] * trapezoidScale(us[0], ue[0]);
    computeIntensityAtZero = false;
    k = 1;
  }
  for (; k < m; ++k) transmission.push_back(sums[k] * trapezoidScale(us[k], ue[k]) / intensityAtZero);
}

// src/curvedWaveGuide2D.cpp:207-219: the march (one device call, then the
// longitudinal filter and downsampling as Solver2D::solve applies them), then the
// transmittivity after every step n = 1 .. Nz-1 from the stored columns.
void CurvedWaveGuideFD::solve()
{
  if (solver == nullptr) throw std::runtime_error("No solver specified!");
  const unsigned nz = nodeNumberLongitudinal();
  solver->solve();
  if (nz > 1) transmittivity.computeSteps(1, nz - 1);
}
