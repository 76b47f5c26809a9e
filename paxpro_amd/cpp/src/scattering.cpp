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

void post::FarField::setExportDimensions(unsigned int rows, unsigned int cols)
{
  exportRows = rows;
  exportCols = cols;
  resizeMatrices = true;
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
  if (padding != Pad_t::ZERO)
    throw std::runtime_error("FarField: only zero padding is implemented on the device");
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
  const int rc = pxg_far_field(plan, 0, holder, 0, Nx, phiMin, phiMax, er, ec, res.memptr());
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
