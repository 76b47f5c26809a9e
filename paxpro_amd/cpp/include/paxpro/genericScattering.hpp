// GenericScattering (include/genericScattering.hpp:11-100 of the reference): a
// Gaussian beam through a material, a reference run without it, and the far field
// of the difference.  The propagation runs on the device (FFTSolver3D or ADI), the
// far field too (post::FarField).  Realtime visualisation, image export, the
// projection solver and the HDF5 post-processing modules are not part of this build.
#pragma once

#include <string>

#include "alternatingDirectionSolver.hpp"
#include "fftSolver3D.hpp"
#include "materialFunction.hpp"
#include "paraxialSimulation.hpp"
#include "paraxialSource.hpp"
#include "postProcessMod.hpp"

class GenericScattering : public ParaxialSimulation {
 public:
  enum class SolverType_t { ADI, FFT, PROJ };
  explicit GenericScattering(const char *name);
  virtual ~GenericScattering();

  void setMaterial(const MaterialFunction &mat) { material = &mat; }
  /** Set the waist of the Gaussian beam */
  void setBeamWaist(double waist) { gbeam.setWaist(waist); }
  /** Sets the maximum scattering angle that will be stored (degrees) */
  void setMaxScatteringAngle(double angMax);
  /** Reference run, then the run with the material (src/genericScattering.cpp:148-213) */
  virtual void solve() override;
  /** Zero during the reference run, the material's values otherwise */
  virtual void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const override;
  using ParaxialSimulation::getXrayMatProp;
  void printInfo() const;
  /** Returns the far field (squared) */
  void getFarField(arma::mat &farf);
  /** GPU extension: true while the reference (vacuum) run is propagated */
  bool isReference() const { return isReferenceRun; }

  unsigned int exportNx{512};
  unsigned int exportNy{512};
  unsigned int FFTPadLength{1024};
  double xmin{-1.0};
  double xmax{1.0};
  double ymin{-1.0};
  double ymax{1.0};
  double zmin{-1.0};
  double zmax{1.0};
  double dx{0.1};
  double dy{0.1};
  double dz{0.1};
  unsigned int downSampleX{1};
  unsigned int downSampleY{1};
  unsigned int downSampleZ{1};
  double wavelength{0.1569};
  double intensityMin{0.0};
  double intensityMax{1.0};
  double phaseMin{-3.14159};
  double phaseMax{3.14159};
  bool subtract_reference{true};
  std::string imgname{""};
  bool realTimeVisualization{false};
  SolverType_t propagator{SolverType_t::FFT};
  bool supressMessages{false};

 private:
  post::FarField ff;
  GaussianBeam gbeam;
  FFTSolver3D fft3Dsolver;
  ADI adisolver;
  arma::cx_mat *reference{nullptr};
  bool isReferenceRun{true};

  /** Has to be called before the simulation is solved */
  void init();
};
