// FFTSolver3D (include/fftSolver3D.hpp:16-99 of the reference): split-step
// Fourier propagation, two fused HBM passes per z-step on the GPU.
#pragma once

#include <string>
#include <vector>

#include "solver3D.hpp"

class FFTSolver3D : public Solver3D {
 public:
  FFTSolver3D() : Solver3D("FFTSolver3D") {}
  virtual ~FFTSolver3D() {}
  void solveStep(unsigned int step) override { Solver3D::solveStep(step); }
  void visualizeRealSpace();
  void visualizeFourierSpace();
  void setIntensityMinMax(double min, double max);
  void setPhaseMinMax(double min, double max);
  void overlayGeometry() {}
  void storeImages(const char *prefix);
  /** Exponential absorbing layer (src/fftSolver3D.cpp:327-349). */
  void absorbingBC(double width, double dampingLength);
  virtual void reset() override;

 protected:
  int gpuKind() const override { return PXG_FFT3D; }
  void configure(pxg_problem &p) const override;
  void samplePlanes(unsigned first, unsigned n, std::vector<double> &db) const override;
  unsigned absorbX{0}, absorbY{0};
  double invDampX{0.0}, invDampY{0.0};
  unsigned nStepsInRefrIntegral{10};
};

class FFT3DSolverDebug : public FFTSolver3D {
 public:
  virtual arma::cx_mat &getLastSolution3D() override { return *prevSolution; }
};
