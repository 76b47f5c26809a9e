// FFTSolver2D (include/fftSolver2D.hpp:13-41 of the reference).
#pragma once

#include "solver2D.hpp"

class FFTSolver2D : public Solver2D {
 public:
  FFTSolver2D() : Solver2D("FFTSolver2D") {}
  virtual ~FFTSolver2D() {}
  void solveStep(unsigned int step) override;

 protected:
  int gpuKind() const override { return PXG_FFT2D; }
};
