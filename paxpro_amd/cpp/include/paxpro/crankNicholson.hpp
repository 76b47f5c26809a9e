// CrankNicholson (include/crankNicholson.hpp:16-43 of the reference).
#pragma once

#include "solver2D.hpp"
#include "thomasAlgorithm.hpp"

class CrankNicholson : public Solver2D {
 public:
  CrankNicholson() : Solver2D("CrankNicholson") {}
  ~CrankNicholson() {}
  virtual void solveStep(unsigned int iz) override final;

 protected:
  ThomasAlgorithm matrixSolver;
  int gpuKind() const override { return PXG_CN; }
};
