// ADI (include/alternatingDirectionSolver.hpp:15-43 of the reference):
// x-implicit then y-implicit half-steps, one tridiagonal line per lane.
#pragma once

#include <vector>

#include "solver3D.hpp"

class ADI : public Solver3D {
 public:
  ADI() : Solver3D("ADI") {}
  virtual void solveStep(unsigned int step) override { Solver3D::solveStep(step); }
  /** Transparent boundary conditions (default), otherwise Dirichlet. */
  bool useTBC{true};
  /** Physics-form switches (PXG_FIX_*); 0 reproduces the reference. */
  unsigned quirks{0};

 protected:
  int gpuKind() const override { return PXG_ADI; }
  void configure(pxg_problem &p) const override;
  void samplePlanes(unsigned first, unsigned n, std::vector<double> &db) const override;
  int planesPerStep() const override { return 2; }
};
