// Solver3D (include/solver3D.hpp:11-73 of the reference).  Ny is forced to Nx;
// the block-averaged solution cube is stored on the device.
#pragma once

#include "solver.hpp"

class Solver3D : public Solver {
 public:
  Solver3D(const char *name) : Solver(name, Dimension_t::THREE_D) {}
  virtual ~Solver3D();

  void setSimulator(ParaxialSimulation &sim) override;
  void setInitialConditions(const arma::cx_mat &values) override;
  void setTimeBetweenStatus(unsigned int sec) { secBetweenStatusMessage = sec; }
  virtual void step() override;
  virtual void solve() override;
  virtual const arma::cx_cube &getSolution3D() const override;
  virtual const arma::cx_mat &getLastSolution3D() const override { return *prevSolution; }
  /** Interactive VISA windows are not part of this build: accepted, no effect. */
  void realTimeVisualization();
  void setPlotLimits(double intensityMin, double intensityMax, double phaseMin, double phaseMax);
  void setPlotLimits(double intensityMin, double intensityMax, double phaseMin, double phaseMax,
                     bool intensityLogScale);
  virtual void updateDimensionsOfArrays() override;
  void reset() override;

 protected:
  mutable arma::cx_cube *solution{nullptr};
  arma::cx_mat *currentSolution{nullptr};
  arma::cx_mat *prevSolution{nullptr};
  unsigned int Nx{0}, Ny{0}, Nz{0};
  unsigned int secBetweenStatusMessage{15};
  mutable bool solutionStale{false};
  bool fieldDirty{true};

  void copyCurrentSolution(unsigned int step);
  virtual void solveStep(unsigned int step);
  virtual int gpuKind() const = 0;
  /** Fills the problem fields specific to the solver (absorber, TBC, ...). */
  virtual void configure(pxg_problem &p) const {}
  /** Host sampling of a MaterialFunction the device cannot evaluate: planes
   *  for steps first..first+n-1 in the layout pxg_set_material_planes expects. */
  virtual void samplePlanes(unsigned first, unsigned n, std::vector<double> &db) const = 0;
  virtual int planesPerStep() const { return 1; }
  void runOnDevice(unsigned first, unsigned n);
  void ensurePlan();
};
