// Solver2D (include/solver2D.hpp:17-131 of the reference).  The z-march runs on
// the GPU; the solution matrix (Nx x Nz) is stored on the device and fetched
// when the caller asks for it.
#pragma once

#include <complex>

#include "solver.hpp"

class ParaxialEquation;
typedef std::complex<double> cdouble;

class Solver2D : public Solver {
 public:
  enum class Comp_t { REAL, IMAG };
  enum class BC_t { DIRICHLET, TRANSPARENT };

  Solver2D(const char *name) : Solver(name, Dimension_t::TWO_D) {}
  virtual ~Solver2D();

  void setEquation(const ParaxialEquation &equation);
  void setSimulator(ParaxialSimulation &newGuide) override;
  const arma::cx_mat &getSolution(unsigned int iz) const { return getSolution(); }
  const arma::cx_mat &getSolution() const override;
  void realPart(double *realsolution) const;
  void imagPart(double *imagsolution) const;
  void getField(arma::mat &field) const;
  void getPhase(arma::mat &phase) const;
  void setInitialConditions(const arma::cx_vec &vec) override;
  void setXBC(const cdouble valuesTop[], const cdouble valuesBottom[]);
  virtual const arma::cx_vec &getLastSolution() const override final { return *prevSolution; }
  void solve() override;
  void step() override;
  void filterInLongitudinalDirection() override;
  void downSampleLongitudinalDirection() override;
  void setBoundaryCondition(BC_t bc) { boundaryCondition = bc; }
  void disableLongitudinalFilter() { longitudinalFilterDisabled = true; }
  void reset() override;

 protected:
  const ParaxialEquation *eq{nullptr};
  mutable arma::cx_mat *solution{nullptr};
  arma::cx_vec *prevSolution{nullptr};
  arma::cx_vec *currentSolution{nullptr};
  unsigned int Nx{0}, Nz{0};
  bool ownsParaxialEquationObject{true};
  double stepX{1.0}, stepZ{1.0}, xmin{0.0}, zmin{0.0}, wavenumber{1.0};
  BC_t boundaryCondition{BC_t::DIRICHLET};
  bool longitudinalFilterDisabled{false};
  mutable bool solutionStale{false};
  bool fieldDirty{true};

  void realOrImagPart(double *solution, Comp_t comp) const;
  void copyCurrentSolution(unsigned int step);
  virtual void solveStep(unsigned int step) = 0;
  void initValuesFromWaveGuide();
  void setLeftBC(const cdouble values[]);

  /** Solver kind in the C ABI (PXG_CN or PXG_FFT2D). */
  virtual int gpuKind() const = 0;
  /** Advances the device fields through steps first..first+n-1 and refreshes
   *  prev/currentSolution; samples host materials for the range when needed. */
  void runOnDevice(unsigned first, unsigned n);
  void ensurePlan();
  void fetchSolution() const;
};
