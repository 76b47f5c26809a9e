// Solver base class (include/solver.hpp:10-74 of the reference), backed by
// libpaxgpu: derived solvers advance their fields on the GPU.
#pragma once

#include <string>

#include "arma_lite.hpp"
#include "device.hpp"

class ParaxialSimulation;

/** Base class for all solvers */
class Solver {
 public:
  enum class Dimension_t { TWO_D, THREE_D };
  Solver(const char *name, Dimension_t dim);
  virtual ~Solver();
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  Dimension_t getDim() const { return dim; }
  std::string getName() const { return name; }
  virtual void setSimulator(ParaxialSimulation &newGuide) { guide = &newGuide; }
  const ParaxialSimulation &getSimulator() const { return *guide; }
  virtual void reset();

  virtual const arma::cx_cube &getSolution3D() const;
  virtual const arma::cx_mat &getSolution() const;
  virtual const arma::cx_vec &getLastSolution() const;
  virtual const arma::cx_mat &getLastSolution3D() const;
  virtual arma::cx_mat &getLastSolution3D();

  virtual void solve() {}
  virtual void step() = 0;
  virtual void setInitialConditions(const arma::cx_vec &) {}
  virtual void setInitialConditions(const arma::cx_mat &) {}
  virtual void filterInLongitudinalDirection() {}
  virtual void downSampleLongitudinalDirection() {}
  virtual void updateDimensionsOfArrays() {}

  /** GPU extension: field precision (PXG_PREC_FP64 or PXG_PREC_MIXED). */
  void setPrecision(int p);
  int getPrecision() const { return precision; }
  /** GPU extension: the device plan holding the current field (null before a run). */
  pxg_plan devicePlan() const { return plan; }

 protected:
  Dimension_t dim;
  std::string name;
  ParaxialSimulation *guide{nullptr};
  unsigned int currentStep{1};
  int precision;
  pxg_plan plan{nullptr};
  void destroyPlan();
};
