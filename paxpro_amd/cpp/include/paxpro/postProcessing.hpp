// post::PostProcessingModule and its two families (include/postProcessing.hpp:9-73,
// src/postProcessing.cpp:7-56 of the reference).  The modules of this build compute
// their results on the device from the solver's plan; addAttrib lists the
// attributes ParaxialSimulation::save attaches to a module's dataset.
#pragma once

#include <string>
#include <vector>

#include "arma_lite.hpp"
#include "h5Attribute.hpp"

class Solver;

namespace post {

class PostProcessingModule {
 public:
  enum class ReturnType_t { vector1D, matrix2D, cube3D };
  explicit PostProcessingModule(const char *name) : name(name) {}
  virtual ~PostProcessingModule() = default;

  /** Returns the result (each module implements the overloads of its kind) */
  virtual void result(const Solver &solver, arma::cube &res) {}
  virtual void result(const Solver &solver, arma::Cube<unsigned char> &res) {}
  virtual void result(const Solver &solver, arma::mat &res) {}
  virtual void result(const Solver &solver, arma::vec &res) {}

  std::string getName() const { return name; }

  /** Attributes of this module's dataset in the HDF5 file (none by default) */
  virtual void addAttrib(std::vector<H5Attr> &attrs) const {}

  /** Sets the size of exported matrices: block means down to rows x cols
   * (resizeMatrix, src/postProcessing.cpp:29-56), for the modules that resize. */
  void setExportDimensions(unsigned int rows, unsigned int cols)
  {
    exportRows = rows;
    exportCols = cols;
    resizeMatrices = true;
  }

  virtual ReturnType_t getReturnType(const Solver &solver) const = 0;

  /** If true, the result type is unsigned char */
  bool isUint8{false};

 protected:
  std::string name;
  unsigned int exportRows{0};
  unsigned int exportCols{0};
  bool resizeMatrices{false};
};

/** 2D for a 2D simulation, 3D for a 3D simulation (intensity, phase) */
class FieldQuantity : public PostProcessingModule {
 public:
  explicit FieldQuantity(const char *name) : PostProcessingModule(name) {}
  ReturnType_t getReturnType(const Solver &solver) const override final;
};

/** One dimension lower than the simulation (exit field, far field) */
class ProjectionQuantity : public PostProcessingModule {
 public:
  explicit ProjectionQuantity(const char *name) : PostProcessingModule(name) {}
  ReturnType_t getReturnType(const Solver &solver) const override final;
};

}  // namespace post
