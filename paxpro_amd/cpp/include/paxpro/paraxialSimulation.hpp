// ParaxialSimulation (include/paraxialSimulation.hpp:20-236 of the reference):
// grid, wavenumber, material dispatch, the initial field, the post-processing
// modules and their HDF5 output.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "arma_lite.hpp"
#include "materialFunction.hpp"

class Solver;
class ParaxialSource;
namespace post {
class PostProcessingModule;
class FarField;
}  // namespace post

struct Disctretization {
  double min{0.0};
  double max{0.0};
  double step{1.0};
  unsigned int downsamplingRatio{1};
};
typedef std::complex<double> cdouble;

class ParaxialSimulation {
 public:
  explicit ParaxialSimulation(const char *name);
  ParaxialSimulation(const ParaxialSimulation &) = delete;
  ParaxialSimulation &operator=(const ParaxialSimulation) = delete;
  virtual ~ParaxialSimulation();

  void setTransverseDiscretization(double xmin, double xmax, double step);
  void setTransverseDiscretization(double xmin, double xmax, double step, unsigned int ratio);
  void setLongitudinalDiscretization(double zmin, double zmax, double step);
  void setLongitudinalDiscretization(double zmin, double zmax, double step, unsigned int ratio);
  void setVerticalDiscretization(double ymin, double ymax, double step);
  unsigned int nodeNumberTransverse() const;
  unsigned int nodeNumberLongitudinal() const;
  unsigned int nodeNumberVertical() const;
  const Disctretization &transverseDiscretization() const { return xDisc; }
  const Disctretization &longitudinalDiscretization() const { return zDisc; }
  const Disctretization &verticalDiscretization() const { return yDisc; }
  double getWavenumber() const { return wavenumber; }
  double getWavelength() const;
  double getEnergy() const;
  void setWavenumber(double k) { wavenumber = k; }
  void setWaveLength(double lambda);
  void setSolver(Solver &solv);
  std::string getName() const { return name; }
  void step();
  /** Grid indices of (x, z), truncated (src/paraxialSimulation.cpp:215-219) */
  void closestIndex(double x, double z, unsigned int &ix, unsigned int &iz) const;
  double getZ(unsigned int iz) const;
  double getX(int ix) const;
  double getY(int iy) const;
  const Solver &getSolver() const { return *solver; }
  void reset();
  virtual void solve();
  virtual void setBoundaryConditions(const ParaxialSource &src);
  virtual void getXrayMatProp(double x, double z, double &delta, double &beta) const;
  virtual void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const;
  /** Writes every module's result to an HDF5 file (src/paraxialSimulation.cpp:137-193):
   * group /data/ with the grid, wavenumber, solver and description attributes, one
   * dataset /data/<module name> per module.  Needs libpaxpro_h5.so (built where the
   * HDF5 C library is found); throws std::runtime_error without it. */
  virtual void save(const char *fname);
  void save(const std::string &fname) { save(fname.c_str()); }
  /** Adds a post-processing module whose result save() writes (src/paraxialSimulation.cpp:431-452) */
  ParaxialSimulation &operator<<(post::PostProcessingModule &module);
  ParaxialSimulation &operator<<(post::FarField &farfield);
  ParaxialSimulation &addPostProcessingModule(post::PostProcessingModule &module) { return *this << module; }
  ParaxialSimulation &addPostProcessingModule(post::FarField &farfield) { return *this << farfield; }

  std::string description{""};
  const MaterialFunction *material{nullptr};

 protected:
  Solver *solver{nullptr};
  Disctretization xDisc, zDisc, yDisc;
  double wavenumber{1.0};
  std::string name;
  const ParaxialSource *src{nullptr};
  std::vector<post::PostProcessingModule *> postProcess;
  std::string groupname{"/data/"};
};
