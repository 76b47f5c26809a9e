// post::FarField (include/postProcessMod.hpp:67-139 of the reference): the far-field
// intensity of a 3D exit field, computed on the device (pxg_far_field).  The other
// post-processing modules write HDF5 datasets and are not part of this build.
#pragma once

#include "arma_lite.hpp"

class Solver;
class ParaxialSimulation;

namespace post {

class FarField {
 public:
  enum class Pad_t { ZERO, LINEAR_FIT, CONSTANT };
  FarField() {}

  /** Padding of the signal; the device path implements ZERO (the default). */
  void setPadding(Pad_t padType) { padding = padType; }
  /** Far field in the 3D case (src/farFieldPost.cpp:141-230): rows follow y, columns x. */
  void result(const Solver &solver, arma::mat &res);
  void setPadLength(unsigned int pad) { signalLength = pad; }
  /** Field subtracted before the transform (the vacuum run of GenericScattering). */
  void setReference(const arma::cx_mat &ref) { reference = &ref; }
  void setAngleRange(double angMin, double angMax);
  /** Throws when phiMax exceeds 2 asin(qmax/k), qmax = pi/min(dx, dy) (src/farFieldPost.cpp:320-333). */
  void verifyConsistentAngles(double dx, double dy, double wavenumber) const;
  void linkParaxialSim(const ParaxialSimulation &simulation) { sim = &simulation; }
  double getMaxScatteringAngle() const { return phiMax; }
  double getMinScatteringAngle() const { return phiMin; }
  /** PostProcessingModule::setExportDimensions (src/postProcessing.cpp:51-56). */
  void setExportDimensions(unsigned int rows, unsigned int cols);

 private:
  unsigned int signalLength{0};
  double phiMin{-90.0};
  double phiMax{90.0};
  const ParaxialSimulation *sim{nullptr};
  Pad_t padding{Pad_t::ZERO};
  const arma::cx_mat *reference{nullptr};
  unsigned int exportRows{0}, exportCols{0};
  bool resizeMatrices{false};
};

}  // namespace post
