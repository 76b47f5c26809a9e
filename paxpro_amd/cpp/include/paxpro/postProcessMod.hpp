// The post-processing modules (include/postProcessMod.hpp:9-194 of the reference),
// computed on the device from the solver's plan: the field quantities through
// pxg_field_quantity / pxg_field_uint8, the far field through pxg_far_field.
// Results are written into the caller's containers as the reference's are; saving
// them (HDF5) is not part of this build.
#pragma once

#include "arma_lite.hpp"
#include "postProcessing.hpp"

class Solver;
class ParaxialSimulation;

namespace post {

/** |u| of the stored solution (src/postProcessMod.cpp:7-15) */
class Intensity : public FieldQuantity {
 public:
  Intensity() : FieldQuantity("amplitude") {}
  using PostProcessingModule::result;
  /** 2D: |getSolution()| */
  void result(const Solver &solver, arma::mat &res) override;
  /** 3D: |getSolution3D()| */
  void result(const Solver &solver, arma::cube &res) override;
};

/** 255 |u| / max|u| of getSolution3D() as bytes (src/postProcessMod.cpp:17-30) */
class IntensityUint8 : public FieldQuantity {
 public:
  IntensityUint8() : FieldQuantity("amplitude") { isUint8 = true; }
  using PostProcessingModule::result;
  void result(const Solver &solver, arma::Cube<unsigned char> &res) override;
};

/** log|u| of getSolution3D() mapped to bytes (src/postProcessMod.cpp:32-88) */
class LogIntensityUint8 : public FieldQuantity {
 public:
  LogIntensityUint8() : FieldQuantity("LogAmplitude") { isUint8 = true; }
  using PostProcessingModule::result;
  void result(const Solver &solver, arma::Cube<unsigned char> &res) override;
  /** Values below this map to 0 (when above the solution's minimum) */
  void setMinValue(double minval)
  {
    minValue = minval;
    minvalSet = true;
  }
  /** Values above this map to 255 (when below the solution's maximum) */
  void setMaxValue(double maxval)
  {
    maxValue = maxval;
    maxvalSet = true;
  }

 private:
  double minValue{0.0}, maxValue{0.0};
  bool minvalSet{false}, maxvalSet{false};
};

/** arg u of the stored solution (src/postProcessMod.cpp:90-104) */
class Phase : public FieldQuantity {
 public:
  Phase() : FieldQuantity("phase") {}
  using PostProcessingModule::result;
  /** 2D: arg getSolution(), resized when export dimensions are set */
  void result(const Solver &solver, arma::mat &res) override;
  /** 3D: arg getSolution3D() */
  void result(const Solver &solver, arma::cube &res) override;
};

/** Far-field intensity of a 3D exit field (include/postProcessMod.hpp:67-139,
 * src/farFieldPost.cpp:141-230). */
class FarField : public ProjectionQuantity {
 public:
  enum class Pad_t { ZERO, LINEAR_FIT, CONSTANT };
  FarField() : ProjectionQuantity("farField") {}
  using PostProcessingModule::result;

  /** Padding of the signal: ZERO (the default), CONSTANT or LINEAR_FIT, all on the device. */
  void setPadding(Pad_t padType) { padding = padType; }
  /** Far field in the 3D case (src/farFieldPost.cpp:141-230): rows follow y, columns x. */
  void result(const Solver &solver, arma::mat &res) override;
  void setPadLength(unsigned int pad) { signalLength = pad; }
  /** Field subtracted before the transform (the vacuum run of GenericScattering). */
  void setReference(const arma::cx_mat &ref) { reference = &ref; }
  void setAngleRange(double angMin, double angMax);
  /** Throws when phiMax exceeds 2 asin(qmax/k), qmax = pi/min(dx, dy) (src/farFieldPost.cpp:320-333). */
  void verifyConsistentAngles(double dx, double dy, double wavenumber) const;
  void linkParaxialSim(const ParaxialSimulation &simulation) { sim = &simulation; }
  double getMaxScatteringAngle() const { return phiMax; }
  double getMinScatteringAngle() const { return phiMin; }
  /** phiMin, phiMax and qmin, qmax = phi k pi / 180 (src/farFieldPost.cpp:131-139) */
  void addAttrib(std::vector<H5Attr> &attrs) const override;

 private:
  unsigned int signalLength{0};
  double phiMin{-90.0};
  double phiMax{90.0};
  const ParaxialSimulation *sim{nullptr};
  Pad_t padding{Pad_t::ZERO};
  const arma::cx_mat *reference{nullptr};
};

/** Re u of the exit field (src/postProcessMod.cpp:106-120) */
class ExitField : public ProjectionQuantity {
 public:
  ExitField() : ProjectionQuantity("exitField") {}
  using PostProcessingModule::result;
  /** 2D: Re getLastSolution() */
  void result(const Solver &solver, arma::vec &res) override;
  /** 3D: Re getLastSolution3D(), resized when export dimensions are set */
  void result(const Solver &solver, arma::mat &res) override;
};

/** |u|^2 of the exit field (src/postProcessMod.cpp:122-136) */
class ExitIntensity : public ProjectionQuantity {
 public:
  ExitIntensity() : ProjectionQuantity("exitIntensity") {}
  using PostProcessingModule::result;
  void result(const Solver &solver, arma::vec &res) override;
  void result(const Solver &solver, arma::mat &res) override;
};

/** arg u of the exit field (src/postProcessMod.cpp:138-151) */
class ExitPhase : public ProjectionQuantity {
 public:
  ExitPhase() : ProjectionQuantity("exitPhase") {}
  using PostProcessingModule::result;
  void result(const Solver &solver, arma::vec &res) override;
  void result(const Solver &solver, arma::mat &res) override;
};

}  // namespace post
