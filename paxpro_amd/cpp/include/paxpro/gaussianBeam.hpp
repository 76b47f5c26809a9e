// Gaussian beam source (include/gaussianBeam.hpp of the reference): waist w0,
// Rayleigh range pi w0^2 / lambda, Gouy phase and curvature radius at z,
// optional tilt about x and y.  Member definitions in src/simulation.cpp.
#pragma once

#include "paraxialSource.hpp"

class GaussianBeam : public ParaxialSource {
 public:
  GaussianBeam() : ParaxialSource("gaussianBeam") {}
  cdouble get(double x, double z) const override final;
  cdouble get(double x, double y, double z) const override final;
  void setWaist(double w) { waist = w; }
  void setCenter(double xc, double yc) { centerX = xc; centerY = yc; }
  void setAngleX(double ang) { angleX = ang; }
  void setAngleY(double ang) { angleY = ang; }
  double beamDivergence() const;
  double getWaist() const { return waist; }

 private:
  double waist{1.0}, centerX{0.0}, centerY{0.0}, angleX{0.0}, angleY{0.0};
  double rayleighRange() const;
  double guoyPhase(double z) const;
  double inverseRadiusOfCurvature(double z) const;
  double spotSize(double z) const;
};
