// Plane wave source (include/planeWave.hpp of the reference): unit amplitude,
// tilted by an angle in degrees to the z axis in the x-z plane.  Member
// definitions in src/simulation.cpp.
#pragma once

#include "paraxialSource.hpp"

class PlaneWave : public ParaxialSource {
 public:
  PlaneWave() : ParaxialSource("planeWave") {}
  cdouble get(double x, double z) const override final;
  cdouble get(double x, double y, double z) const override final;
  void setAngleDeg(double angle);

 private:
  double angleWithZAxis{0.0};
};
