// Base of the 2D waveguide simulations (include/waveGuideFDSimulation.hpp of
// the reference): guide length, cladding and core materials.
#pragma once

#include "cladding.hpp"
#include "paraxialSimulation.hpp"
#include "transmittivity.hpp"

class WaveGuideFDSimulation : public ParaxialSimulation {
 public:
  WaveGuideFDSimulation() : ParaxialSimulation("noname") {}
  WaveGuideFDSimulation(const char *name) : ParaxialSimulation(name) {}
  void setWaveguideLength(double wgl) { wglength = wgl; }
  void setCladding(const Cladding &clad) { cladding = &clad; }
  void setInsideMaterial(const Cladding &mat) { insideMaterial = &mat; }
  const Cladding &getCladding() const { return *cladding; }
  virtual bool isInsideGuide(double x, double z) const { return true; }
  virtual void getXrayMatProp(double x, double z, double &delta, double &beta) const override;
  double getWaveguideLength() const { return wglength; }
  const Cladding *getInsideMaterial() const { return insideMaterial; }

 protected:
  double wglength{1E10};
  const Cladding *cladding{nullptr};
  const Cladding *insideMaterial{nullptr};
  virtual bool waveguideEnded(double x, double z) const { return z > wglength; }
};
