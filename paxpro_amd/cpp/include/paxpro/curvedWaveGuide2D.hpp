// Curved 2D waveguide (include/curvedWaveGuide2D.hpp of the reference): a
// channel of width `width` bent on radius R, its left wall at x = -z^2/(2R).
// The guide's material is evaluated on the device, and so are the per-step
// transmittivity sums; border tracking, eigenmode projection and HDF5 output
// are not part of this build.
#pragma once

#include "waveGuideFDSimulation.hpp"

class CurvedWaveGuideFD : public WaveGuideFDSimulation {
 public:
  CurvedWaveGuideFD() : WaveGuideFDSimulation("CurvedWaveGuide2D") { transmittivity.linkWaveguide(*this); }
  /** Returns the transmittivity module */
  const post::Transmittivity &getTransmittivity() { return transmittivity; }
  /** The z-march, then the transmittivity of every step (src/curvedWaveGuide2D.cpp:207-219) */
  virtual void solve() override;
  void setRadiusOfCurvature(double newR) { R = newR; }
  void setWidth(double newWidth) { width = newWidth; }
  double getWidth() const { return width; }
  double getRadiusOfCurvature() const { return R; }
  double smoothedWG(double x, double z) const;
  void useSmoothedWG() { useSmoothed = true; }
  bool usesSmoothedWG() const { return useSmoothed; }
  virtual bool isInsideGuide(double x, double z) const override;
  virtual void getXrayMatProp(double x, double z, double &delta, double &beta) const override;
  virtual double waveGuideStartX(double z) const { return -0.5 * z * z / R; }
  virtual double waveGuideEndX(double z) const { return waveGuideStartX(z) + width; }

 protected:
  double R{1.0};
  double width{1.0};
  bool useSmoothed{false};
  post::Transmittivity transmittivity;
};
