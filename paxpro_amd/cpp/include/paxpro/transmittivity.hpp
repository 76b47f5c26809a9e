// post::Transmittivity (include/transmittivity.hpp:10-38 of the reference): the
// guided intensity after each z-step of a curved waveguide, normalised by the
// intensity across the guide's z = 0 extent.
#pragma once

#include <vector>

class CurvedWaveGuideFD;

namespace post {

class Transmittivity {
 public:
  /** Links a waveguide instance to the transmittivity computation */
  void linkWaveguide(const CurvedWaveGuideFD &wg) { guide = &wg; }
  /** One step, from the solver's current field (src/transmittivity.cpp:44-83) */
  void compute(double z);
  /** GPU extension: every step first..first+n-1 of a finished march at once, from the
   *  solution columns stored on the device (pxg_intensity_trapezoid); the same values
   *  compute(getZ(step)) gives after each of those steps */
  void computeSteps(unsigned int first, unsigned int n);
  /** Returns the result */
  const std::vector<double> &get() const { return transmission; }
  /** Returns the intensity at zero */
  double getIntensityAtZero() const { return intensityAtZero; }

 private:
  const CurvedWaveGuideFD *guide{nullptr};
  std::vector<double> transmission;
  double intensityAtZero{1E80};
  bool computeIntensityAtZero{true};

  double trapezoidalIntegrateIntensityZ(unsigned int ixStart, unsigned int ixEnd) const;
  double trapezoidScale(unsigned int ixStart, unsigned int ixEnd) const;
};

}  // namespace post
