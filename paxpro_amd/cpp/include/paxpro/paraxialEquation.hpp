// Coefficient hooks of the CN discretisation (include/paraxialEquation.hpp:16-35):
// F(x,z) dA/dz = a G(x,z) d/dx (H(x,z) dA/dx) - K A + i k J(x,z) A.  A subclass
// with other coefficients is sampled on the host where CrankNicholson evaluates it
// and handed to the device as coefficient planes (pxg_set_equation_planes).
#pragma once

#include <complex>

typedef std::complex<double> cdouble;

class ParaxialEquation {
 public:
  ParaxialEquation() {}
  virtual ~ParaxialEquation() {}
  virtual double F(double x, double z) const { return 1.0; }
  virtual double G(double x, double z) const { return 1.0; }
  virtual double H(double x, double z) const { return 1.0; }
  virtual double J(double x, double z) const { return 0.0; }
  virtual cdouble phaseFactor(double k, double z) const { return std::exp(cdouble(0.0, 1.0) * k * z); }
};
