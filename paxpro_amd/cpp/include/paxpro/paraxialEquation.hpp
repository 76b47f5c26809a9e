// Coefficient hooks of the CN discretisation (include/paraxialEquation.hpp:16-35).
// The GPU path implements the default equation (F = G = H = 1, J = 0); a
// subclass with other coefficients is rejected when the plan is built.
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
