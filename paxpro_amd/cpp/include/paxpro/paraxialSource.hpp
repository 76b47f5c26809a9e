// The source base class (include/paraxialSource.hpp of the reference); the
// Gaussian beam and the plane wave have headers of their own.
#pragma once

#include <complex>
#include <string>

typedef std::complex<double> cdouble;

class ParaxialSource {
 public:
  enum class Dim_t { TWO_D, THREE_D };
  ParaxialSource(const char *name) : name(name) {}
  virtual ~ParaxialSource() {}
  virtual cdouble get(double x, double z) const;
  virtual cdouble get(double x, double y, double z) const;
  void setWavenumber(double k) { wavenumber = k; }
  double getWavenumber() const { return wavenumber; }
  void setWavelength(double lambda);
  void setOrigin(double xOrig, double zOrig) { z0 = zOrig; }
  void setAmplitude(double amp) { amplitude = amp; }
  void setDim(Dim_t newdim) { dim = newdim; }
  Dim_t getDim() const { return dim; }
  std::string getName() const { return name; }
  double getWavelength() const;

 protected:
  double wavenumber{0.0};
  double z0{0.0};
  double x0{0.0};
  double amplitude{1.0};
  Dim_t dim{Dim_t::TWO_D};

 private:
  std::string name;
};
