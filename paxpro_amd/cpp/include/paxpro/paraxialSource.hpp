// Sources (include/paraxialSource.hpp, gaussianBeam.hpp, planeWave.hpp of the
// reference).
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

class PlaneWave : public ParaxialSource {
 public:
  PlaneWave() : ParaxialSource("planeWave") {}
  cdouble get(double x, double z) const override final;
  cdouble get(double x, double y, double z) const override final;
  void setAngleDeg(double angle);

 private:
  double angleWithZAxis{0.0};
};
