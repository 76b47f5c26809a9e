// Problem definition of the drop-in API: discretisation, sources, CSG geometry
// and waveguide materials.  Behaviour follows the reference files cited per
// function; the solvers in solvers.cpp hand these to libpaxgpu.
#include <dlfcn.h>

#include <cmath>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "paxpro/h5writer.h"
#include "paxpro/paxpro.hpp"

namespace {
const double PI = std::acos(-1.0);
}

// ------------------------------------------------------------ ParaxialSimulation
ParaxialSimulation::ParaxialSimulation(const char *name) : name(name) {}
ParaxialSimulation::~ParaxialSimulation() {}

void ParaxialSimulation::solve() { solver->solve(); }
void ParaxialSimulation::step() { solver->step(); }
void ParaxialSimulation::reset()
{
  if (solver != nullptr) solver->reset();
}

// src/paraxialSimulation.cpp:68-120: k = 2 pi / lambda; node counts truncate.
void ParaxialSimulation::setWaveLength(double lambda) { wavenumber = 2.0 * PI / lambda; }
void ParaxialSimulation::setTransverseDiscretization(double a, double b, double s)
{
  setTransverseDiscretization(a, b, s, 1);
}
void ParaxialSimulation::setTransverseDiscretization(double a, double b, double s, unsigned r)
{
  xDisc = {a, b, s, r};
  yDisc.downsamplingRatio = r;
}
void ParaxialSimulation::setVerticalDiscretization(double a, double b, double s)
{
  yDisc.min = a;
  yDisc.max = b;
  yDisc.step = s;
}
void ParaxialSimulation::setLongitudinalDiscretization(double a, double b, double s)
{
  setLongitudinalDiscretization(a, b, s, 1);
}
void ParaxialSimulation::setLongitudinalDiscretization(double a, double b, double s, unsigned r)
{
  zDisc = {a, b, s, r};
}
unsigned ParaxialSimulation::nodeNumberTransverse() const { return (xDisc.max - xDisc.min) / xDisc.step + 1.0; }
unsigned ParaxialSimulation::nodeNumberVertical() const { return (yDisc.max - yDisc.min) / yDisc.step + 1.0; }
unsigned ParaxialSimulation::nodeNumberLongitudinal() const { return (zDisc.max - zDisc.min) / zDisc.step + 1.0; }

void ParaxialSimulation::setSolver(Solver &solv)
{
  solver = &solv;
  solver->setSimulator(*this);
}

double ParaxialSimulation::getZ(unsigned iz) const { return zDisc.min + iz * zDisc.step; }
double ParaxialSimulation::getX(int ix) const { return xDisc.min + ix * xDisc.step; }
double ParaxialSimulation::getY(int iy) const { return yDisc.min + iy * yDisc.step; }
double ParaxialSimulation::getWavelength() const { return 2.0 * PI / wavenumber; }
double ParaxialSimulation::getEnergy() const { return 12.398 * 100.0 / getWavelength(); }

// src/paraxialSimulation.cpp:356-399: sample the source at z = 0, then take k
// from the source.
void ParaxialSimulation::setBoundaryConditions(const ParaxialSource &source)
{
  if (solver == nullptr) throw std::runtime_error("Solver needs to be set before boundary conditions!");
  src = &source;
  const unsigned Nx = nodeNumberTransverse();
  if (source.getDim() == ParaxialSource::Dim_t::TWO_D) {
    arma::cx_vec values(Nx, cdouble(1.0, 0.0));
    for (unsigned i = 0; i < Nx; ++i) values[i] = src->get(getX(i), 0.0);
    solver->setInitialConditions(values);
  } else {
    const unsigned Ny = nodeNumberVertical();
    arma::cx_mat values(Nx, Ny);
#pragma omp parallel for
    for (long indx = 0; indx < static_cast<long>(Nx) * Ny; ++indx) {
      const unsigned i = indx / Ny, j = indx % Ny;
      values(j, i) = src->get(getX(i), getY(j), 0.0);
    }
    solver->setInitialConditions(values);
  }
  setWaveLength(source.getWavelength());
}

// src/paraxialSimulation.cpp:513-536
void ParaxialSimulation::getXrayMatProp(double x, double z, double &delta, double &beta) const
{
  if (material != nullptr) {
    material->getXrayMatProp(x, 0.0, z, delta, beta);
  } else {
    delta = 0.0;
    beta = 0.0;
  }
}
void ParaxialSimulation::getXrayMatProp(double x, double y, double z, double &delta, double &beta) const
{
  if (material != nullptr) {
    material->getXrayMatProp(x, y, z, delta, beta);
  } else {
    delta = 0.0;
    beta = 0.0;
  }
}

// src/paraxialSimulation.cpp:431-452
ParaxialSimulation &ParaxialSimulation::operator<<(post::PostProcessingModule &module)
{
  postProcess.push_back(&module);
  return *this;
}

ParaxialSimulation &ParaxialSimulation::operator<<(post::FarField &farfield)
{
  farfield.linkParaxialSim(*this);
  postProcess.push_back(&farfield);
  return *this;
}

namespace {

typedef int (*pxh5_write_fn)(const char *, const char *, const pxh5_group_attr *, int, const pxh5_dataset *,
                             int, char *, size_t);

// pxh5_write of libpaxpro_h5.so, looked up next to this library on first use
pxh5_write_fn h5_writer()
{
  static std::mutex mu;
  static pxh5_write_fn fn = nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (fn) return fn;
  Dl_info info{};
  std::string dir = ".";
  if (dladdr(reinterpret_cast<void *>(&h5_writer), &info) && info.dli_fname) {
    dir = info.dli_fname;
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? "." : dir.substr(0, slash);
  }
  const std::string path = dir + "/libpaxpro_h5.so";
  void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h)
    throw std::runtime_error("ParaxialSimulation::save: the HDF5 writer " + path +
                             " is not available (built only where the HDF5 C library is found): " +
                             std::string(dlerror()));
  fn = reinterpret_cast<pxh5_write_fn>(dlsym(h, "pxh5_write"));
  if (!fn) throw std::runtime_error("ParaxialSimulation::save: " + path + " has no pxh5_write");
  return fn;
}

}  // namespace

// src/paraxialSimulation.cpp:137-193 (the module loop), :454-511 (setGroupAttributes),
// :286-317 (saveArray) with the dataspaces of include/hdf5DataspaceCreator.hpp:30-66:
// a vector is {n_elem}, a matrix {n_cols, n_rows} and a cube {n_rows, n_cols, n_slices},
// each written from the column-major buffer as it is stored.  The 8-bit cubes go out as
// NATIVE_UINT8 without attributes (:180-182).
void ParaxialSimulation::save(const char *h5fname)
{
  if (solver == nullptr) throw std::runtime_error("No solver specified!\n");
  if (src == nullptr) throw std::runtime_error("No source specified!\n");
  const pxh5_write_fn write = h5_writer();

  std::vector<pxh5_group_attr> ga = {
      {"xmin", nullptr, xDisc.min}, {"xmax", nullptr, xDisc.max}, {"dx", nullptr, xDisc.step},
      {"zmin", nullptr, zDisc.min}, {"zmax", nullptr, zDisc.max}, {"dz", nullptr, zDisc.step},
      {"ymin", nullptr, yDisc.min}, {"ymax", nullptr, yDisc.max}, {"dy", nullptr, yDisc.step},
      {"wavenumber", nullptr, wavenumber}};
  const std::string solvername = solver->getName();
  ga.push_back({"solver", solvername.c_str(), 0.0});
  if (description != "") ga.push_back({"description", description.c_str(), 0.0});

  // results and attributes stay alive until the file is written
  const size_t nm = postProcess.size();
  std::vector<arma::vec> r1(nm);
  std::vector<arma::mat> r2(nm);
  std::vector<arma::cube> r3(nm);
  std::vector<arma::Cube<unsigned char>> r8(nm);
  std::vector<std::vector<H5Attr>> attrs(nm);
  std::vector<std::vector<pxh5_attr>> cattrs(nm);
  std::vector<std::string> names(nm);
  std::vector<pxh5_dataset> ds(nm);
  for (size_t i = 0; i < nm; ++i) {
    post::PostProcessingModule &m = *postProcess[i];
    names[i] = m.getName();
    pxh5_dataset &d = ds[i];
    d = pxh5_dataset{};
    d.name = names[i].c_str();
    m.addAttrib(attrs[i]);
    switch (m.getReturnType(*solver)) {
      case post::PostProcessingModule::ReturnType_t::vector1D:
        m.result(*solver, r1[i]);
        d.rank = 1;
        d.dims[0] = r1[i].n_elem;
        d.data = r1[i].memptr();
        break;
      case post::PostProcessingModule::ReturnType_t::matrix2D:
        m.result(*solver, r2[i]);
        d.rank = 2;
        d.dims[0] = r2[i].n_cols;
        d.dims[1] = r2[i].n_rows;
        d.data = r2[i].memptr();
        break;
      case post::PostProcessingModule::ReturnType_t::cube3D:
        d.rank = 3;
        if (m.isUint8) {
          m.result(*solver, r8[i]);
          d.dims[0] = r8[i].n_rows;
          d.dims[1] = r8[i].n_cols;
          d.dims[2] = r8[i].n_slices;
          d.uint8 = 1;
          d.data = r8[i].memptr();
          attrs[i].clear();
        } else {
          m.result(*solver, r3[i]);
          d.dims[0] = r3[i].n_rows;
          d.dims[1] = r3[i].n_cols;
          d.dims[2] = r3[i].n_slices;
          d.data = r3[i].memptr();
        }
        break;
    }
    for (const H5Attr &a : attrs[i])
      cattrs[i].push_back({a.name.c_str(), a.dtype == H5Attr::Type::NATIVE_INT, a.value});
    d.attrs = cattrs[i].data();
    d.n_attrs = static_cast<int>(cattrs[i].size());
  }
  std::string group = groupname;
  while (group.size() > 1 && group.back() == '/') group.pop_back();
  char err[512] = {0};
  if (write(h5fname, group.c_str(), ga.data(), static_cast<int>(ga.size()), ds.data(), static_cast<int>(ds.size()),
            err, sizeof err) != 0)
    throw std::runtime_error(std::string("ParaxialSimulation::save: ") + err);
  for (size_t i = 0; i < nm; ++i) std::clog << "Dataset " << names[i] << " added to HDF5 file\n";
}

// ------------------------------------------------------------------- sources
// src/paraxialSource.cpp, src/gaussianBeam.cpp:7-63, src/planeWave.cpp:5-14
void ParaxialSource::setWavelength(double lambda) { wavenumber = 2.0 * PI / lambda; }
double ParaxialSource::getWavelength() const { return 2.0 * PI / wavenumber; }
cdouble ParaxialSource::get(double, double) const
{
  throw std::runtime_error("2D version of the ParaxialSorce get function is not implemented!");
}
cdouble ParaxialSource::get(double, double, double) const
{
  throw std::runtime_error("3D version of the ParaxialSource get function is not implemented!");
}

double GaussianBeam::rayleighRange() const { return PI * waist * waist / getWavelength(); }
double GaussianBeam::guoyPhase(double z) const { return std::atan(z / rayleighRange()); }
double GaussianBeam::inverseRadiusOfCurvature(double z) const
{
  const double zR = rayleighRange();
  return z / (z * z + zR * zR);
}
double GaussianBeam::spotSize(double z) const
{
  const double zR = rayleighRange();
  return waist * std::sqrt(1.0 + std::pow(z / zR, 2));
}
cdouble GaussianBeam::get(double x, double z) const
{
  x -= centerX;
  const cdouble im(0.0, 1.0);
  const double waistRatio = waist / spotSize(z - z0);
  const double gaussianFactor = std::exp(-std::pow(x / spotSize(z - z0), 2));
  const cdouble phaseFactor = std::exp(-im * wavenumber * z0 +
                                       0.5 * im * wavenumber * x * x * inverseRadiusOfCurvature(z - z0) -
                                       im * guoyPhase(z - z0));
  const cdouble angX = std::exp(im * getWavenumber() * std::sin(angleX * PI / 180.0) * x);
  return amplitude * waistRatio * gaussianFactor * phaseFactor * angX;
}
cdouble GaussianBeam::get(double x, double y, double z) const
{
  y -= centerY;
  const double gaussY = std::exp(-std::pow(y / spotSize(z - z0), 2));
  const cdouble angY = std::exp(cdouble(0.0, 1.0) * getWavenumber() * std::sin(angleY * PI / 180.0) * y);
  return get(x, z) * gaussY * angY;
}
double GaussianBeam::beamDivergence() const { return getWavelength() / (PI * waist); }

cdouble PlaneWave::get(double x, double z) const
{
  const cdouble im(0.0, 1.0);
  return amplitude * std::exp(im * wavenumber * x * std::sin(angleWithZAxis) - im * wavenumber * z0);
}
cdouble PlaneWave::get(double x, double y, double z) const { return get(x, z); }
void PlaneWave::setAngleDeg(double angle) { angleWithZAxis = angle * PI / 180.0; }

// ------------------------------------------------------------------ geometry
// src/shape.cpp:10-70,126-187
geom::Shape::Shape(const char *n) : name(n), transformation(4, 4)
{
  for (int i = 0; i < 4; ++i) transformation(i, i) = 1.0;
}

static void mat4mul(const arma::mat &A, const arma::mat &B, arma::mat &C)
{
  arma::mat R(4, 4);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A(i, k) * B(k, j);
      R(i, j) = s;
    }
  C = R;
}

void geom::Shape::translate(double x, double y, double z)
{
  arma::mat M(4, 4);
  for (int i = 0; i < 4; ++i) M(i, i) = 1.0;
  M(0, 3) = -x;
  M(1, 3) = -y;
  M(2, 3) = -z;
  mat4mul(transformation, M, transformation);
}

void geom::Shape::rotate(double angleDeg, Axis_t axis)
{
  const double a = angleDeg * PI / 180.0;
  arma::mat M(4, 4);
  for (int i = 0; i < 4; ++i) M(i, i) = 1.0;
  const double c = std::cos(a), s = std::sin(a);
  if (axis == Axis_t::X) {
    M(1, 1) = c; M(1, 2) = s; M(2, 1) = -s; M(2, 2) = c;
  } else if (axis == Axis_t::Y) {
    M(0, 0) = c; M(0, 2) = -s; M(2, 0) = s; M(2, 2) = c;
  } else {
    M(0, 0) = c; M(0, 1) = s; M(1, 0) = -s; M(1, 1) = c;
  }
  mat4mul(transformation, M, transformation);
}

void geom::Shape::scale(double factor, Axis_t axis)
{
  arma::mat M(4, 4);
  for (int i = 0; i < 4; ++i) M(i, i) = 1.0;
  const int k = axis == Axis_t::X ? 0 : (axis == Axis_t::Y ? 1 : 2);
  M(k, k) = 1.0 / factor;
  mat4mul(M, transformation, transformation);
}

void geom::Shape::transform(double &x, double &y, double &z) const
{
  const double v[4] = {x, y, z, 1.0};
  double o[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 4; ++j) s += transformation(i, j) * v[j];
    o[i] = s;
  }
  x = o[0];
  y = o[1];
  z = o[2];
}

bool geom::Sphere::isInside(double x, double y, double z) const
{
  transform(x, y, z);
  return std::sqrt(x * x + y * y + z * z) < radius;
}
bool geom::Box::isInside(double x, double y, double z) const
{
  transform(x, y, z);
  return (x > -Lx / 2.0) && (x < Lx / 2.0) && (y > -Ly / 2.0) && (y < Ly / 2.0) && (z > -Lz / 2.0) &&
         (z < Lz / 2.0);
}
bool geom::Cylinder::isInside(double x, double y, double z) const
{
  transform(x, y, z);
  if (!((z < height / 2.0) && (z > -height / 2.0))) return false;
  const double radius = (r2 - r1) / height * (z + height / 2.0) + r1;
  return std::sqrt(x * x + y * y) < radius;
}

// src/geometry.cpp:8-108: parts own clones after a copy; the last shape that
// contains the point decides.
geom::Part::~Part()
{
  if (ownShapeObjects)
    for (Shape *s : shapes) delete s;
}
geom::Part::Part(const Part &o) : delta(o.delta), beta(o.beta), name(o.name), ownShapeObjects(true)
{
  for (size_t i = 0; i < o.shapes.size(); ++i) {
    shapes.push_back(o.shapes[i]->clone());
    operations.push_back(o.operations[i]);
  }
}
void geom::Part::add(Shape &shape)
{
  shapes.push_back(&shape);
  operations.push_back(Operation_t::UNION);
}
void geom::Part::difference(Shape &shape)
{
  if (shapes.empty())
    throw std::runtime_error(
        "There is no shape to make difference with. You have to add at least one shape before...");
  shapes.push_back(&shape);
  operations.push_back(Operation_t::DIFFERENCE);
}
bool geom::Part::isInside(double x, double y, double z) const
{
  for (size_t i = 0; i < shapes.size(); ++i) {
    const size_t k = shapes.size() - 1 - i;
    if (shapes[k]->isInside(x, y, z)) return operations[k] == Operation_t::UNION;
  }
  return false;
}
void geom::Part::translate(double x, double y, double z)
{
  for (Shape *s : shapes) s->translate(x, y, z);
}
void geom::Part::rotate(double a, Axis_t ax)
{
  for (Shape *s : shapes) s->rotate(a, ax);
}
void geom::Part::scale(double f, Axis_t ax)
{
  for (Shape *s : shapes) s->scale(f, ax);
}

// src/geometry.cpp:110-234
geom::Module::~Module()
{
  if (ownPartObjects)
    for (Part *p : parts) delete p;
}
geom::Module::Module(const Module &o) : name(o.name), ownPartObjects(true)
{
  for (size_t i = 0; i < o.parts.size(); ++i) {
    parts.push_back(new Part(*o.parts[i]));
    operations.push_back(o.operations[i]);
  }
}
void geom::Module::add(Part &p)
{
  parts.push_back(&p);
  operations.push_back(Operation_t::UNION);
}
void geom::Module::difference(Part &p)
{
  if (parts.size() <= 1)
    throw std::runtime_error("There have to be at least one part present before making difference!");
  parts.push_back(&p);
  operations.push_back(Operation_t::DIFFERENCE);
}
bool geom::Module::getXrayMatProp(double x, double y, double z, double &delta, double &beta) const
{
  delta = 0.0;
  beta = 0.0;
  for (size_t i = 0; i < parts.size(); ++i) {
    const size_t k = parts.size() - 1 - i;
    if (parts[k]->isInside(x, y, z)) {
      if (operations[k] == Operation_t::UNION) {
        delta = parts[k]->delta;
        beta = parts[k]->beta;
      }
      return true;
    }
  }
  return false;
}
void geom::Module::translate(double x, double y, double z)
{
  for (Part *p : parts) p->translate(x, y, z);
}
void geom::Module::rotate(double a, Axis_t ax)
{
  for (Part *p : parts) p->rotate(a, ax);
}
void geom::Module::scale(double f, Axis_t ax)
{
  for (Part *p : parts) p->scale(f, ax);
}

// src/materialFunction.cpp:3-13
void CSGMaterial::getXrayMatProp(double x, double y, double z, double &delta, double &beta) const
{
  delta = 0.0;
  beta = 0.0;
  if (isReferenceRun) return;
  for (const geom::Module *m : modules)
    if (m->getXrayMatProp(x, y, z, delta, beta)) return;
}
void CSGMaterial::addModule(const geom::Module &m) { modules.push_back(&m); }

void CSGMaterial::flatten(std::vector<pxg_shape> &shapes, std::vector<pxg_part> &parts,
                          std::vector<pxg_module> &mods) const
{
  shapes.clear();
  parts.clear();
  mods.clear();
  for (const geom::Module *m : modules) {
    pxg_module pm{static_cast<int>(parts.size()), static_cast<int>(m->getParts().size())};
    for (size_t ip = 0; ip < m->getParts().size(); ++ip) {
      const geom::Part *p = m->getParts()[ip];
      pxg_part pp{};
      pp.first_shape = static_cast<int>(shapes.size());
      pp.n_shapes = static_cast<int>(p->getShapes().size());
      pp.op = m->getOperations()[ip] == geom::Operation_t::UNION ? PXG_OP_UNION : PXG_OP_DIFFERENCE;
      pp.delta = p->delta;
      pp.beta = p->beta;
      for (size_t is = 0; is < p->getShapes().size(); ++is) {
        const geom::Shape *s = p->getShapes()[is];
        pxg_shape ps{};
        ps.type = s->typeId();
        ps.op = p->getOperations()[is] == geom::Operation_t::UNION ? PXG_OP_UNION : PXG_OP_DIFFERENCE;
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) ps.T[4 * r + c] = s->getTransformation()(r, c);
        s->params(ps.p);
        shapes.push_back(ps);
      }
      parts.push_back(pp);
    }
    mods.push_back(pm);
  }
}

// --------------------------------------------------------------- waveguides
// src/waveGuideFDSimulation.cpp:219-265
void WaveGuideFDSimulation::getXrayMatProp(double x, double z, double &delta, double &beta) const
{
  if (cladding == nullptr) throw std::runtime_error("WaveGuideFDSimulation: no cladding set");
  const double dx = xDisc.step, dz = zDisc.step;
  if (waveguideEnded(x, z)) {
    beta = 0.0;
    delta = 0.0;
    return;
  }
  const bool inside = isInsideGuide(x, z);
  const bool nb = isInsideGuide(x + dx, z) && isInsideGuide(x - dx, z) && isInsideGuide(x, z + dz) &&
                  isInsideGuide(x, z - dz);
  const double bIn = insideMaterial ? insideMaterial->getBeta() : 0.0;
  const double dIn = insideMaterial ? insideMaterial->getDelta() : 0.0;
  if (inside && nb) {
    beta = bIn;
    delta = dIn;
  } else if (inside) {
    delta = 0.5 * (cladding->getDelta() + dIn);
    beta = 0.5 * (cladding->getBeta() + bIn);
  } else {
    delta = cladding->getDelta();
    beta = cladding->getBeta();
  }
}

// src/curvedWaveGuide2D.cpp:49-52,179-205
bool CurvedWaveGuideFD::isInsideGuide(double x, double z) const
{
  return (2.0 * x * R + z * z > 0.0) && (2.0 * x * R + z * z < 2.0 * width * R);
}
double CurvedWaveGuideFD::smoothedWG(double x, double z) const
{
  const double w = 0.1 * width;
  const double d = 2.0 * R * x + z * z;
  const double f1 = 1.0 / (1.0 + std::exp(-d / w));
  const double f2 = 1.0 / (1.0 + std::exp((2 * R * width - d) / w));
  return 1.0 - (f1 - f2);
}
void CurvedWaveGuideFD::getXrayMatProp(double x, double z, double &delta, double &beta) const
{
  if (!useSmoothed) {
    WaveGuideFDSimulation::getXrayMatProp(x, z, delta, beta);
    return;
  }
  delta = cladding->getDelta() * smoothedWG(x, z);
  beta = cladding->getBeta() * smoothedWG(x, z);
}
