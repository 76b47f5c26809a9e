// Test driver of the C++ drop-in classes: builds simulations the way the
// reference's examples do (Examples/sphere.cpp, Examples/waveguide2D.cpp,
// Examples/circularWG.py) and writes the exit fields as raw complex128 so the
// Python tests can compare them with the CPU oracle.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "paxpro/paxpro.hpp"

static void write(const char *path, const cdouble *p, size_t n)
{
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), n * sizeof(cdouble));
}

static void write_raw(const std::string &path, const void *p, size_t bytes)
{
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), bytes);
}

// The post-processing modules of a 3D run: exit-field quantities (the phase resized
// to 20 x 30), the cube's intensity and phase, and the 8-bit images.
static void post3d(const Solver &solver, const std::string &out)
{
  post::ExitIntensity ei;
  post::ExitPhase ep;
  post::ExitField ef;
  post::Intensity in;
  post::Phase ph;
  post::IntensityUint8 u8;
  post::LogIntensityUint8 lu8;
  if (ei.getReturnType(solver) != post::PostProcessingModule::ReturnType_t::matrix2D ||
      in.getReturnType(solver) != post::PostProcessingModule::ReturnType_t::cube3D || !u8.isUint8 ||
      lu8.getName() != "LogAmplitude" || ep.getName() != "exitPhase")
    throw std::runtime_error("post module kinds");
  arma::mat m;
  ei.result(solver, m);
  write_raw(out + ".exitint", m.memptr(), m.n_elem * 8);
  ep.setExportDimensions(20, 30);
  ep.result(solver, m);
  write_raw(out + ".exitphase_rs", m.memptr(), m.n_elem * 8);
  ef.result(solver, m);
  write_raw(out + ".exitfield", m.memptr(), m.n_elem * 8);
  arma::cube c;
  in.result(solver, c);
  write_raw(out + ".intcube", c.memptr(), c.n_elem * 8);
  ph.result(solver, c);
  write_raw(out + ".phasecube", c.memptr(), c.n_elem * 8);
  arma::Cube<unsigned char> b;
  u8.result(solver, b);
  write_raw(out + ".u8", b.memptr(), b.n_elem);
  lu8.setMinValue(1e-3);
  lu8.setMaxValue(0.9);
  lu8.result(solver, b);
  write_raw(out + ".logu8", b.memptr(), b.n_elem);
  bool threw = false;
  try {
    arma::vec v;
    ei.result(solver, v);  // the 2D version on a 3D solver
  } catch (const std::runtime_error &) {
    threw = true;
  }
  if (!threw) throw std::runtime_error("2D result on a 3D solver did not throw");
}

// The modules of a 2D run: the solution's intensity and phase (resized to 40 x 7),
// and the exit quantities as vectors.
static void post2d(const Solver &solver, const std::string &out)
{
  post::Intensity in;
  post::Phase ph;
  post::ExitIntensity ei;
  post::ExitPhase ep;
  post::ExitField ef;
  if (in.getReturnType(solver) != post::PostProcessingModule::ReturnType_t::matrix2D ||
      ei.getReturnType(solver) != post::PostProcessingModule::ReturnType_t::vector1D)
    throw std::runtime_error("post module kinds");
  arma::mat m;
  in.result(solver, m);
  write_raw(out + ".int2d", m.memptr(), m.n_elem * 8);
  ph.setExportDimensions(40, 7);
  ph.result(solver, m);
  write_raw(out + ".phase2d_rs", m.memptr(), m.n_elem * 8);
  arma::vec v;
  ei.result(solver, v);
  write_raw(out + ".exitint", v.memptr(), v.n_elem * 8);
  ep.result(solver, v);
  write_raw(out + ".exitphase", v.memptr(), v.n_elem * 8);
  ef.result(solver, v);
  write_raw(out + ".exitfield", v.memptr(), v.n_elem * 8);
}

static int kats()
{
  struct K { int op; double a, b, c; double p[3]; double e[3]; } ks[] = {
      {0, 90.0, 2, 0, {1, 0, 0}, {0, -1, 0}},
      {0, 90.0, 1, 0, {1, 0, 0}, {0, 0, 1}},
      {0, 90.0, 0, 0, {0, 1, 0}, {0, 0, -1}},
      {1, 10.0, -10.0, -3.0, {1, -1, 10}, {-9, 9, 13}}};
  for (const K &k : ks) {
    geom::Sphere s(10.0);
    if (k.op == 0) s.rotate(k.a, static_cast<geom::Axis_t>(static_cast<int>(k.b)));
    else s.translate(k.a, k.b, k.c);
    double x = k.p[0], y = k.p[1], z = k.p[2];
    s.transform(x, y, z);
    if (std::fabs(x - k.e[0]) > 1e-6 || std::fabs(y - k.e[1]) > 1e-6 || std::fabs(z - k.e[2]) > 1e-6) {
      std::printf("KAT FAIL %f %f %f\n", x, y, z);
      return 1;
    }
  }
  // CSG membership through the Part/Module/CSGMaterial chain
  geom::Sphere sp(4.0);
  sp.translate(1.0, -2.0, 0.5);
  geom::Part part("p");
  part.add(sp);
  part.delta = 5e-6;
  part.beta = 1e-7;
  geom::Module mod("m");
  mod.add(part);
  CSGMaterial csg;
  csg.addModule(mod);
  double d, b;
  csg.getXrayMatProp(1.0, -2.0, 0.5, d, b);
  if (d != 5e-6 || b != 1e-7) return 1;
  csg.getXrayMatProp(6.0, -2.0, 0.5, d, b);
  if (d != 0.0 || b != 0.0) return 1;
  std::printf("KAT OK\n");
  return 0;
}

// A sphere as a user MaterialFunction (Examples/sphere.cpp:6-24): host-sampled.
class UserSphere : public MaterialFunction {
 public:
  explicit UserSphere(double r) : r(r) {}
  void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const override
  {
    if (x * x + y * y + z * z < r * r) {
      delta = 5e-6;
      beta = 1e-7;
    } else {
      delta = beta = 0.0;
    }
  }
  double r;
};

// Slab of Examples/waveguide2D.cpp:9-27 centred on 0.
class Slab : public MaterialFunction {
 public:
  void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const override
  {
    if (x < 50.0 && x > -50.0) {
      delta = 0.0;
      beta = 0.0;
    } else {
      delta = 4.5e-5;
      beta = 3.5e-6;
    }
  }
};

static int run3d(const std::string &mode, int n, int nsteps, const char *out, int prec)
{
  const double extent = 20000.0, dx = extent / n, xmin = -extent / 2, dz = 10.0;
  const double zmin = -nsteps * dz / 2;
  ParaxialSimulation sim("sphere");
  sim.setTransverseDiscretization(xmin, xmin + (n - 1) * dx, dx, 8);
  sim.setVerticalDiscretization(xmin, xmin + (n - 1) * dx, dx);
  sim.setLongitudinalDiscretization(zmin, zmin + nsteps * dz, dz, nsteps / 2);
  GaussianBeam gb;
  gb.setWaist(3000.0);
  gb.setWavelength(0.1);
  gb.setDim(ParaxialSource::Dim_t::THREE_D);
  geom::Sphere sphere(4000.0);
  geom::Part part("sphere");
  part.add(sphere);
  part.delta = 5e-6;
  part.beta = 1e-7;
  geom::Module mod("phantom");
  mod.add(part);
  CSGMaterial csg;
  csg.addModule(mod);
  UserSphere user(4000.0);
  FFTSolver3D fft;
  ADI adi;
  Solver3D *solver = mode.rfind("adi", 0) == 0 ? static_cast<Solver3D *>(&adi) : &fft;
  solver->setPrecision(prec);
  sim.setSolver(*solver);
  sim.material = (mode == "fft3d_user") ? static_cast<const MaterialFunction *>(&user) : &csg;
  sim.setBoundaryConditions(gb);
  if (mode == "fft3d_abs") fft.absorbingBC(600.0, 300.0);
  if (mode == "fft3d_reset" || mode == "adi_reset") {
    // a first propagation through another material, then ParaxialSimulation::reset
    // (src/paraxialSimulation.cpp:50-54: the solver's reset, currentStep back to 1),
    // new boundary conditions and the material of the checked run, as
    // GenericScattering::solve does between its two runs (src/genericScattering.cpp:148-213)
    UserSphere other(2500.0);
    sim.material = &other;
    sim.solve();
    sim.reset();
    sim.material = &csg;
    sim.setBoundaryConditions(gb);
  }
  if (mode == "fft3d_step") {
    for (int i = 1; i < n * 0 + nsteps + 1; ++i) sim.step();
  } else {
    sim.solve();
  }
  const arma::cx_mat &last = solver->getLastSolution3D();
  write(out, last.memptr(), last.n_elem);
  const arma::cx_cube &cube = solver->getSolution3D();
  write((std::string(out) + ".cube").c_str(), cube.memptr(), cube.n_elem);
  if (mode == "fft3d") {
    post3d(*solver, out);
    // post::FarField with constant padding (no reference), pad length 256, +-1 degree
    post::FarField ff;
    ff.linkParaxialSim(sim);
    ff.setPadding(post::FarField::Pad_t::CONSTANT);
    ff.setPadLength(256);
    ff.setAngleRange(-1.0, 1.0);
    arma::mat far;
    ff.result(*solver, far);
    write_raw(std::string(out) + ".ffconst", far.memptr(), far.n_elem * 8);
    std::printf("ffconst %zux%zu\n", (size_t)far.n_rows, (size_t)far.n_cols);
    // the same modules through ParaxialSimulation::save
    post::ExitIntensity ei;
    post::ExitPhase ep;
    ep.setExportDimensions(20, 30);
    post::Intensity in;
    post::LogIntensityUint8 lu8;  // IntensityUint8 shares Intensity's dataset name "amplitude"
    lu8.setMinValue(1e-3);
    lu8.setMaxValue(0.9);
    post::FarField ff2;
    ff2.setPadding(post::FarField::Pad_t::CONSTANT);
    ff2.setPadLength(256);
    ff2.setAngleRange(-1.0, 1.0);
    sim.description = "drop-in sphere";
    sim << ei << ep << in << lu8 << ff2;
    sim.save(std::string(out) + ".h5");
  }
  std::printf("%s n=%d steps=%d cube=%zux%zux%zu\n", mode.c_str(), n, nsteps, (size_t)cube.n_rows,
              (size_t)cube.n_cols, (size_t)cube.n_slices);
  return 0;
}

// A ParaxialEquation with coefficients varying in x and z (the same polynomials as
// the GPU test of the equation planes): F = 1 + 1e-4 x + 1e-7 z, G = 1 + 2e-4 x,
// H = 1 - 1e-4 x + 2e-8 z, J = 3e-6 + 1e-9 x.
struct PolyEquation : ParaxialEquation {
  double F(double x, double z) const override { return 1.0 + 1e-4 * x + 1e-7 * z; }
  double G(double x, double z) const override { return 1.0 + 2e-4 * x + 0.0 * z; }
  double H(double x, double z) const override { return 1.0 + -1e-4 * x + 2e-8 * z; }
  double J(double x, double z) const override { return 3e-6 + 1e-9 * x + 0.0 * z; }
};

static int run2d(const std::string &mode, int nx, int nsteps, const char *out)
{
  const double dz = 500.0;
  CurvedWaveGuideFD wg;
  ParaxialSimulation plain("wg2D");
  ParaxialSimulation &sim = mode == "cn_curved" ? static_cast<ParaxialSimulation &>(wg) : plain;
  double lo = -500.0, hi = 500.0;
  Cladding clad;
  clad.setRefractiveIndex(4.5e-5, 3.5e-6);
  if (mode == "cn_curved") {
    wg.setRadiusOfCurvature(5e6);
    wg.setWidth(120.0);
    wg.setCladding(clad);
    const double L = nsteps * dz;
    lo = -L * L / (2 * 5e6) - 2 * 120.0 - 200;
    hi = 120.0 + 200;
  }
  const double dx = (hi - lo) / (nx - 1);
  sim.setTransverseDiscretization(lo, hi, dx);
  sim.setLongitudinalDiscretization(0.0, nsteps * dz, dz);
  Slab slab;
  sim.material = mode == "cn_curved" ? nullptr : &slab;
  PlaneWave pw;
  pw.setWavelength(0.1);
  CrankNicholson cn;
  FFTSolver2D f2;
  Solver2D *solver = mode == "fft2d" ? static_cast<Solver2D *>(&f2) : &cn;
  if (mode != "fft2d") cn.setBoundaryCondition(Solver2D::BC_t::TRANSPARENT);
  PolyEquation poly;
  sim.setSolver(*solver);
  // after setSolver: Solver2D::setSimulator installs a default equation
  // (src/solver2D.cpp:41), so the reference's users set theirs afterwards
  if (mode == "cn_eq") cn.setEquation(poly);
  sim.setBoundaryConditions(pw);
  if (mode == "cn_step") {
    for (int i = 1; i <= nsteps; ++i) sim.step();
  } else {
    sim.solve();
  }
  const arma::cx_vec &last = solver->getLastSolution();
  write(out, last.memptr(), last.n_elem);
  const arma::cx_mat &sol = solver->getSolution();
  write((std::string(out) + ".sol").c_str(), sol.memptr(), sol.n_elem);
  if (mode == "cn") post2d(*solver, out);
  if (mode == "cn_curved") {
    // CurvedWaveGuideFD::solve also computed the transmittivity after every step
    std::vector<double> tr(1, wg.getTransmittivity().getIntensityAtZero());
    const std::vector<double> &t = wg.getTransmittivity().get();
    tr.insert(tr.end(), t.begin(), t.end());
    std::ofstream f((std::string(out) + ".trans").c_str(), std::ios::binary);
    f.write(reinterpret_cast<const char *>(tr.data()), tr.size() * sizeof(double));
  }
  std::printf("%s nx=%d steps=%d sol=%zux%zu\n", mode.c_str(), nx, nsteps, (size_t)sol.n_rows, (size_t)sol.n_cols);
  return 0;
}

// Examples/sphere.cpp with n points per side and nsteps z-steps: GenericScattering
// of a 500 nm sphere (a CSG part, or the example's own MaterialFunction subclass in
// "scatter_user"), far field within 0.05 degrees, written as raw doubles
// (column-major) with its shape printed.
static int scatter(const std::string &mode, int n, int nsteps, const char *out, int prec)
{
  paxpro::setDefaultPrecision(prec);  // before the simulation constructs its solvers
  const double r = 500.0;
  const double xmin = -1.5 * r, ext = 3.0 * r;
  const double dx = ext / n;
  const double zmin = -1.05 * r, zmax = 1.05 * r;
  geom::Sphere sphere(r);
  geom::Part part("sphere");
  part.add(sphere);
  part.delta = 8.9e-6;
  part.beta = 7e-7;
  geom::Module mod("phantom");
  mod.add(part);
  CSGMaterial csg;
  csg.addModule(mod);
  struct Ball : MaterialFunction {
    double rad;
    explicit Ball(double r) : rad(r) {}
    void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const override
    {
      if (x * x + y * y + z * z < rad * rad) {
        delta = 8.9e-6;
        beta = 7e-7;
        return;
      }
      delta = 0.0;
      beta = 0.0;
    }
  } ball(r);
  GenericScattering simulation("sphere");
  simulation.setBeamWaist(400.0 * r);
  simulation.setMaxScatteringAngle(0.05);
  simulation.xmin = xmin;
  simulation.xmax = xmin + (n - 1) * dx;
  simulation.ymin = xmin;
  simulation.ymax = xmin + (n - 1) * dx;
  simulation.zmin = zmin;
  simulation.zmax = zmax;
  simulation.dx = dx;
  simulation.dy = dx;
  simulation.dz = (zmax - zmin) / nsteps;
  simulation.downSampleX = n;
  simulation.downSampleY = n;
  simulation.downSampleZ = nsteps;
  simulation.wavelength = 0.1569;
  simulation.supressMessages = true;
  if (mode == "scatter_user")
    simulation.setMaterial(ball);
  else
    simulation.setMaterial(csg);
  if (mode == "scatter_adi") simulation.propagator = GenericScattering::SolverType_t::ADI;
  simulation.solve();
  arma::mat ff;
  simulation.getFarField(ff);
  std::ofstream f(out, std::ios::binary);
  f.write(reinterpret_cast<const char *>(ff.memptr()), ff.n_elem * sizeof(double));
  std::printf("%s n=%d steps=%d farfield=%zux%zu\n", mode.c_str(), n, nsteps, (size_t)ff.n_rows,
              (size_t)ff.n_cols);
  return 0;
}

// User-defined modules with known results (the reference lets any
// PostProcessingModule subclass be added): a vector, a matrix with attributes, a
// cube and an 8-bit cube, to check ParaxialSimulation::save without a device.
struct KnownVector : post::ProjectionQuantity {
  KnownVector() : post::ProjectionQuantity("knownVector") {}
  using post::PostProcessingModule::result;
  void result(const Solver &, arma::vec &res) override
  {
    res.set_size(5);
    for (unsigned i = 0; i < 5; ++i) res[i] = 0.5 * i - 1.0;
  }
};
struct KnownMatrix : post::ProjectionQuantity {
  KnownMatrix() : post::ProjectionQuantity("knownMatrix") {}
  using post::PostProcessingModule::result;
  void result(const Solver &, arma::mat &res) override
  {
    res.set_size(3, 4);  // element (r, c) = 10 r + c, column-major
    for (unsigned c = 0; c < 4; ++c)
      for (unsigned r = 0; r < 3; ++r) res(r, c) = 10.0 * r + c;
  }
  void addAttrib(std::vector<H5Attr> &attrs) const override
  {
    attrs.push_back(makeAttr("scale", 2.5));
    attrs.push_back(makeAttr("count", 7));
  }
};
struct KnownCube : post::FieldQuantity {
  KnownCube() : post::FieldQuantity("knownCube") {}
  using post::PostProcessingModule::result;
  void result(const Solver &, arma::cube &res) override
  {
    res.set_size(2, 3, 4);  // element (r, c, s) = 100 r + 10 c + s
    for (unsigned sl = 0; sl < 4; ++sl)
      for (unsigned c = 0; c < 3; ++c)
        for (unsigned r = 0; r < 2; ++r) res(r, c, sl) = 100.0 * r + 10.0 * c + sl;
  }
  void addAttrib(std::vector<H5Attr> &attrs) const override { attrs.push_back(makeAttr("slices", 4)); }
};
struct KnownBytes : post::FieldQuantity {
  KnownBytes() : post::FieldQuantity("knownBytes") { isUint8 = true; }
  using post::PostProcessingModule::result;
  void result(const Solver &, arma::Cube<unsigned char> &res) override
  {
    res.set_size(2, 2, 3);
    for (unsigned i = 0; i < res.n_elem; ++i) res.memptr()[i] = static_cast<unsigned char>(20 * i + 1);
  }
  // the reference writes 8-bit cubes without attributes, whatever the module lists
  void addAttrib(std::vector<H5Attr> &attrs) const override { attrs.push_back(makeAttr("dropped", 3)); }
};

// A field quantity of a 2D simulation is a matrix: the same values as KnownMatrix.
struct KnownField2D : post::FieldQuantity {
  KnownField2D() : post::FieldQuantity("knownField2D") {}
  using post::PostProcessingModule::result;
  void result(const Solver &s, arma::mat &res) override { KnownMatrix().result(s, res); }
};

// ParaxialSimulation::save of a 3D simulation (FFTSolver3D, Gaussian source, not
// propagated) and of a 2D one (CrankNicholson, plane wave) holding the known
// modules: no device work is needed.  A projection quantity is a matrix in 3D and
// a vector in 2D (ProjectionQuantity::getReturnType), a field quantity a cube in
// 3D and a matrix in 2D.
static int h5_known(const std::string &out)
{
  {
    ParaxialSimulation sim2("h5test2d");
    sim2.setTransverseDiscretization(-40.0, 40.0, 2.5);
    sim2.setLongitudinalDiscretization(0.0, 100.0, 10.0);
    PlaneWave pw;
    pw.setWavelength(0.1);
    CrankNicholson cn;
    sim2.setSolver(cn);
    sim2.setBoundaryConditions(pw);
    KnownVector v;
    KnownField2D f;
    KnownMatrix m;  // a projection quantity: a vector in 2D, and it has no vector result
    sim2 << v << f << m;
    sim2.save(out + ".2d.h5");
  }
  ParaxialSimulation sim("h5test");
  sim.setTransverseDiscretization(-100.0, 100.0, 12.5, 2);
  sim.setVerticalDiscretization(-100.0, 100.0, 12.5);
  sim.setLongitudinalDiscretization(0.0, 50.0, 5.0, 5);
  sim.description = "known modules";
  GaussianBeam gb;
  gb.setWaist(50.0);
  gb.setWavelength(0.2);
  gb.setDim(ParaxialSource::Dim_t::THREE_D);
  FFTSolver3D fft;
  sim.setSolver(fft);
  sim.setBoundaryConditions(gb);
  KnownMatrix m;
  KnownCube c;
  KnownBytes b;
  sim << m;
  sim.addPostProcessingModule(c).addPostProcessingModule(b);
  sim.save(out);
  std::printf("h5 saved %s\n", out.c_str());
  return 0;
}

int main(int argc, char **argv)
{
  try {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "kat") return kats();
    if (mode == "h5known" && argc >= 3) return h5_known(argv[2]);
    if (argc < 5) return 2;
    const int n = std::atoi(argv[2]), nsteps = std::atoi(argv[3]);
    const int prec = argc > 5 ? std::atoi(argv[5]) : PXG_PREC_FP64;
    if (mode.rfind("scatter", 0) == 0) return scatter(mode, n, nsteps, argv[4], prec);
    if (mode.rfind("fft3d", 0) == 0 || mode.rfind("adi", 0) == 0) return run3d(mode, n, nsteps, argv[4], prec);
    return run2d(mode, n, nsteps, argv[4]);
  } catch (const std::exception &e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
}
