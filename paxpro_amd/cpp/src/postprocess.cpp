// The field-quantity post-processing modules of the drop-in API
// (src/postProcessMod.cpp:7-151, src/postProcessing.cpp:7-27 of the reference).
// Each reads the field or the stored solution where the solver keeps it, on the
// device, and receives only the result.
#include <stdexcept>
#include <string>

#include "paxpro/paxpro.hpp"

using paxpro::check;
typedef post::PostProcessingModule::ReturnType_t ret_t;

post::PostProcessingModule::ReturnType_t post::FieldQuantity::getReturnType(const Solver &solver) const
{
  return solver.getDim() == Solver::Dimension_t::TWO_D ? ret_t::matrix2D : ret_t::cube3D;
}

post::PostProcessingModule::ReturnType_t post::ProjectionQuantity::getReturnType(const Solver &solver) const
{
  return solver.getDim() == Solver::Dimension_t::TWO_D ? ret_t::vector1D : ret_t::matrix2D;
}

namespace {

pxg_plan device_plan(const Solver &solver, const std::string &module)
{
  const pxg_plan plan = solver.devicePlan();
  if (plan == nullptr) throw std::runtime_error(module + ": the solver holds no field on the device");
  return plan;
}

void require_dim(const Solver &solver, Solver::Dimension_t dim, const std::string &what)
{
  // the reference's Solver base throws for the dimension a solver does not have
  if (solver.getDim() != dim) throw std::runtime_error("The " + what + " is not implemented!");
}

// quantity q of `source`, resized to er x ec when both are nonzero (matrices only)
void quantity_dims(pxg_plan plan, int source, unsigned er, unsigned ec, int d[3])
{
  check(pxg_field_quantity_dims(plan, 0, source, er, ec, d));
}

void quantity(const Solver &solver, const std::string &module, int source, int q, unsigned er,
              unsigned ec, arma::mat &res)
{
  const pxg_plan plan = device_plan(solver, module);
  int d[3];
  quantity_dims(plan, source, er, ec, d);
  res.set_size(d[0], d[1]);
  check(pxg_field_quantity(plan, 0, source, q, er, ec, res.memptr()));
}

void quantity(const Solver &solver, const std::string &module, int source, int q, arma::vec &res)
{
  const pxg_plan plan = device_plan(solver, module);
  int d[3];
  quantity_dims(plan, source, 0, 0, d);
  res.set_size(static_cast<arma::uword>(d[0]) * d[1]);
  check(pxg_field_quantity(plan, 0, source, q, 0, 0, res.memptr()));
}

void quantity(const Solver &solver, const std::string &module, int source, int q, arma::cube &res)
{
  const pxg_plan plan = device_plan(solver, module);
  int d[3];
  quantity_dims(plan, source, 0, 0, d);
  res.set_size(d[0], d[1], d[2]);
  check(pxg_field_quantity(plan, 0, source, q, 0, 0, res.memptr()));
}

void uint8_image(const Solver &solver, const std::string &module, int log_scale, const double *lo,
                 const double *hi, arma::Cube<unsigned char> &res)
{
  require_dim(solver, Solver::Dimension_t::THREE_D, "3D version of getSolution()");
  const pxg_plan plan = device_plan(solver, module);
  int d[3];
  quantity_dims(plan, PXG_SRC_SOLUTION, 0, 0, d);
  res.set_size(d[0], d[1], d[2]);
  check(pxg_field_uint8(plan, 0, PXG_SRC_SOLUTION, log_scale, lo, hi, res.memptr()));
}

}  // namespace

void post::Intensity::result(const Solver &solver, arma::mat &res)
{
  require_dim(solver, Solver::Dimension_t::TWO_D, "2D version of getSolution()");
  quantity(solver, name, PXG_SRC_SOLUTION, PXG_Q_ABS, 0, 0, res);
}

void post::Intensity::result(const Solver &solver, arma::cube &res)
{
  require_dim(solver, Solver::Dimension_t::THREE_D, "3D version of getSolution()");
  quantity(solver, name, PXG_SRC_SOLUTION, PXG_Q_ABS, res);
}

void post::IntensityUint8::result(const Solver &solver, arma::Cube<unsigned char> &res)
{
  uint8_image(solver, name, 0, nullptr, nullptr, res);
}

void post::LogIntensityUint8::result(const Solver &solver, arma::Cube<unsigned char> &res)
{
  uint8_image(solver, name, 1, minvalSet ? &minValue : nullptr, maxvalSet ? &maxValue : nullptr, res);
}

void post::Phase::result(const Solver &solver, arma::mat &res)
{
  require_dim(solver, Solver::Dimension_t::TWO_D, "2D version of getSolution()");
  const unsigned er = resizeMatrices ? exportRows : 0, ec = resizeMatrices ? exportCols : 0;
  quantity(solver, name, PXG_SRC_SOLUTION, PXG_Q_ARG, er, ec, res);
}

void post::Phase::result(const Solver &solver, arma::cube &res)
{
  require_dim(solver, Solver::Dimension_t::THREE_D, "3D version of getSolution()");
  quantity(solver, name, PXG_SRC_SOLUTION, PXG_Q_ARG, res);
}

namespace {

void exit_vec(const Solver &solver, const std::string &module, int q, arma::vec &res)
{
  require_dim(solver, Solver::Dimension_t::TWO_D, "2D version of getLastSolution()");
  quantity(solver, module, PXG_SRC_FIELD, q, res);
}

void exit_mat(const Solver &solver, const std::string &module, int q, bool resize, unsigned er,
              unsigned ec, arma::mat &res)
{
  require_dim(solver, Solver::Dimension_t::THREE_D, "3D version of getLastSolution3D()");
  quantity(solver, module, PXG_SRC_FIELD, q, resize ? er : 0, resize ? ec : 0, res);
}

}  // namespace

void post::ExitField::result(const Solver &solver, arma::vec &res) { exit_vec(solver, name, PXG_Q_REAL, res); }
void post::ExitField::result(const Solver &solver, arma::mat &res)
{
  exit_mat(solver, name, PXG_Q_REAL, resizeMatrices, exportRows, exportCols, res);
}
void post::ExitIntensity::result(const Solver &solver, arma::vec &res) { exit_vec(solver, name, PXG_Q_ABS2, res); }
void post::ExitIntensity::result(const Solver &solver, arma::mat &res)
{
  exit_mat(solver, name, PXG_Q_ABS2, resizeMatrices, exportRows, exportCols, res);
}
void post::ExitPhase::result(const Solver &solver, arma::vec &res) { exit_vec(solver, name, PXG_Q_ARG, res); }
void post::ExitPhase::result(const Solver &solver, arma::mat &res)
{
  exit_mat(solver, name, PXG_Q_ARG, resizeMatrices, exportRows, exportCols, res);
}
