// C entry point around the reference's own ThomasAlgorithm template, compiled
// straight from include/thomasAlgorithm.{hpp,tpp} under the reference tree by
// oracle/Makefile into oracle/_ref/.  Used only by tests to pin orc_thomas
// bit-exactly; it is never part of the product library.
#include <complex>
#include "thomasAlgorithm.hpp"

extern "C" void ref_thomas_c128(double *diag, const double *sub, double *rhs, unsigned n)
{
  ThomasAlgorithm solver;
  solver.solve(reinterpret_cast<std::complex<double> *>(diag),
               reinterpret_cast<const std::complex<double> *>(sub),
               reinterpret_cast<std::complex<double> *>(rhs), n);
}

extern "C" void ref_thomas_f64(double *diag, const double *sub, double *rhs, unsigned n)
{
  ThomasAlgorithm solver;
  solver.solve(diag, sub, rhs, n);
}
