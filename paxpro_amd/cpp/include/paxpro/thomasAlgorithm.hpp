// ThomasAlgorithm (include/thomasAlgorithm.hpp:5-14 of the reference): solves
// a symmetric tridiagonal system (sup = sub) without pivoting on the GPU.  On
// return the solution is stored in diag and rhs is modified.
#pragma once

#include <complex>
#include <type_traits>
#include <vector>

#include "device.hpp"

class ThomasAlgorithm {
 public:
  ThomasAlgorithm() {}
  template <class T>
  void solve(T diag[], const T subdiag[], T rhs[], unsigned int N) const
  {
    static_assert(std::is_same<T, std::complex<double>>::value || std::is_same<T, double>::value,
                  "ThomasAlgorithm supports double and std::complex<double>");
    if (N == 0) return;
    std::vector<std::complex<double>> d(diag, diag + N), s(subdiag, subdiag + (N > 1 ? N - 1 : 0)),
        r(rhs, rhs + N);
    paxpro::check(pxg_thomas_batched(paxpro::context(), PXG_LAYOUT_CONTIGUOUS, 1, N,
                                     reinterpret_cast<double *>(d.data()),
                                     reinterpret_cast<const double *>(s.data()),
                                     reinterpret_cast<double *>(r.data())));
    for (unsigned i = 0; i < N; ++i) {
      assign(diag[i], d[i]);
      assign(rhs[i], r[i]);
    }
  }

 private:
  static void assign(std::complex<double> &dst, const std::complex<double> &v) { dst = v; }
  static void assign(double &dst, const std::complex<double> &v) { dst = v.real(); }
};
