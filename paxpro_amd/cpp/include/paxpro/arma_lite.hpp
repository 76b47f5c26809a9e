// Minimal column-major containers with the subset of Armadillo's interface the
// PaxPro solver API exposes (cx_vec, cx_mat, cx_cube, vec, mat): n_rows,
// n_cols, n_slices, n_elem, memptr(), element access and set_size.  Armadillo
// is not available to this build; the layout (column-major, interleaved
// complex<double>) is Armadillo's, so memptr() buffers map 1:1 to the C ABI.
#pragma once

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace arma {

using uword = std::size_t;

template <class T>
class Col;

template <class T>
class Mat {
 public:
  Mat() = default;
  Mat(uword r, uword c) { set_size(r, c); }
  Mat(uword r, uword c, const T &fill) : n_rows(r), n_cols(c), n_elem(r * c), v_(r * c, fill) {}
  void set_size(uword r, uword c)
  {
    n_rows = r;
    n_cols = c;
    n_elem = r * c;
    v_.assign(n_elem, T());
  }
  void zeros() { std::fill(v_.begin(), v_.end(), T()); }
  T *memptr() { return v_.data(); }
  const T *memptr() const { return v_.data(); }
  T &operator()(uword i) { return v_[i]; }
  const T &operator()(uword i) const { return v_[i]; }
  T &operator[](uword i) { return v_[i]; }
  const T &operator[](uword i) const { return v_[i]; }
  T &operator()(uword r, uword c) { return v_[c * n_rows + r]; }
  const T &operator()(uword r, uword c) const { return v_[c * n_rows + r]; }
  T &at(uword r, uword c)
  {
    if (r >= n_rows || c >= n_cols) throw std::out_of_range("Mat::at");
    return (*this)(r, c);
  }
  bool is_square() const { return n_rows == n_cols; }
  uword n_rows = 0, n_cols = 0, n_elem = 0;

 protected:
  std::vector<T> v_;
};

template <class T>
class Col : public Mat<T> {
 public:
  Col() = default;
  explicit Col(uword n) : Mat<T>(n, 1) {}
  Col(uword n, const T &fill) : Mat<T>(n, 1, fill) {}
  void set_size(uword n) { Mat<T>::set_size(n, 1); }
  using Mat<T>::operator();
};

template <class T>
class Cube {
 public:
  Cube() = default;
  Cube(uword r, uword c, uword s) { set_size(r, c, s); }
  void set_size(uword r, uword c, uword s)
  {
    n_rows = r;
    n_cols = c;
    n_slices = s;
    n_elem = r * c * s;
    v_.assign(n_elem, T());
  }
  T *memptr() { return v_.data(); }
  const T *memptr() const { return v_.data(); }
  T &operator()(uword r, uword c, uword s) { return v_[(s * n_cols + c) * n_rows + r]; }
  const T &operator()(uword r, uword c, uword s) const { return v_[(s * n_cols + c) * n_rows + r]; }
  uword n_rows = 0, n_cols = 0, n_slices = 0, n_elem = 0;

 private:
  std::vector<T> v_;
};

using cx_double = std::complex<double>;
using cx_vec = Col<cx_double>;
using cx_mat = Mat<cx_double>;
using cx_cube = Cube<cx_double>;
using vec = Col<double>;
using mat = Mat<double>;
using cube = Cube<double>;

}  // namespace arma
