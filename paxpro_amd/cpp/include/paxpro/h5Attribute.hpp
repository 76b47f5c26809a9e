// H5Attr and makeAttr (include/h5Attribute.hpp, src/h5Attribute.cpp of the
// reference): a scalar attribute that a post-processing module attaches to its
// dataset in ParaxialSimulation::save.  The type is a tag rather than an
// H5::PredType, so this header needs no HDF5 headers.
#pragma once

#include <string>

struct H5Attr {
  enum class Type { NATIVE_DOUBLE, NATIVE_INT };
  std::string name{""};
  double value{0.0};
  Type dtype{Type::NATIVE_DOUBLE};
};

inline H5Attr makeAttr(const char *name, double value) { return H5Attr{name, value, H5Attr::Type::NATIVE_DOUBLE}; }
inline H5Attr makeAttr(const char *name, int value)
{
  return H5Attr{name, static_cast<double>(value), H5Attr::Type::NATIVE_INT};
}
