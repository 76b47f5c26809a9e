// Cladding (include/cladding.hpp of the reference): n = 1 - delta + i beta.
#pragma once

class Cladding {
 public:
  Cladding() {}
  void setRefractiveIndex(double d, double b) { delta = d; beta = b; }
  double getDelta() const { return delta; }
  double getBeta() const { return beta; }

 private:
  double delta{0.0}, beta{0.0};
};
