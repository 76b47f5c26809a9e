// CSG parts and modules (include/geometry.hpp:7-107 of the reference).
#pragma once

#include <string>
#include <vector>

#include "shapes.hpp"

namespace geom {
enum class Operation_t { UNION, DIFFERENCE };

class Part {
 public:
  Part(const char *name) : name(name) {}
  ~Part();
  Part(const Part &other);
  Part &operator=(const Part &rhs) = delete;
  void add(Shape &shape);
  void difference(Shape &shape);
  void translate(double x, double y, double z);
  void rotate(double angle, Axis_t axis);
  void scale(double factor, Axis_t axis);
  bool isInside(double x, double y, double z) const;
  const std::string &getName() const { return name; }
  double delta{0.0};
  double beta{0.0};
  const std::vector<Shape *> &getShapes() const { return shapes; }
  const std::vector<Operation_t> &getOperations() const { return operations; }

 private:
  std::string name;
  std::vector<Shape *> shapes;
  std::vector<Operation_t> operations;
  bool ownShapeObjects{false};
};

class Module {
 public:
  Module(const char *name) : name(name) {}
  ~Module();
  Module(const Module &other);
  Module &operator=(const Module &rhs) = delete;
  void add(Part &newpart);
  void difference(Part &newpart);
  bool getXrayMatProp(double x, double y, double z, double &delta, double &beta) const;
  void translate(double x, double y, double z);
  void rotate(double angleDeg, Axis_t axis);
  void scale(double factor, Axis_t axis);
  const std::vector<Part *> &getParts() const { return parts; }
  const std::vector<Operation_t> &getOperations() const { return operations; }

 private:
  std::vector<Part *> parts;
  std::vector<Operation_t> operations;
  std::string name;
  bool ownPartObjects{false};
};
}  // namespace geom
