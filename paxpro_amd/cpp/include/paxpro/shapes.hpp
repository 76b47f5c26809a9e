// CSG primitives (include/shapes.hpp:7-113 of the reference).  A shape keeps a
// 4x4 world->shape transform: translate and rotate right-multiply, scale
// left-multiplies; a point maps to shape coordinates as p' = T p.
#pragma once

#include <string>

#include "arma_lite.hpp"

namespace geom {
enum class Axis_t { X, Y, Z };

class Shape {
 public:
  Shape(const char *name);
  virtual ~Shape() {}
  virtual bool isInside(double x, double y, double z) const = 0;
  virtual Shape *clone() = 0;
  void translate(double x, double y, double z);
  void rotate(double angleDeg, Axis_t axis);
  void scale(double factor, Axis_t axis);
  void transform(double &x, double &y, double &z) const;
  const arma::mat &getTransformation() const { return transformation; }
  const std::string &getName() const { return name; }
  /** Shape type and parameters for the device descriptor (PXG_SHAPE_*). */
  virtual int typeId() const = 0;
  virtual void params(double p[3]) const = 0;

 protected:
  std::string name;
  arma::mat transformation;
};

class Sphere : public Shape {
 public:
  Sphere(double radius) : Shape("sphere"), radius(radius) {}
  bool isInside(double x, double y, double z) const override final;
  Shape *clone() override { return new Sphere(*this); }
  int typeId() const override { return 0; }
  void params(double p[3]) const override { p[0] = radius; p[1] = p[2] = 0.0; }

 protected:
  double radius{0.0};
};

class Box : public Shape {
 public:
  Box(double Lx, double Ly, double Lz) : Shape("Cube"), Lx(Lx), Ly(Ly), Lz(Lz) {}
  bool isInside(double x, double y, double z) const override final;
  Shape *clone() override { return new Box(*this); }
  int typeId() const override { return 1; }
  void params(double p[3]) const override { p[0] = Lx; p[1] = Ly; p[2] = Lz; }

 protected:
  double Lx, Ly, Lz;
};

class Cylinder : public Shape {
 public:
  Cylinder(double radius, double height) : Shape("Cylinder"), r1(radius), r2(radius), height(height) {}
  Cylinder(double r1, double r2, double height) : Shape("Cylinder"), r1(r1), r2(r2), height(height) {}
  bool isInside(double x, double y, double z) const override final;
  Shape *clone() override { return new Cylinder(*this); }
  int typeId() const override { return 2; }
  void params(double p[3]) const override { p[0] = r1; p[1] = r2; p[2] = height; }

 protected:
  double r1, r2, height;
};
}  // namespace geom
