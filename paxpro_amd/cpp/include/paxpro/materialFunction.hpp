// Materials (include/materialFunction.hpp:6-29 of the reference).  A CSGMaterial
// is evaluated on the device from its flattened shape/part/module arrays; any
// other MaterialFunction is sampled on the host through its virtual call.
#pragma once

#include <vector>

#include "geometry.hpp"
#include "paxgpu.h"

class MaterialFunction {
 public:
  virtual ~MaterialFunction() {}
  virtual void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const = 0;
};

class CSGMaterial : public MaterialFunction {
 public:
  CSGMaterial() {}
  virtual void getXrayMatProp(double x, double y, double z, double &delta, double &beta) const override;
  void addModule(const geom::Module &newmodule);
  double deltaSurrounding{0.0};
  double betaSurrounding{0.0};
  bool isReferenceRun{false};
  /** Flat descriptor in the layout of pxg_set_material_csg. */
  void flatten(std::vector<pxg_shape> &shapes, std::vector<pxg_part> &parts,
               std::vector<pxg_module> &modules) const;

 protected:
  std::vector<const geom::Module *> modules;
};
