// DataClass / UpdatableDataClass (include/DataClass.h:38-178): a per-block
// array with vc ghost layers; an updatable one carries the VCUpdater that
// exchanges its ghosts.
#ifndef BCM_FACADE_DATACLASS_H
#define BCM_FACADE_DATACLASS_H

#include <iostream>

#include "BCMTools.h"
#include "Vec3.h"

class VCUpdater;

class DataClass {
 protected:
  Vec3class::Vec3i size;
  int vc;

 public:
  DataClass(const Vec3class::Vec3i& size_, int vc_) : size(size_), vc(vc_) {
    if (size.x / 2 < vc || size.y / 2 < vc || size.z / 2 < vc) {
      std::cout << "*** error: cell size must be > 2*vc" << std::endl;
      Exit(EX_FAILURE);
    }
  }
  virtual ~DataClass() {}
  const Vec3class::Vec3i& getSize() const { return size; }
  int getSizeX() const { return size[0]; }
  int getSizeY() const { return size[1]; }
  int getSizeZ() const { return size[2]; }
  int getVCSize() const { return vc; }
};

class UpdatableDataClass : public DataClass {
  VCUpdater* updater;

 public:
  UpdatableDataClass(const Vec3class::Vec3i& size_, int vc_) : DataClass(size_, vc_), updater(0) {}
  virtual ~UpdatableDataClass();
  void setUpdater(VCUpdater* u);
  VCUpdater* getUpdater() const { return updater; }
};

#endif
