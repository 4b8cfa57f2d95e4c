#ifndef BCM_FACADE_SCALAR3DUPDATER_H
#define BCM_FACADE_SCALAR3DUPDATER_H

#include "Scalar3D.h"
#include "VCUpdater.h"

template <typename T>
class Scalar3DUpdater : public VCUpdater {
 public:
  explicit Scalar3DUpdater(const NeighborInfo* ni) : VCUpdater(ni) {}
};

#endif
