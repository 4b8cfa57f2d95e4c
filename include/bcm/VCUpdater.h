// VCUpdater / Scalar3DUpdater<T> (include/VCUpdater.h, include/Scalar3DUpdater.h).
// In this build the updater is a tag: it marks a data class as exchangeable
// and records its neighbour table.  BlockManager::updateVC* performs every
// face's copy / F2C / C2F on the device in one launch (bcm_ghost_fill).
#ifndef BCM_FACADE_VCUPDATER_H
#define BCM_FACADE_VCUPDATER_H

#include "DataClass.h"
#include "NeighborInfo.h"

class VCUpdater {
 protected:
  const NeighborInfo* neighborInfo;
  DataClass* dataClass;

 public:
  explicit VCUpdater(const NeighborInfo* ni) : neighborInfo(ni), dataClass(0) {}
  virtual ~VCUpdater() {}
  virtual void setDataClass(DataClass* dc) { dataClass = dc; }
  const NeighborInfo* getNeighborInfo() const { return neighborInfo; }
};

inline UpdatableDataClass::~UpdatableDataClass() { delete updater; }

inline void UpdatableDataClass::setUpdater(VCUpdater* u) {
  updater = u;
  updater->setDataClass(this);
}

#endif
