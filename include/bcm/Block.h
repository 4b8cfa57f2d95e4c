// Block (Utils/include/Block.h:25-50): a BlockBase with the boundary
// condition of each face; the device block table takes the face types and
// boundary-value IDs from here.
#ifndef BCM_FACADE_BLOCK_H
#define BCM_FACADE_BLOCK_H

#include "BlockBase.h"
#include "BoundaryInfo.h"

class Block : public BlockBase {
  BoundaryInfo* boundaryInfo;

 public:
  Block(const Vec3i& size_, const Vec3r& origin_, const Vec3r& blockSize_, int level_, NeighborInfo* ni,
        BoundaryInfo* bi)
      : BlockBase(size_, origin_, blockSize_, level_, ni), boundaryInfo(bi) {}
  ~Block() { delete[] boundaryInfo; }
  const BoundaryInfo* getBoundaryInfo() const { return boundaryInfo; }
  int boundaryType(int f) const { return boundaryInfo ? int(boundaryInfo[f].getType()) : BCM_BC_INNER; }
  int boundaryID(int f) const { return boundaryInfo ? boundaryInfo[f].getID() : -1; }
};

#endif
