// BlockBase (include/BlockBase.h:40-138): a block's cell count, origin,
// size, level, neighbour table and registered data classes.
#ifndef BCM_FACADE_BLOCKBASE_H
#define BCM_FACADE_BLOCKBASE_H

#include <iostream>
#include <vector>

#include "BCMTools.h"
#include "DataClass.h"
#include "NeighborInfo.h"
#include "Vec3.h"

using namespace Vec3class;

class BlockBase {
  Vec3i size;
  Vec3r origin, blockSize, cellSize;
  int level;
  NeighborInfo* neighborInfo;
  std::vector<DataClass*> dataClassTable;

 public:
  BlockBase(const Vec3i& size_, const Vec3r& origin_, const Vec3r& blockSize_, int level_, NeighborInfo* ni)
      : size(size_), origin(origin_), blockSize(blockSize_), level(level_), neighborInfo(ni) {
    cellSize.x = blockSize.x / size.x;
    cellSize.y = blockSize.y / size.y;
    cellSize.z = blockSize.z / size.z;
    if (size.x % 2 != 0 || size.y % 2 != 0 || size.z % 2 != 0) {
      std::cout << "*** error: cell size must be even number" << std::endl;
      Exit(EX_FAILURE);
    }
  }
  virtual ~BlockBase() {
    delete[] neighborInfo;
    for (size_t i = 0; i < dataClassTable.size(); ++i) delete dataClassTable[i];
  }
  int setDataClass(DataClass* dc) {
    dataClassTable.push_back(dc);
    return int(dataClassTable.size()) - 1;
  }
  const Vec3i& getSize() const { return size; }
  const Vec3r& getOrigin() const { return origin; }
  const Vec3r& getBlockSize() const { return blockSize; }
  const Vec3r& getCellSize() const { return cellSize; }
  DataClass* getDataClass(int id) {
    if (id < 0 || id >= int(dataClassTable.size())) return 0;
    return dataClassTable[id];
  }
  int getLevel() const { return level; }
  const NeighborInfo* getNeighborInfo() const { return neighborInfo; }
  // Outer-face condition seen by the device kernels (INNER by default).
  virtual int boundaryType(int) const { return BCM_BC_INNER; }
  virtual int boundaryID(int) const { return -1; }

 private:
  BlockBase(const BlockBase&);
  BlockBase& operator=(const BlockBase&);
};

#endif
