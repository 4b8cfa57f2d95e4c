// BlockFactory (Utils/include/BlockFactory.h:26-100): a Block for a leaf
// node -- origin (tree origin x rootLength + rootOrigin), size, level, the
// driver's boundary conditions and the neighbour table.
#ifndef BCM_FACADE_BLOCKFACTORY_H
#define BCM_FACADE_BLOCKFACTORY_H

#include "BCMOctree.h"
#include "Block.h"
#include "BoundaryConditionSetterBase.h"
#include "Partition.h"

class BlockFactory {
  const BCMOctree* tree;
  const Partition* partition;
  BoundaryConditionSetterBase* boundaryConditionSetter;
  const Vec3r rootOrigin;
  REAL_TYPE rootLength;
  const Vec3i size;

 public:
  BlockFactory(const BCMOctree* tree_, const Partition* partition_, BoundaryConditionSetterBase* setter,
               const Vec3r& rootOrigin_, REAL_TYPE rootLength_, const Vec3i& size_)
      : tree(tree_), partition(partition_), boundaryConditionSetter(setter), rootOrigin(rootOrigin_),
        rootLength(rootLength_), size(size_) {}
  BlockFactory(const BCMOctree* tree_, const Partition* partition_, BoundaryConditionSetterBase* setter,
               const Vec3i& size_)
      : BlockFactory(tree_, partition_, setter, Vec3r(0.0, 0.0, 0.0), 1.0, size_) {}

  Block* makeBlock(const Node* node) {
    const Vec3r origin = tree->getOrigin(node) * rootLength + rootOrigin;
    const Vec3r blockSize = node->getBlockSize() * rootLength;
    BoundaryInfo* boundaryInfo = (*boundaryConditionSetter)(node, tree);
    NeighborInfo* neighborInfo = tree->makeNeighborInfo(node, partition);
    return new Block(size, origin, blockSize, node->getLevel(), neighborInfo, boundaryInfo);
  }
};

#endif
