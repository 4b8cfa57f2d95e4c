// FlatDivider and SimpleDivider (Utils/include/SimpleDivider.h:30-151).
//  * FlatDivider: every node down to maxLevel.
//  * SimpleDivider: down to minLevel everywhere, then down to maxLevel only
//    the nodes touching the selected outer faces of the domain.
#ifndef BCM_FACADE_SIMPLEDIVIDER_H
#define BCM_FACADE_SIMPLEDIVIDER_H

#include "Divider.h"
#include "RootGrid.h"

class FlatDivider : public Divider {
  const RootGrid* rootGrid;
  int maxLevel;

 public:
  FlatDivider(const RootGrid* rootGrid_, int maxLevel_) : rootGrid(rootGrid_), maxLevel(maxLevel_) {}
  NodeType operator()(const Pedigree& pedigree) {
    return int(pedigree.getLevel()) >= maxLevel ? LEAF_ACTIVE : BRANCH;
  }
};

class SimpleDivider : public Divider {
 public:
  typedef unsigned char Direction;
  static const Direction NONE = 0;
  static const Direction XM = 1;
  static const Direction XP = 2;
  static const Direction YM = 4;
  static const Direction YP = 8;
  static const Direction ZM = 16;
  static const Direction ZP = 32;
  static const Direction ALL = XM | XP | YM | YP | ZM | ZP;

 private:
  const RootGrid* rootGrid;
  const int minLevel;
  const int maxLevel;
  Direction direction;

 public:
  SimpleDivider(const RootGrid* rootGrid_, int minLevel_, int maxLevel_, Direction direction_ = ALL)
      : rootGrid(rootGrid_), minLevel(minLevel_), maxLevel(maxLevel_), direction(direction_) {
    if (minLevel < 0 || maxLevel < minLevel) {
      std::cout << "*** error: SimpleDivider needs 0 <= minLevel <= maxLevel" << std::endl;
      Exit(EX_FAILURE);
    }
  }
  NodeType operator()(const Pedigree& pedigree) {
    const int level = int(pedigree.getLevel());
    if (level < minLevel) return BRANCH;
    if (level >= maxLevel) return LEAF_ACTIVE;
    const unsigned top = pedigree.getUpperBound() - 1;
    const unsigned c[3] = {pedigree.getX(), pedigree.getY(), pedigree.getZ()};
    const int root = int(pedigree.getRootID());
    for (int f = 0; f < NUM_FACE; ++f) {
      if (!(direction & (1u << f))) continue;
      const unsigned edge = (f & 1) ? top : 0u;
      if (c[f / 2] == edge && rootGrid->isOuterBoundary(root, Face(f))) return BRANCH;
    }
    return LEAF_ACTIVE;
  }
};

#endif
