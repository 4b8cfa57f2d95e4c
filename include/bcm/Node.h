// Node (Utils/include/Node.h:28-110): an active leaf of the octree as the
// facade's BCMOctree hands it out -- its pedigree and block ID.  The tree
// itself is built and kept by the library (bcm_tree_build_user), so a
// Node here has no parent or child links.
#ifndef BCM_FACADE_NODE_H
#define BCM_FACADE_NODE_H

#include "Pedigree.h"
#include "Vec3.h"

using namespace Vec3class;

class Node {
  int id;
  bool active;
  Pedigree pedigree;

 public:
  Node(const Pedigree& p, int blockID) : id(blockID), active(true), pedigree(p) {}
  bool isRootNode() const { return pedigree.getLevel() == 0; }
  bool isLeafNode() const { return true; }
  bool isActive() const { return active; }
  void setActive(bool on = true) { active = on; }
  int getBlockID() const { return id; }
  void setBlockID(int i) { id = i; }
  const Pedigree& getPedigree() const { return pedigree; }
  int getLevel() const { return int(pedigree.getLevel()); }
  // Edge of the node in root lengths: 1 / 2^level.
  Vec3r getBlockSize() const {
    const int ub = int(pedigree.getUpperBound());
    return Vec3r(1.0 / ub, 1.0 / ub, 1.0 / ub);
  }
};

#endif
