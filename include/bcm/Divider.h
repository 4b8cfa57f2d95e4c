// Divider (Utils/include/Divider.h:28-48): decides for each octree node,
// from its pedigree, whether to split it (BRANCH) or keep it as an active
// or inactive leaf.
#ifndef BCM_FACADE_DIVIDER_H
#define BCM_FACADE_DIVIDER_H

#include "Pedigree.h"
#include "bcm_hip.h"

class Divider {
 public:
  enum NodeType {
    BRANCH = BCM_NODE_BRANCH,
    LEAF_ACTIVE = BCM_NODE_LEAF_ACTIVE,
    LEAF_NO_ACTIVE = BCM_NODE_LEAF_INACTIVE,
  };
  Divider() {}
  virtual ~Divider() {}
  virtual NodeType operator()(const Pedigree& pedigree) = 0;
};

#endif
