// BoundaryConditionSetterBase (Utils/include/BoundaryConditionSetterBase.h):
// a driver's setter returns a new BoundaryInfo[NUM_FACE] for a leaf node;
// the Block made from the node takes ownership.
#ifndef BCM_FACADE_BOUNDARYCONDITIONSETTERBASE_H
#define BCM_FACADE_BOUNDARYCONDITIONSETTERBASE_H

class BoundaryInfo;
class Node;
class BCMOctree;

class BoundaryConditionSetterBase {
 public:
  BoundaryConditionSetterBase() {}
  virtual ~BoundaryConditionSetterBase() {}
  virtual BoundaryInfo* operator()(const Node* node, const BCMOctree* tree) = 0;
};

#endif
