// Per-face neighbour record (include/NeighborInfo.h:36-228), single GPU:
// the rank of a neighbour is 0 when it exists.
#ifndef BCM_FACADE_NEIGHBORINFO_H
#define BCM_FACADE_NEIGHBORINFO_H

#include "BCMTools.h"

class NeighborInfo {
  int neighborID[NUM_SUBFACE];
  int levelDifference;
  int neighborSubface;
  bool outerBoundary;

 public:
  NeighborInfo() : levelDifference(0), neighborSubface(0), outerBoundary(false) {
    for (int i = 0; i < NUM_SUBFACE; ++i) neighborID[i] = -1;
  }
  void setLevelDifference(int d) { levelDifference = d; }
  int getLevelDifference() const { return levelDifference; }
  void setID(int id) { neighborID[0] = id; }
  void setID(Subface s, int id) { neighborID[s] = id; }
  int getID() const { return neighborID[0]; }
  int getID(Subface s) const { return neighborID[s]; }
  int getRank() const { return neighborID[0] >= 0 ? 0 : -1; }
  int getRank(Subface s) const { return neighborID[s] >= 0 ? 0 : -1; }
  void setNeighborSubface(Subface s) { neighborSubface = s; }
  Subface getNeighborSubface() const { return Subface(neighborSubface); }
  void setOuterBoundary(bool f = true) { outerBoundary = f; }
  bool isOuterBoundary() const { return outerBoundary; }
  bool exists() const { return neighborID[0] >= 0; }
  static Face reverseFace(Face f) { return Face(f ^ 1); }
};

#endif
