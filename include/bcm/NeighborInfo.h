// Per-face neighbour record (include/NeighborInfo.h:36-228): neighbour
// block IDs and ranks per subface, the level difference (-1, 0, +1), the
// subface this block occupies on a coarser neighbour, and the outer-face
// flag.  One GPU holds every block, so a neighbour's rank is 0.
#ifndef BCM_FACADE_NEIGHBORINFO_H
#define BCM_FACADE_NEIGHBORINFO_H

#include <iostream>

#include "BCMTools.h"
#include "mpi.h"

class NeighborInfo {
  int neighborID[NUM_SUBFACE];
  int neighborRank[NUM_SUBFACE];
  int levelDifference;
  int neighborSubface;
  bool outerBoundary;

 public:
  NeighborInfo() : levelDifference(0), neighborSubface(0), outerBoundary(false) {
    for (int i = 0; i < NUM_SUBFACE; ++i) {
      neighborID[i] = -1;
      neighborRank[i] = MPI::PROC_NULL;
    }
  }
  void setLevelDifference(int d) { levelDifference = d; }
  int getLevelDifference() const { return levelDifference; }
  void setID(int id) { neighborID[0] = id; }
  void setID(Subface s, int id) { neighborID[s] = id; }
  int getID() const { return neighborID[0]; }
  int getID(Subface s) const { return neighborID[s]; }
  void setRank(int rank) { neighborRank[0] = rank; }
  void setRank(Subface s, int rank) { neighborRank[s] = rank; }
  int getRank() const { return neighborRank[0]; }
  int getRank(Subface s) const { return neighborRank[s]; }
  void setNeighborSubface(Subface s) { neighborSubface = s; }
  Subface getNeighborSubface() const { return Subface(neighborSubface); }
  void setOuterBoundary(bool f = true) { outerBoundary = f; }
  bool isOuterBoundary() const { return outerBoundary; }
  bool exists() const { return neighborID[0] >= 0; }
  void print() const {
    std::cout << "levelDIff=" << levelDifference << ", neighborSubface=" << neighborSubface << std::endl;
    std::cout << "  (ID,rank)=";
    for (int i = 0; i < NUM_SUBFACE; ++i) std::cout << "(" << neighborID[i] << "," << neighborRank[i] << ")";
    std::cout << std::endl;
  }
  // Child slot (x + 2y + 4z) of a finer neighbour that lies against my face
  // `face` at subface `subface` (include/NeighborInfo.h:176-193).
  static int getNeighborChildId(Face face, Subface subface) {
    static const int table[NUM_FACE][NUM_SUBFACE] = {
        {1, 3, 5, 7}, {0, 2, 4, 6}, {2, 6, 3, 7}, {0, 4, 1, 5}, {4, 5, 6, 7}, {0, 1, 2, 3}};
    return table[face][subface];
  }
  // Subface on face `face` occupied by child slot `childId`
  // (include/NeighborInfo.h:196-211).
  static Subface childIdToSubface(Face face, int childId) {
    static const int table[3][8] = {{0, 0, 1, 1, 2, 2, 3, 3}, {0, 2, 0, 2, 1, 3, 1, 3}, {0, 1, 2, 3, 0, 1, 2, 3}};
    return Subface(table[face / 2][childId]);
  }
  static Face reverseFace(Face f) { return Face(f ^ 1); }
};

#endif
