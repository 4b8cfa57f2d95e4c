// RootGrid (Utils/include/RootGrid.h:35-230): nx x ny x nz root blocks,
// root ID = ix + nx (iy + ny iz), with periodic flags per axis.
#ifndef BCM_FACADE_ROOTGRID_H
#define BCM_FACADE_ROOTGRID_H

#include "BCMTools.h"
#include "Vec3.h"

using namespace Vec3class;

class RootGrid {
  int n[3];
  bool periodic[3];

 public:
  RootGrid(int nx, int ny, int nz) {
    n[0] = nx;
    n[1] = ny;
    n[2] = nz;
    periodic[0] = periodic[1] = periodic[2] = false;
  }
  explicit RootGrid(const Vec3i& v) : RootGrid(v.x, v.y, v.z) {}
  int getSize() const { return n[0] * n[1] * n[2]; }
  int getSizeX() const { return n[0]; }
  int getSizeY() const { return n[1]; }
  int getSizeZ() const { return n[2]; }
  void setPeriodicX() { periodic[0] = true; }
  void setPeriodicY() { periodic[1] = true; }
  void setPeriodicZ() { periodic[2] = true; }
  void clearPeriodicX() { periodic[0] = false; }
  void clearPeriodicY() { periodic[1] = false; }
  void clearPeriodicZ() { periodic[2] = false; }
  bool isPeriodicX() const { return periodic[0]; }
  bool isPeriodicY() const { return periodic[1]; }
  bool isPeriodicZ() const { return periodic[2]; }
  int rootID2indexX(int rootID) const { return rootID % n[0]; }
  int rootID2indexY(int rootID) const { return (rootID / n[0]) % n[1]; }
  int rootID2indexZ(int rootID) const { return (rootID / n[0]) / n[1]; }
  int index2rootID(int ix, int iy, int iz) const { return ix + n[0] * (iy + n[1] * iz); }
  // The root across `face`, wrapping on a periodic axis; -1 outside.
  int getNeighborRoot(int rootID, Face face) const {
    int i[3] = {rootID2indexX(rootID), rootID2indexY(rootID), rootID2indexZ(rootID)};
    const int a = face / 2;
    i[a] += (face & 1) ? 1 : -1;
    if (i[a] < 0 || i[a] >= n[a]) {
      if (!periodic[a]) return -1;
      i[a] = (i[a] + n[a]) % n[a];
    }
    return index2rootID(i[0], i[1], i[2]);
  }
  // Whether the root lies on the domain boundary at `face` (by position,
  // periodic or not).
  bool isOuterBoundary(int rootID, Face face) const {
    const int i[3] = {rootID2indexX(rootID), rootID2indexY(rootID), rootID2indexZ(rootID)};
    const int a = face / 2;
    return (face & 1) ? i[a] == n[a] - 1 : i[a] == 0;
  }
};

#endif
