// BoundaryInfo (Utils/include/BoundaryInfo.h:26-95): the condition type of
// one block face and the index of its boundary value.
#ifndef BCM_FACADE_BOUNDARYINFO_H
#define BCM_FACADE_BOUNDARYINFO_H

#include <iostream>

#include "BCMTools.h"

class BoundaryInfo {
 public:
  enum Type { INNER = BCM_BC_INNER, DIRICHLET = BCM_BC_DIRICHLET, NEUMANN = BCM_BC_NEUMANN, PERIODIC = BCM_BC_PERIODIC };

 private:
  Type type;
  int id;

 public:
  BoundaryInfo() : type(INNER), id(-1) {}
  void setType(Type t) { type = t; }
  void setID(int i) { id = i; }
  Type getType() const { return type; }
  int getID() const { return id; }
  void print() const {
    static const char* names[] = {"INNER", "DIRICHLET", "NEUMANN", "PERIODIC"};
    std::cout << "type=" << names[type] << ", id=" << id << std::endl;
  }
  static void print(const BoundaryInfo* bInfo) {
    static const char* faces[] = {"X_M", "X_P", "Y_M", "Y_P", "Z_M", "Z_P"};
    for (int i = 0; i < NUM_FACE; ++i) {
      std::cout << faces[i] << ": ";
      bInfo[i].print();
    }
  }
};

#endif
