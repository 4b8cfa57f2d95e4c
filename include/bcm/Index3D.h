// Index3DS / Index3DV (include/Index3D.h:35-100): (i, j, k) -> offset in a
// ghosted block, x fastest; Index3DV for three interleaved components
// (component l at operator()(i,j,k) + l).
#ifndef BCM_FACADE_INDEX3D_H
#define BCM_FACADE_INDEX3D_H

#include "Vec3.h"

class Index3DS {
  int c_j, c_k, c_0;

 public:
  Index3DS(const Vec3class::Vec3i& size, int vc) {
    const int nx0 = size[0] + 2 * vc, ny0 = size[1] + 2 * vc;
    c_j = nx0;
    c_k = nx0 * ny0;
    c_0 = (1 + nx0 + nx0 * ny0) * vc;
  }
  int operator()(int i, int j, int k) const { return i + c_j * j + c_k * k + c_0; }
};

class Index3DV {
  int c_j, c_k, c_0;

 public:
  Index3DV(const Vec3class::Vec3i& size, int vc) {
    const int nx0 = size[0] + 2 * vc, ny0 = size[1] + 2 * vc;
    c_j = 3 * nx0;
    c_k = 3 * nx0 * ny0;
    c_0 = 3 * (1 + nx0 + nx0 * ny0) * vc;
  }
  int operator()(int i, int j, int k) const { return 3 * i + c_j * j + c_k * k + c_0; }
  int operator()(int i, int j, int k, int l) const { return l + 3 * i + c_j * j + c_k * k + c_0; }
};

#endif
