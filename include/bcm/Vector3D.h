// Vector3D<T> (include/Vector3D.h:36-110): a ghosted 3-component array per
// block, components interleaved (component l of cell (i,j,k) at
// Index3DV(i,j,k) + l).  As Scalar3D, the facade's BlockManager gives it a
// slice of a contiguous host arena mirrored by a device vector field
// (bcm_vfield_t); the (size, vc, data0) constructor is that slice.
#ifndef BCM_FACADE_VECTOR3D_H
#define BCM_FACADE_VECTOR3D_H

#include "DataClass.h"
#include "Index3D.h"
#include "VCUpdater.h"

template <typename T>
class Vector3D : public UpdatableDataClass {
  int nx0, ny0, nz0;
  T* data;
  bool owner;
  Index3DV index;

 public:
  typedef T value_type;
  static const int ncomp = 3;  // values per cell
  Vector3D(const Vec3class::Vec3i& size_, int vc_)
      : UpdatableDataClass(size_, vc_), owner(true), index(size_, vc_) {
    nx0 = size_[0] + 2 * vc_;
    ny0 = size_[1] + 2 * vc_;
    nz0 = size_[2] + 2 * vc_;
    data = new T[3 * size_t(nx0) * ny0 * nz0]();
  }
  Vector3D(const Vec3class::Vec3i& size_, int vc_, T* data0)
      : UpdatableDataClass(size_, vc_), data(data0), owner(false), index(size_, vc_) {
    nx0 = size_[0] + 2 * vc_;
    ny0 = size_[1] + 2 * vc_;
    nz0 = size_[2] + 2 * vc_;
  }
  ~Vector3D() {
    if (owner) delete[] data;
  }
  T* getData() const { return data; }
  Index3DV getIndex() const { return index; }
  // The first component of cell (i,j,k), and component l.
  T& operator()(int i, int j, int k) { return data[index(i, j, k)]; }
  const T& operator()(int i, int j, int k) const { return data[index(i, j, k)]; }
  T& operator()(int i, int j, int k, int l) { return data[index(i, j, k, l)]; }
  const T& operator()(int i, int j, int k, int l) const { return data[index(i, j, k, l)]; }

  // Box copies, three values per cell (include/Vector3D.h:120-190).
  void copyToBuffer(int i0, int j0, int k0, int nx, int ny, int nz, T* buffer) const {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          for (int l = 0; l < 3; ++l)
            buffer[l + 3 * (i + nx * (j + ny * k))] = data[index(i0 + i, j0 + j, k0 + k, l)];
  }
  void copyFromBuffer(int i0, int j0, int k0, int nx, int ny, int nz, const T* buffer) {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          for (int l = 0; l < 3; ++l)
            data[index(i0 + i, j0 + j, k0 + k, l)] = buffer[l + 3 * (i + nx * (j + ny * k))];
  }
  void copyFromDataClass(int i0, int j0, int k0, int i1, int j1, int k1, int nx, int ny, int nz,
                         const DataClass* dataClass) {
    const Vector3D<T>* s = dynamic_cast<const Vector3D<T>*>(dataClass);
    const T* sData = s->getData();
    const Index3DV sIndex = s->getIndex();
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          for (int l = 0; l < 3; ++l)
            data[index(i0 + i, j0 + j, k0 + k, l)] = sData[sIndex(i1 + i, j1 + j, k1 + k, l)];
  }

 private:
  Vector3D(const Vector3D&);
  Vector3D& operator=(const Vector3D&);
};

#endif
