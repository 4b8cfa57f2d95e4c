// Scalar3D<T> (include/Scalar3D.h:36-277): a ghosted scalar array per
// block.  Host storage is a slice of the BlockManager's contiguous arena,
// which mirrors the device field; the slice is not owned (the reference's
// delete[] of a slice, include/Scalar3D.h:88-90, is not repeated).
#ifndef BCM_FACADE_SCALAR3D_H
#define BCM_FACADE_SCALAR3D_H

#include "DataClass.h"
#include "Index3D.h"
#include "VCUpdater.h"

template <typename T>
class Scalar3D : public UpdatableDataClass {
  int nx0, ny0, nz0;
  T* data;
  bool owner;
  Index3DS index;

 public:
  typedef T value_type;
  static const int ncomp = 1;  // values per cell
  Scalar3D(const Vec3class::Vec3i& size_, int vc_)
      : UpdatableDataClass(size_, vc_), owner(true), index(size_, vc_) {
    nx0 = size_[0] + 2 * vc_;
    ny0 = size_[1] + 2 * vc_;
    nz0 = size_[2] + 2 * vc_;
    data = new T[size_t(nx0) * ny0 * nz0]();
  }
  Scalar3D(const Vec3class::Vec3i& size_, int vc_, T* data0)
      : UpdatableDataClass(size_, vc_), data(data0), owner(false), index(size_, vc_) {
    nx0 = size_[0] + 2 * vc_;
    ny0 = size_[1] + 2 * vc_;
    nz0 = size_[2] + 2 * vc_;
  }
  ~Scalar3D() {
    if (owner) delete[] data;
  }
  T* getData() const { return data; }
  int getVCsize() { return vc; }
  Index3DS getIndex() const { return index; }
  T& operator()(int i, int j, int k) { return data[index(i, j, k)]; }
  const T& operator()(int i, int j, int k) const { return data[index(i, j, k)]; }

  // Box copies between the block and a dense buffer (x fastest), and from
  // another Scalar3D (include/Scalar3D.h:120-137).
  void copyToBuffer(int i0, int j0, int k0, int nx, int ny, int nz, T* buffer) const {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) buffer[i + nx * (j + ny * k)] = data[index(i0 + i, j0 + j, k0 + k)];
  }
  void copyFromBuffer(int i0, int j0, int k0, int nx, int ny, int nz, const T* buffer) {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) data[index(i0 + i, j0 + j, k0 + k)] = buffer[i + nx * (j + ny * k)];
  }
  void copyFromDataClass(int i0, int j0, int k0, int i1, int j1, int k1, int nx, int ny, int nz,
                         const DataClass* dataClass) {
    const Scalar3D<T>* s = dynamic_cast<const Scalar3D<T>*>(dataClass);
    const T* sData = s->getData();
    const Index3DS sIndex = s->getIndex();
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) data[index(i0 + i, j0 + j, k0 + k)] = sData[sIndex(i1 + i, j1 + j, k1 + k)];
  }

 private:
  Scalar3D(const Scalar3D&);
  Scalar3D& operator=(const Scalar3D&);
};

#endif
