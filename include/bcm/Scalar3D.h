// Scalar3D<T> (include/Scalar3D.h:36-277): a ghosted scalar array per
// block.  Host storage is a slice of the BlockManager's contiguous arena,
// which mirrors the device field; the slice is not owned (the reference's
// delete[] of a slice, include/Scalar3D.h:88-90, is not repeated).
#ifndef BCM_FACADE_SCALAR3D_H
#define BCM_FACADE_SCALAR3D_H

#include "DataClass.h"
#include "Index3D.h"

template <typename T>
class Scalar3D : public UpdatableDataClass {
  int nx0, ny0, nz0;
  T* data;
  bool owner;
  Index3DS index;

 public:
  typedef T value_type;
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

 private:
  Scalar3D(const Scalar3D&);
  Scalar3D& operator=(const Scalar3D&);
};

#endif
