// Minimal 3-vector with the interface the drivers use (include/Vec3.h):
// x/y/z members, operator[], element-wise arithmetic, Vec3i/Vec3r/Vec3d.
#ifndef BCM_FACADE_VEC3_H
#define BCM_FACADE_VEC3_H

#include <ostream>

#define REAL_TYPE double

namespace Vec3class {

template <typename T>
class Vec3 {
 public:
  T x, y, z;
  Vec3(T v = 0) : x(v), y(v), z(v) {}
  Vec3(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
  T& operator[](int a) { return (&x)[a]; }
  const T& operator[](int a) const { return (&x)[a]; }
  Vec3 operator+(const Vec3& v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
  Vec3 operator-(const Vec3& v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
  Vec3 operator*(T s) const { return Vec3(x * s, y * s, z * s); }
  Vec3 operator*(const Vec3& v) const { return Vec3(x * v.x, y * v.y, z * v.z); }
  Vec3 operator/(T s) const { return Vec3(x / s, y / s, z / s); }
  bool operator==(const Vec3& v) const { return x == v.x && y == v.y && z == v.z; }
  bool operator!=(const Vec3& v) const { return !(*this == v); }
};

template <typename T>
inline std::ostream& operator<<(std::ostream& os, const Vec3<T>& v) {
  return os << "(" << v.x << ", " << v.y << ", " << v.z << ")";
}

typedef Vec3<int> Vec3i;
typedef Vec3<double> Vec3d;
typedef Vec3<REAL_TYPE> Vec3r;

}  // namespace Vec3class

#endif
