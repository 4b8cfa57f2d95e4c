// Restriction / prolongation arithmetic shared by the ghost-fill kernel and
// the halo-gathering sweep.  Values are gathered first (v[q], q = di + 2 dj
// + 4 dk relative to the base cell) and combined in the reference's operand
// order, so the two kernels produce identical bits.
#pragma once

#include <hip/hip_runtime.h>

namespace bcm {

// interpolateF2C (include/Scalar3DUpdater.h:190-198): the 8 fine cells
// (i,j,k), (i+1,j,k), (i,j+1,k), (i+1,j+1,k), (i,j,k+1), ... summed left
// to right, then * 0.125 in double.
template <typename T>
__device__ __forceinline__ T f2c_combine(const T* v) {
  T s = v[0];
  s = s + v[1];
  s = s + v[2];
  s = s + v[3];
  s = s + v[4];
  s = s + v[5];
  s = s + v[6];
  s = s + v[7];
  return T(0.125 * s);
}

// linearInterpolate (include/Scalar3DUpdater.h:242-266).
__device__ __forceinline__ void c2f_weight(int i, int n, int& I, double& r) {
  I = min(max(i / 2 - 1 + i % 2, 0), n - 2);
  r = -0.25 + 0.5 * double(i) - double(I);
}

// interpolateC2F (include/Scalar3DUpdater.h:221-236) on gathered values
// v = c000, c100, c010, c110, c001, c101, c011, c111, in double.
__device__ __forceinline__ double c2f_combine(const double* v, double r, double s, double t) {
  const double r1 = 1.0 - r, s1 = 1.0 - s, t1 = 1.0 - t;
  const double lo = s1 * (r1 * v[0] + r * v[1]) + s * (r1 * v[2] + r * v[3]);
  const double hi = s1 * (r1 * v[4] + r * v[5]) + s * (r1 * v[6] + r * v[7]);
  return t1 * lo + t * hi;
}

template <typename T, typename Get>
__device__ __forceinline__ T f2c_sum(const Get& get, int I, int J, int K) {
  T v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = get(2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + (q >> 2));
  return f2c_combine<T>(v);
}

template <typename T, typename Get>
__device__ __forceinline__ T c2f_interp(const Get& get, int n, int i, int j, int k) {
  int I, J, K;
  double r, s, t;
  c2f_weight(i, n, I, r);
  c2f_weight(j, n, J, s);
  c2f_weight(k, n, K, t);
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = double(get(I + (q & 1), J + ((q >> 1) & 1), K + (q >> 2)));
  return T(c2f_combine(v, r, s, t));
}

}  // namespace bcm
