// Matrix-free Poisson coefficients and the per-cell operators of
// examples/SolverTest/blas.f90, shared by the cell sweeps and the fused
// Krylov engine so both evaluate every cell with the same operation order.
#pragma once

#include <hip/hip_runtime.h>

#include "bcm_internal.h"

namespace bcm {

// Poisson operator (examples/SolverTest/bcut.f90:2848-2866): l = 1/h^2 per
// axis, Ap = -(l0+l1)-(l2+l3)-(l4+l5); then, face by face in X_M..Z_P
// order (examples/SolverTest/BlockScalar3D.h:207-214), the outer-face fold:
// Dirichlet Ap -= Anb (bca.f90:12-202), Neumann Ap += Anb (bca.f90:204-394);
// Anb = 0.
template <typename T>
struct Coef7 {
  T p, w, e, s, n, b, t;
};

template <typename T>
__device__ __forceinline__ void fold_face(T& p, T& a, bool neumann) {
  p = neumann ? p + a : p - a;
  a = T(0);
}

template <typename T>
__device__ __forceinline__ Coef7<T> poisson_coef(const BlockMeta& me, int i, int j, int k, int n) {
  const T lx = T(me.cx), ly = T(me.cy), lz = T(me.cz);
  Coef7<T> a;
  a.w = lx; a.e = lx; a.s = ly; a.n = ly; a.b = lz; a.t = lz;
  a.p = ((-(a.w + a.e)) - (a.s + a.n)) - (a.b + a.t);
  const int bm = me.dir_mask | me.neu_mask;
  if (bm) {
    const int nm = me.neu_mask;
    if ((bm & 1) && i == 0) fold_face(a.p, a.w, nm & 1);
    if ((bm & 2) && i == n - 1) fold_face(a.p, a.e, nm & 2);
    if ((bm & 4) && j == 0) fold_face(a.p, a.s, nm & 4);
    if ((bm & 8) && j == n - 1) fold_face(a.p, a.n, nm & 8);
    if ((bm & 16) && k == 0) fold_face(a.p, a.b, nm & 16);
    if ((bm & 32) && k == n - 1) fold_face(a.p, a.t, nm & 32);
  }
  return a;
}

// The folded right-hand side of one cell: rhs, then per outer face in
// X_M..Z_P order Dirichlet b -= Anb*v*2 (bca.f90:33) and Neumann
// b += Anb*xc on a minus face, b -= Anb*xc on a plus face (bca.f90:225,257),
// with xc = v * cellSize.x (examples/SolverTest/BlockScalar3D.h:98-100).
__device__ __forceinline__ double fold_rhs(const BlockMeta& me, int i, int j, int k, int n, double bv,
                                           const double* __restrict__ bval) {
  const int bm = me.dir_mask | me.neu_mask;
  if (!bm) return bv;
  const double l[6] = {me.cx, me.cx, me.cy, me.cy, me.cz, me.cz};
  const int at[6] = {i == 0, i == n - 1, j == 0, j == n - 1, k == 0, k == n - 1};
  for (int f = 0; f < 6; ++f) {
    if (!((bm >> f & 1) && at[f])) continue;
    const double v = bval[me.bc_id[f]];
    if (me.neu_mask >> f & 1) {
      const double xc = v * me.hx;
      bv = (f & 1) ? bv - l[f] * xc : bv + l[f] * xc;
    } else {
      bv = bv - l[f] * v * 2.0;
    }
  }
  return bv;
}

// jacobi_smoother (blas.f90:485-493) on the cell's value xc and its six
// neighbours' values.
template <typename T>
__device__ __forceinline__ T jacobi_values(const Coef7<T>& a, T xw, T xe, T xs, T xn, T xb, T xt, T xc, T bv, T om) {
  T r = bv - a.w * xw;
  r = r - a.e * xe;
  r = r - a.s * xs;
  r = r - a.n * xn;
  r = r - a.b * xb;
  r = r - a.t * xt;
  const T xtent = r / a.p;
  return xc - om * (xc - xtent);
}

// The same on a ghosted array (cell offset o, row and plane strides).
template <typename T>
__device__ __forceinline__ T jacobi_cell(const Coef7<T>& a, const T* __restrict__ x, long long o, int sy, int sz,
                                         T bv, T om) {
  return jacobi_values(a, x[o - 1], x[o + 1], x[o - sy], x[o + sy], x[o - sz], x[o + sz], x[o], bv, om);
}

// calc_ax (blas.f90:584-591): neighbours first, centre last.
template <typename T>
__device__ __forceinline__ T ax_cell(const Coef7<T>& a, const T* __restrict__ x, long long o, int sy, int sz) {
  T v = a.w * x[o - 1];
  v = v + a.e * x[o + 1];
  v = v + a.s * x[o - sy];
  v = v + a.n * x[o + sy];
  v = v + a.b * x[o - sz];
  v = v + a.t * x[o + sz];
  return v + a.p * x[o];
}

// calc_r (blas.f90:632-639).
template <typename T>
__device__ __forceinline__ T r_cell(const Coef7<T>& a, const T* __restrict__ x, long long o, int sy, int sz, T bv) {
  T r = bv - a.w * x[o - 1];
  r = r - a.e * x[o + 1];
  r = r - a.s * x[o - sy];
  r = r - a.n * x[o + sy];
  r = r - a.b * x[o - sz];
  r = r - a.t * x[o + sz];
  return r - a.p * x[o];
}

}  // namespace bcm

namespace bcm {

// The same cell operators on gathered neighbour values (W, E, S, N, B, T
// and the centre), for kernels whose neighbours come partly from a ghost
// frame in LDS; the operation order is that of the pointer forms above.
template <typename T>
__device__ __forceinline__ T jacobi_vals(const Coef7<T>& a, T w, T e, T s, T n, T b, T t, T c, T bv, T om) {
  T r = bv - a.w * w;
  r = r - a.e * e;
  r = r - a.s * s;
  r = r - a.n * n;
  r = r - a.b * b;
  r = r - a.t * t;
  const T xt = r / a.p;
  return c - om * (c - xt);
}

template <typename T>
__device__ __forceinline__ T ax_vals(const Coef7<T>& a, T w, T e, T s, T n, T b, T t, T c) {
  T v = a.w * w;
  v = v + a.e * e;
  v = v + a.s * s;
  v = v + a.n * n;
  v = v + a.b * b;
  v = v + a.t * t;
  return v + a.p * c;
}

}  // namespace bcm
