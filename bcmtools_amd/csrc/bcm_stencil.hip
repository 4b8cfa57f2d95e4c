// Per-block stencil sweeps, BLAS-1 and reductions over all blocks.
//
// Data layout: block-major, each block a ghosted (n+2vc)^3 array with x
// fastest (include/Index3D.h:35-59).  Every kernel is HBM-bound (well
// below 1 flop/byte), so the design goal is one read of each input and
// one write of each output per sweep: coalesced x-rows, LDS staging of
// the z-planes for the red-black sweep, and no MFMA.
//
// Arithmetic follows the reference operand order with FP contraction
// disabled at compile time, so FP64 results are bit-identical to the
// CPU path (no fused multiply-add, IEEE division).
#include <cmath>
#include <mutex>
#include <utility>

#include "bcm_internal.h"
#include "bcm_poisson.h"

namespace bcm {

struct CellGeo {
  int n, vc, nx0;
  long long bvol;
  long long ncell;  // nblocks * n^3
};

// Cells per thread and cells per work-group of a cell sweep.
constexpr int kCPT = 4;
constexpr int kWG = 256;
constexpr int kCellsPerWG = kCPT * kWG;

__device__ __forceinline__ long long cell_offset(const CellGeo& g, long long c, int& b, int& i, int& j, int& k) {
  const long long n3 = (long long)g.n * g.n * g.n;
  b = int(c / n3);
  const int rem = int(c - (long long)b * n3);
  i = rem % g.n;
  j = (rem / g.n) % g.n;
  k = rem / (g.n * g.n);
  return (long long)b * g.bvol + (i + g.vc) + (long long)g.nx0 * ((j + g.vc) + (long long)g.nx0 * (k + g.vc));
}

// Deterministic work-group sum of one double per thread (wave64 butterfly,
// then a fixed-order sum of the four wave totals).
__device__ __forceinline__ double wg_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = ((red[0] + red[1]) + red[2]) + red[3];
  return s;
}

static int cell_grid(const bcm_mesh* m) { return int((m->ncell + kCellsPerWG - 1) / kCellsPerWG); }

static CellGeo cell_geo(const bcm_mesh* m) {
  return CellGeo{m->n, m->vc, m->nx0, (long long)m->bvol, (long long)m->ncell};
}

// ---------------------------------------------------------------------
// Red-black SOR
// ---------------------------------------------------------------------
// Per-block coefficients {c0, cx, cy, cz}:
//   sor1 (examples/sor1/Solver.cpp:98-106): c = 2(cx+cy+cz); cx = w cx / c ...; c0 = 1 - w
//   sor2 (examples/sor2/Solver.cpp:99-102): cx = 1/dx^2 ...; c0 = 1 / (2 (cx+cy+cz))
static int sor_coef(bcm_mesh* m, double omega, int form) {
  if (m->d_sor_coef && m->sor_omega == omega && m->sor_form == form) return BCM_OK;
  std::vector<double> h(size_t(m->nblocks) * 4);
  for (int b = 0; b < m->nblocks; ++b) {
    const bcm_block_info& bi = m->blocks[b];
    const double dx = bi.block_size[0] / m->n, dy = bi.block_size[1] / m->n, dz = bi.block_size[2] / m->n;
    double cx = 1.0 / (dx * dx);
    double cy = 1.0 / (dy * dy);
    double cz = 1.0 / (dz * dz);
    double c0;
    if (form == 1) {
      const double c = 2.0 * (cx + cy + cz);
      cx = omega * cx / c;
      cy = omega * cy / c;
      cz = omega * cz / c;
      c0 = 1.0 - omega;
    } else {
      c0 = 1.0 / (2.0 * (cx + cy + cz));
    }
    h[4 * b + 0] = c0;
    h[4 * b + 1] = cx;
    h[4 * b + 2] = cy;
    h[4 * b + 3] = cz;
  }
  if (!m->d_sor_coef) BCM_TRY(dev_alloc(reinterpret_cast<void**>(&m->d_sor_coef), h.size() * sizeof(double)));
  BCM_HIP_TRY(hipMemcpyAsync(m->d_sor_coef, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, m->stream));
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  m->sor_omega = omega;
  m->sor_form = form;
  return BCM_OK;
}

// One cell update.  sor1: f = c0 f + cx(W+E) + cy(S+N) + cz(B+T);
// sor2: fNew = c0 (cx(W+E) + cy(S+N) + cz(B+T) - s); f = w fNew + (1-w) f.
template <typename T, bool SRC>
__device__ __forceinline__ T sor_update(T fc, T fw, T fe, T fs, T fn, T fb, T ft, T sv, T c0, T cx, T cy, T cz,
                                        T om, T om1) {
  if (SRC) {
    const T fnew = c0 * (((cx * (fw + fe) + cy * (fs + fn)) + cz * (fb + ft)) - sv);
    return om * fnew + om1 * fc;
  } else {
    return ((c0 * fc + cx * (fw + fe)) + cy * (fs + fn)) + cz * (fb + ft);
  }
}

static __device__ __forceinline__ void store2(double* p, double a, double b) {
  *reinterpret_cast<double2*>(p) = make_double2(a, b);
}
static __device__ __forceinline__ void store2(float* p, float a, float b) {
  *reinterpret_cast<float2*>(p) = make_float2(a, b);
}

// Fused two-colour sweep of one block per work-group, vc == 1.
// Thread t owns the 2x2 patch (2pi..2pi+1, 2pj..2pj+1) of every z-plane,
// so each thread updates two red and two black cells per plane.  A ring
// of four ghosted z-planes sits in LDS; at step k the red cells of plane k
// and then the black cells of plane k-1 are updated (black cells read the
// red values of planes k-2..k, all final by then), and plane k-1 is
// written back.  Plane k+2 is prefetched into registers during step k.
template <typename T, int N, bool SRC>
__global__ __launch_bounds__(N* N / 4) void k_rbsor_fused(T* __restrict__ f, const T* __restrict__ s,
                                                         const double* __restrict__ coef, double omega,
                                                         T* __restrict__ cout, T* __restrict__ xout,
                                                         T* __restrict__ scout) {
  constexpr int NX = N + 2;
  constexpr int PL = NX * NX;
  constexpr int NT = N * N / 4;
  constexpr int HALF = N / 2;
  constexpr int PER = (PL + NT - 1) / NT;
  __shared__ T lds[4][PL];
  const int b = blockIdx.x;
  const long long base = (long long)b * PL * NX;
  T* __restrict__ fb = f + base;
  const T* __restrict__ sb = SRC ? s + base : nullptr;
  const T c0 = T(coef[4 * b + 0]), cx = T(coef[4 * b + 1]), cy = T(coef[4 * b + 2]), cz = T(coef[4 * b + 3]);
  const T om = T(omega), om1 = T(1.0 - omega);
  const int t = threadIdx.x;
  const int i0 = 2 * (t % HALF), j0 = 2 * (t / HALF);
  T pre[PER];
  auto fetch = [&](int p) {
    const T* src = fb + (long long)(p + 1) * PL;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + q * NT;
      if (e < PL) pre[q] = src[e];
    }
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = t + q * NT;
      if (e < PL) lds[slot][e] = pre[q];
    }
  };
  auto L = [](int i, int j) { return (j + 1) * NX + (i + 1); };
  // Update the two cells of colour `col` in plane p of my patch.
  auto update = [&](int p, int col, const T* sv) {
    const int P = p & 3, Pm = (p + 3) & 3, Pp = (p + 1) & 3;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int a = q & 1, bb = q >> 1;
      if (((a + bb + p) & 1) != col) continue;
      const int i = i0 + a, j = j0 + bb;
      const int c = L(i, j);
      const T v = sor_update<T, SRC>(lds[P][c], lds[P][c - 1], lds[P][c + 1], lds[P][c - NX], lds[P][c + NX],
                                     lds[Pm][c], lds[Pp][c], SRC ? sv[q] : T(0), c0, cx, cy, cz, om, om1);
      lds[P][c] = v;
    }
  };
  // The swept plane back to f, or (cout given) to the compact iterate of
  // the halo schedule together with its X-face export (columns 0, 1, N-2,
  // N-1); f then keeps its old values.  scout (with cout, SRC only)
  // receives the plane's source values compact, for the later sweeps.
  auto store = [&](int p, const T* sv) {
    const int P = p & 3;
    if (cout) {
      // Each row pair (i0, i0+1) starts at an even compact index: one
      // aligned two-value store per row and array.
      const long long cp = (long long)b * N * N * N + (long long)p * N * N;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int j = j0 + r;
        const T v0 = lds[P][L(i0, j)], v1 = lds[P][L(i0 + 1, j)];
        store2(cout + cp + i0 + N * j, v0, v1);
        if (SRC && scout) store2(scout + cp + i0 + N * j, sv[2 * r], sv[2 * r + 1]);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int i = i0 + a;
          const int col = i < 2 ? i : (i >= N - 2 ? i - (N - 4) : -1);
          if (col >= 0) xout[((long long)(b * 4 + col) * N + p) * N + j] = a ? v1 : v0;
        }
      }
      return;
    }
    T* dst = fb + (long long)(p + 1) * PL;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + (q & 1), j = j0 + (q >> 1);
      dst[L(i, j)] = lds[P][L(i, j)];
    }
    // The rows' X ghost cells go back too, with the values read into LDS
    // (the sweep never changes a ghost, and only this work-group writes the
    // block), so the rows j = 0..N-1 of the plane are stored as one
    // contiguous span instead of rows with holes at both ends: whole cache
    // lines, no partial-line read-modify-write.
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int j = j0 + r;
      if (i0 == 0) dst[L(-1, j)] = lds[P][L(-1, j)];
      if (i0 == N - 2) dst[L(N, j)] = lds[P][L(N, j)];
    }
  };
  auto load_s = [&](int p, T* sv) {
    if (!SRC) return;
    const T* src = sb + (long long)(p + 1) * PL;
#pragma unroll
    for (int q = 0; q < 4; ++q) sv[q] = src[L(i0 + (q & 1), j0 + (q >> 1))];
  };

  T s_cur[4] = {T(0), T(0), T(0), T(0)}, s_prev[4] = {T(0), T(0), T(0), T(0)};
  fetch(-1);
  commit(3);
  fetch(0);
  commit(0);
  fetch(1);
  for (int k = 0; k < N; ++k) {
    commit((k + 1) & 3);
    if (k + 2 <= N) fetch(k + 2);
    load_s(k, s_cur);
    __syncthreads();
    update(k, 0, s_cur);
    __syncthreads();
    if (k >= 1) {
      update(k - 1, 1, s_prev);
      store(k - 1, s_prev);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) s_prev[q] = s_cur[q];
  }
  update(N - 1, 1, s_prev);
  store(N - 1, s_prev);
}

// Generic one-colour pass for any n, vc: one thread per cell of the colour.
template <typename T, bool SRC>
__global__ __launch_bounds__(256) void k_rbsor_colour(T* __restrict__ f, const T* __restrict__ s,
                                                      const double* __restrict__ coef, double omega, CellGeo g,
                                                      int colour) {
  const long long half = g.ncell / 2;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= half) return;
  // Enumerate (block, k, j, ih) with i = 2 ih + (colour + j + k) % 2.
  const int n = g.n, nh = n / 2;
  const long long per_block = (long long)n * n * nh;
  const int b = int(c / per_block);
  int rem = int(c - (long long)b * per_block);
  const int ih = rem % nh;
  const int j = (rem / nh) % n;
  const int k = rem / (nh * n);
  const int i = 2 * ih + ((colour + j + k) & 1);
  const int sy = g.nx0, sz = g.nx0 * g.nx0;
  const long long o = (long long)b * g.bvol + (i + g.vc) + (long long)sy * (j + g.vc) + (long long)sz * (k + g.vc);
  const T c0 = T(coef[4 * b + 0]), cx = T(coef[4 * b + 1]), cy = T(coef[4 * b + 2]), cz = T(coef[4 * b + 3]);
  f[o] = sor_update<T, SRC>(f[o], f[o - 1], f[o + 1], f[o - sy], f[o + sy], f[o - sz], f[o + sz],
                            SRC ? s[o] : T(0), c0, cx, cy, cz, T(omega), T(1.0 - omega));
}

// cout (vc = 1, n = 16 or 32 only): write the swept iterate compact, with
// its X-face export, instead of back into f; scout: the source term
// compact as well.
template <typename T>
static int rbsor_raw(bcm_mesh* m, T* f, const T* s, double omega, bool colour_passes = false, T* cout = nullptr,
                     T* xout = nullptr, T* scout = nullptr) {
  const int form = s ? 2 : 1;
  BCM_TRY(sor_coef(m, omega, form));
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_SOR, &ev));
  const int nb = m->nblocks;
  if (colour_passes) {
    // fall through to the generic kernel below
  } else if (m->vc == 1 && m->n == 32) {
    if (s) hipLaunchKernelGGL((k_rbsor_fused<T, 32, true>), dim3(nb), dim3(256), 0, m->stream, f, s, m->d_sor_coef, omega, cout, xout, scout);
    else hipLaunchKernelGGL((k_rbsor_fused<T, 32, false>), dim3(nb), dim3(256), 0, m->stream, f, s, m->d_sor_coef, omega, cout, xout, scout);
  } else if (m->vc == 1 && m->n == 16) {
    if (s) hipLaunchKernelGGL((k_rbsor_fused<T, 16, true>), dim3(nb), dim3(64), 0, m->stream, f, s, m->d_sor_coef, omega, cout, xout, scout);
    else hipLaunchKernelGGL((k_rbsor_fused<T, 16, false>), dim3(nb), dim3(64), 0, m->stream, f, s, m->d_sor_coef, omega, cout, xout, scout);
  }
  if (colour_passes || !(m->vc == 1 && (m->n == 32 || m->n == 16))) {
    const CellGeo g = cell_geo(m);
    const long long half = g.ncell / 2;
    const int grid = int((half + 255) / 256);
    for (int colour = 0; colour < 2; ++colour) {
      if (s) hipLaunchKernelGGL((k_rbsor_colour<T, true>), dim3(grid), dim3(256), 0, m->stream, f, s, m->d_sor_coef, omega, g, colour);
      else hipLaunchKernelGGL((k_rbsor_colour<T, false>), dim3(grid), dim3(256), 0, m->stream, f, s, m->d_sor_coef, omega, g, colour);
    }
  }
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

// Relative residual of lap_h f = s: r = s - (cx(W+E-2C) + cy(..) + cz(..)).
// The sor2 coefficients 1/h^2 per axis come from the per-block table
// {c0, cx, cy, cz} (FROM_META false) or from the mesh's BlockMeta, which
// holds the same values computed the same way (FROM_META true).
template <bool FROM_META>
__global__ __launch_bounds__(kWG) void k_sor_residual(const double* __restrict__ f, const double* __restrict__ s,
                                                      const double* __restrict__ coef_sor2,
                                                      const BlockMeta* __restrict__ meta, CellGeo g,
                                                      double* __restrict__ parts) {
  __shared__ double red[4];
  const int sy = g.nx0, sz = g.nx0 * g.nx0;
  double rr = 0.0, ss = 0.0;
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) break;
    int b, i, j, k;
    const long long o = cell_offset(g, c, b, i, j, k);
    const double cx = FROM_META ? meta[b].cx : coef_sor2[4 * b + 1];
    const double cy = FROM_META ? meta[b].cy : coef_sor2[4 * b + 2];
    const double cz = FROM_META ? meta[b].cz : coef_sor2[4 * b + 3];
    const double fc = f[o];
    const double lap = (cx * ((f[o - 1] + f[o + 1]) - 2.0 * fc) + cy * ((f[o - sy] + f[o + sy]) - 2.0 * fc)) +
                       cz * ((f[o - sz] + f[o + sz]) - 2.0 * fc);
    const double r = s[o] - lap;
    rr += r * r;
    ss += s[o] * s[o];
  }
  const double a = wg_sum(rr, red);
  __syncthreads();
  const double bsum = wg_sum(ss, red);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = a;
    parts[gridDim.x + blockIdx.x] = bsum;
  }
}

// ---------------------------------------------------------------------
// 7-point operators.  Stored coefficients follow blas.f90; the Poisson
// operator computes the same coefficients on the fly: l = 1/h^2 per
// axis, Ap = -(l0+l1)-(l2+l3)-(l4+l5) (bcut.f90:2857-2859) and, on an
// outer Dirichlet face, Ap -= Anb; Anb = 0 (bca.f90:12-42).
// ---------------------------------------------------------------------
template <typename T>
struct StoredA {
  const T *p, *w, *e, *s, *n, *b, *t;
};

template <typename T>
__device__ __forceinline__ Coef7<T> stored_coef(const StoredA<T>& A, long long o) {
  return Coef7<T>{A.p[o], A.w[o], A.e[o], A.s[o], A.n[o], A.b[o], A.t[o]};
}

template <typename T, bool STORED>
__global__ __launch_bounds__(kWG) void k_jacobi(T* __restrict__ x1, const T* __restrict__ x0, StoredA<T> A,
                                                const BlockMeta* __restrict__ meta, const T* __restrict__ b,
                                                double omega, CellGeo g, const T* __restrict__ mask) {
  const int sy = g.nx0, sz = g.nx0 * g.nx0;
  const T om = T(omega);
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) return;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    if (mask && mask[o] != T(1)) continue;
    const Coef7<T> a = STORED ? stored_coef(A, o) : poisson_coef<T>(meta[bk], i, j, k, g.n);
    x1[o] = jacobi_cell(a, x0, o, sy, sz, b[o], om);
  }
}

// mode 0: Ax; mode 1: r = b - Ax; mode 2: sum r^2 only.
template <typename T, bool STORED, int MODE>
__global__ __launch_bounds__(kWG) void k_apply(T* __restrict__ out, const T* __restrict__ x, StoredA<T> A,
                                               const BlockMeta* __restrict__ meta, const T* __restrict__ b,
                                               CellGeo g, double* __restrict__ parts, const T* __restrict__ mask) {
  __shared__ double red[4];
  const int sy = g.nx0, sz = g.nx0 * g.nx0;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) break;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    if (mask && mask[o] != T(1)) continue;
    const Coef7<T> a = STORED ? stored_coef(A, o) : poisson_coef<T>(meta[bk], i, j, k, g.n);
    if (MODE == 0) {
      out[o] = ax_cell(a, x, o, sy, sz);
    } else {
      const T r = r_cell(a, x, o, sy, sz, b[o]);
      if (MODE == 1) out[o] = r;
      else acc += double(r) * double(r);
    }
  }
  if (MODE == 2) {
    const double s = wg_sum(acc, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
  }
}

// FP64 residual r = b - A x of the Poisson operator, written to FP32
// (the refinement step's right-hand side), with sum r^2 in FP64.
__global__ __launch_bounds__(kWG) void k_residual_d2f(float* __restrict__ r32, const double* __restrict__ x,
                                                      const BlockMeta* __restrict__ meta,
                                                      const double* __restrict__ b, CellGeo g,
                                                      double* __restrict__ parts) {
  __shared__ double red[4];
  const int sy = g.nx0, sz = g.nx0 * g.nx0;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) break;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    const Coef7<double> a = poisson_coef<double>(meta[bk], i, j, k, g.n);
    const double r = r_cell(a, x, o, sy, sz, b[o]);
    if (r32) r32[o] = float(r);
    acc += r * r;
  }
  const double s = wg_sum(acc, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// dot (blas.f90:379-412): per-thread accumulation in T, work-group and
// global sums in FP64 by a fixed tree.
template <typename T>
__global__ __launch_bounds__(kWG) void k_dot(const T* __restrict__ x, const T* __restrict__ y, CellGeo g,
                                             double* __restrict__ parts, const T* __restrict__ mask) {
  __shared__ double red[4];
  T acc = T(0);
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) break;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    if (mask && mask[o] != T(1)) continue;
    acc = acc + x[o] * y[o];
  }
  const double s = wg_sum(double(acc), red);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// blas.f90:242-377 (interior cells only).
template <typename T, int OP>
__global__ __launch_bounds__(kWG) void k_blas(T* __restrict__ z, const T* __restrict__ x, const T* __restrict__ y,
                                              double a, double bcoef, CellGeo g, const T* __restrict__ mask) {
  const T ta = T(a), tb = T(bcoef);
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) return;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    if (mask && mask[o] != T(1)) continue;
    if (OP == OP_AXPY) z[o] = z[o] + ta * x[o];                       // y = y + a x
    else if (OP == OP_XPAY) z[o] = x[o] + ta * z[o];                  // y = x + a y
    else if (OP == OP_AXPYZ) z[o] = ta * x[o] + y[o];                 // z = a x + y
    else if (OP == OP_AXPBYPZ) z[o] = (ta * x[o] + tb * y[o]) + z[o]; // z = a x + b y + z
    else z[o] = x[o];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_all(T* __restrict__ x, T v, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] = v;
}

// setup_mask: 1 on interior cells, 0 on ghosts, every block.
template <typename T>
__global__ __launch_bounds__(256) void k_setup_mask(T* __restrict__ x, int n, int vc, long long count) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  const int nx0 = n + 2 * vc;
  const int r = int(c % ((long long)nx0 * nx0 * nx0));
  const int i = r % nx0 - vc, j = (r / nx0) % nx0 - vc, k = r / (nx0 * nx0) - vc;
  const bool inside = i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n;
  x[c] = inside ? T(1) : T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_all(T* __restrict__ y, const T* __restrict__ x, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = x[i];
}

// Interior conversions for the refinement loop.
__global__ __launch_bounds__(kWG) void k_convert(void* __restrict__ dst, const void* __restrict__ src, int op,
                                                 CellGeo g) {
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) return;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    if (op == CONV_D2F) static_cast<float*>(dst)[o] = float(static_cast<const double*>(src)[o]);
    else static_cast<double*>(dst)[o] = static_cast<double*>(dst)[o] + double(static_cast<const float*>(src)[o]);
  }
}

// Poisson assembly: stored A (optional) and folded b
// (bcut.f90:2848-2865 without cut cells; the bca.f90 Dirichlet and Neumann folds).
__global__ __launch_bounds__(kWG) void k_assemble(double* Ap, double* Aw, double* Ae, double* As, double* An,
                                                  double* Ab, double* At, double* __restrict__ bout,
                                                  const double* __restrict__ rhs, const BlockMeta* __restrict__ meta,
                                                  const double* __restrict__ bval, CellGeo g) {
#pragma unroll
  for (int q = 0; q < kCPT; ++q) {
    const long long c = (long long)blockIdx.x * kCellsPerWG + q * kWG + threadIdx.x;
    if (c >= g.ncell) return;
    int bk, i, j, k;
    const long long o = cell_offset(g, c, bk, i, j, k);
    const BlockMeta me = meta[bk];
    bout[o] = fold_rhs(me, i, j, k, g.n, rhs[o], bval);
    if (Ap) {
      const Coef7<double> a = poisson_coef<double>(me, i, j, k, g.n);
      Ap[o] = a.p; Aw[o] = a.w; Ae[o] = a.e; As[o] = a.s; An[o] = a.n; Ab[o] = a.b; At[o] = a.t;
    }
  }
}

// ---------------------------------------------------------------------
// Raw launchers
// ---------------------------------------------------------------------
template <typename T>
int jacobi_poisson_raw(bcm_mesh* m, T* x1, const T* x0, const T* b, double omega) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_JACOBI, &ev));
  hipLaunchKernelGGL((k_jacobi<T, false>), dim3(cell_grid(m)), dim3(kWG), 0, m->stream, x1, x0, StoredA<T>{},
                     m->d_meta, b, omega, cell_geo(m), (const T*)nullptr);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}
template int jacobi_poisson_raw<double>(bcm_mesh*, double*, const double*, const double*, double);
template int jacobi_poisson_raw<float>(bcm_mesh*, float*, const float*, const float*, double);

template <typename T>
int spmv_poisson_raw(bcm_mesh* m, T* ax, const T* x) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_SPMV, &ev));
  hipLaunchKernelGGL((k_apply<T, false, 0>), dim3(cell_grid(m)), dim3(kWG), 0, m->stream, ax, x, StoredA<T>{},
                     m->d_meta, (const T*)nullptr, cell_geo(m), m->d_partials, (const T*)nullptr);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}
template int spmv_poisson_raw<double>(bcm_mesh*, double*, const double*);
template int spmv_poisson_raw<float>(bcm_mesh*, float*, const float*);

template <typename T>
int residual_poisson_raw(bcm_mesh* m, T* r, const T* x, const T* b, bool want_norm, int slot) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_RESID, &ev));
  const int grid = cell_grid(m);
  if (want_norm)
    hipLaunchKernelGGL((k_apply<T, false, 2>), dim3(grid), dim3(kWG), 0, m->stream, r, x, StoredA<T>{}, m->d_meta, b,
                       cell_geo(m), m->d_partials, (const T*)nullptr);
  else
    hipLaunchKernelGGL((k_apply<T, false, 1>), dim3(grid), dim3(kWG), 0, m->stream, r, x, StoredA<T>{}, m->d_meta, b,
                       cell_geo(m), m->d_partials, (const T*)nullptr);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  if (want_norm) BCM_TRY(reduce_partials(m, grid, slot));
  return BCM_OK;
}
template int residual_poisson_raw<double>(bcm_mesh*, double*, const double*, const double*, bool, int);
template int residual_poisson_raw<float>(bcm_mesh*, float*, const float*, const float*, bool, int);

int residual_d2f_raw(bcm_mesh* m, float* r32, const double* x, const double* b, int slot) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_RESID, &ev));
  const int grid = cell_grid(m);
  hipLaunchKernelGGL(k_residual_d2f, dim3(grid), dim3(kWG), 0, m->stream, r32, x, m->d_meta, b, cell_geo(m),
                     m->d_partials);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  return reduce_partials(m, grid, slot);
}

template <typename T>
int dot_raw(bcm_mesh* m, const T* x, const T* y, int slot, const T* mask) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_DOT, &ev));
  const int grid = cell_grid(m);
  hipLaunchKernelGGL(k_dot<T>, dim3(grid), dim3(kWG), 0, m->stream, x, y, cell_geo(m), m->d_partials, mask);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  return reduce_partials(m, grid, slot);
}
template int dot_raw<double>(bcm_mesh*, const double*, const double*, int, const double*);
template int dot_raw<float>(bcm_mesh*, const float*, const float*, int, const float*);

template <typename T>
int blas_raw(bcm_mesh* m, int op, T* z, const T* x, const T* y, double a, double b, const T* mask) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_BLAS1, &ev));
  const dim3 grid(cell_grid(m)), blk(kWG);
  const CellGeo g = cell_geo(m);
  switch (op) {
    case OP_AXPY: hipLaunchKernelGGL((k_blas<T, OP_AXPY>), grid, blk, 0, m->stream, z, x, y, a, b, g, mask); break;
    case OP_XPAY: hipLaunchKernelGGL((k_blas<T, OP_XPAY>), grid, blk, 0, m->stream, z, x, y, a, b, g, mask); break;
    case OP_AXPYZ: hipLaunchKernelGGL((k_blas<T, OP_AXPYZ>), grid, blk, 0, m->stream, z, x, y, a, b, g, mask); break;
    case OP_AXPBYPZ: hipLaunchKernelGGL((k_blas<T, OP_AXPBYPZ>), grid, blk, 0, m->stream, z, x, y, a, b, g, mask); break;
    default: hipLaunchKernelGGL((k_blas<T, OP_COPY_INTERIOR>), grid, blk, 0, m->stream, z, x, y, a, b, g, mask); break;
  }
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}
template int blas_raw<double>(bcm_mesh*, int, double*, const double*, const double*, double, double, const double*);
template int blas_raw<float>(bcm_mesh*, int, float*, const float*, const float*, double, double, const float*);

template <typename T>
int fill_raw(bcm_mesh* m, T* x, double a) {
  const long long count = (long long)m->nblocks * (long long)m->bvol;
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_BLAS1, &ev));
  hipLaunchKernelGGL(k_fill_all<T>, dim3(unsigned((count + 255) / 256)), dim3(256), 0, m->stream, x, T(a), count);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}
template int fill_raw<double>(bcm_mesh*, double*, double);
template int fill_raw<float>(bcm_mesh*, float*, double);

int convert_raw(bcm_mesh* m, int op, void* dst, const void* src) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_BLAS1, &ev));
  hipLaunchKernelGGL(k_convert, dim3(cell_grid(m)), dim3(kWG), 0, m->stream, dst, src, op, cell_geo(m));
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

int assemble_poisson_b_raw(bcm_mesh* m, double* b, const double* rhs) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_ASSEMBLE, &ev));
  hipLaunchKernelGGL(k_assemble, dim3(cell_grid(m)), dim3(kWG), 0, m->stream, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr, (double*)nullptr, b, rhs,
                     m->d_meta, m->d_bvalues, cell_geo(m));
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

}  // namespace bcm

using namespace bcm;

// ---------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------
namespace {

bool same_mesh(bcm_field_t a, bcm_field_t b) { return a && b && a->mesh == b->mesh && a->dtype == b->dtype; }

template <typename T>
StoredA<T> stored(const bcm_field_t A[7]) {
  return StoredA<T>{static_cast<const T*>(A[0]->data), static_cast<const T*>(A[1]->data),
                    static_cast<const T*>(A[2]->data), static_cast<const T*>(A[3]->data),
                    static_cast<const T*>(A[4]->data), static_cast<const T*>(A[5]->data),
                    static_cast<const T*>(A[6]->data)};
}

bool check_A(bcm_field_t ref, const bcm_field_t A[7]) {
  if (!A) return false;
  for (int i = 0; i < 7; ++i)
    if (!same_mesh(ref, A[i])) return false;
  return true;
}

}  // namespace

extern "C" int bcm_rbsor_sweep(bcm_field_t f, bcm_field_t source, double omega) {
  if (!f || (source && !same_mesh(f, source))) return BCM_ERR_ARG;
  if (f->dtype == BCM_F64)
    return rbsor_raw<double>(f->mesh, static_cast<double*>(f->data),
                             source ? static_cast<const double*>(source->data) : nullptr, omega);
  return rbsor_raw<float>(f->mesh, static_cast<float*>(f->data),
                          source ? static_cast<const float*>(source->data) : nullptr, omega);
}

// `iters` >= 2 outer iterations with the halo-gathering schedule.  Sweep 1
// uses the ghosts stored in f, as the reference's first iteration does,
// and writes its result compact with the X-face export; sweeps 2..iters
// build each block's ghost frame from the previous compact iterate
// (ping-pong between two workspace buffers, the source term compact too),
// the last one writing f's interior; then f's ghosts are completed.  f's
// own ghost cells are not read after sweep 1: every face of a mesh this
// schedule runs on has a neighbour or a Dirichlet condition.
template <typename T>
static int sor_iterate_halo(bcm_field_t f, bcm_field_t source, double omega, int iters) {
  bcm_mesh* m = f->mesh;
  int rc = BCM_OK;
  const size_t xbytes = size_t(m->nblocks) * 4 * m->n * m->n * sizeof(double);
  for (int i = 0; i < 2; ++i)
    if (!m->d_xexp[i]) BCM_TRY(dev_alloc(&m->d_xexp[i], xbytes));
  const size_t cbytes = size_t(m->ncell) * sizeof(T);
  const bool f64 = sizeof(T) == 8;
  T* A = static_cast<T*>(workspace(m, f64 ? WS_SOR_CA_F64 : WS_SOR_CA_F32, cbytes, &rc));
  if (rc != BCM_OK) return rc;
  T* B = static_cast<T*>(workspace(m, f64 ? WS_SOR_CB_F64 : WS_SOR_CB_F32, cbytes, &rc));
  if (rc != BCM_OK) return rc;
  T* d = static_cast<T*>(f->data);
  const T* sg = source ? static_cast<const T*>(source->data) : nullptr;
  T* s = nullptr;
  if (source) {
    s = static_cast<T*>(workspace(m, f64 ? WS_SOR_CS_F64 : WS_SOR_CS_F32, cbytes, &rc));
    if (rc != BCM_OK) return rc;
  }
  T* xe[2] = {static_cast<T*>(m->d_xexp[0]), static_cast<T*>(m->d_xexp[1])};
  // The first sweep reads f's stored ghosts and writes the compact iterate,
  // its X export and the compact source term.
  BCM_TRY(rbsor_raw<T>(m, d, sg, omega, false, A, xe[0], s));
  int cur = 0;
  for (int it = 1; it < iters; ++it) {
    BCM_TRY(halo_sweep_raw<T>(m, A, B, s, omega, xe[cur], xe[cur ^ 1], it == iters - 1 ? d : nullptr));
    std::swap(A, B);
    cur ^= 1;
  }
  // Complete the ghosts of the final iterate: exchange, then the mirror.
  BCM_TRY(ghost_fill_raw<T>(m, d, BCM_GF_ALL));
  return dirichlet_device(m, f);
}

static bool variant_supported(const bcm_mesh* m, int v) {
  switch (v) {
    case BCM_SOR_AUTO:
    case BCM_SOR_FUSED_TASKS:
    case BCM_SOR_COLOUR_TASKS: return true;
    case BCM_SOR_FUSED_GATHER: return m->vc == 1 && (m->n == 4 || m->n == 8 || m->n == 16 || m->n == 32);
    case BCM_SOR_HALO: return m->vc == 1 && (m->n == 16 || m->n == 32) && m->halo_ok;
    default: return false;
  }
}

// `inner` sweeps + in-block BC per block between exchanges (nLoopInner).
// The halo-gathering schedule reads the neighbours' latest iterate, which
// only matches the exchange order for inner == 1; otherwise the fused
// sweeps over f's own (frozen) ghosts run.
template <typename T>
static int sor_iterate_t(bcm_field_t f, bcm_field_t source, double omega, const double* bvalues, int nbvalues,
                         int iters, int variant, int inner = 1) {
  bcm_mesh* m = f->mesh;
  T* d = static_cast<T*>(f->data);
  const T* s = source ? static_cast<const T*>(source->data) : nullptr;
  if (variant == BCM_SOR_AUTO) variant = variant_supported(m, BCM_SOR_HALO) ? BCM_SOR_HALO : BCM_SOR_FUSED_TASKS;
  if (variant == BCM_SOR_HALO && inner > 1) variant = BCM_SOR_FUSED_TASKS;
  if (variant == BCM_SOR_HALO && iters >= 2) return sor_iterate_halo<T>(f, source, omega, iters);
  for (int it = 0; it < iters; ++it) {
    d = static_cast<T*>(f->data);
    for (int in = 0; in < inner; ++in) {
      BCM_TRY(rbsor_raw<T>(m, d, s, omega, variant == BCM_SOR_COLOUR_TASKS));
      BCM_TRY(dirichlet_device(m, f));
    }
    if (variant == BCM_SOR_FUSED_GATHER) BCM_TRY(ghost_gather_raw<T>(m, d));
    else BCM_TRY(ghost_fill_raw<T>(m, d, BCM_GF_ALL));
    BCM_TRY(dirichlet_device(m, f));
  }
  return BCM_OK;
}

extern "C" int bcm_sor_iterate_inner(bcm_field_t f, bcm_field_t source, double omega, const double* bvalues,
                                     int nbvalues, int iters, int inner) {
  if (!f || iters < 0 || inner < 1 || (source && !same_mesh(f, source))) return BCM_ERR_ARG;
  BCM_TRY(bcm_dirichlet_mirror(f, bvalues, nbvalues));  // validates and uploads the table
  bcm_mesh* m = f->mesh;
  const int v = variant_supported(m, m->sor_variant) ? m->sor_variant : BCM_SOR_AUTO;
  return f->dtype == BCM_F64 ? sor_iterate_t<double>(f, source, omega, bvalues, nbvalues, iters, v, inner)
                             : sor_iterate_t<float>(f, source, omega, bvalues, nbvalues, iters, v, inner);
}

extern "C" int bcm_sor_iterate(bcm_field_t f, bcm_field_t source, double omega, const double* bvalues, int nbvalues,
                               int iters) {
  return bcm_sor_iterate_inner(f, source, omega, bvalues, nbvalues, iters, 1);
}

extern "C" int bcm_sor_set_variant(bcm_mesh_t m, int variant) {
  if (!m || variant < 0 || variant >= BCM_SOR_NUM_VARIANTS || !variant_supported(m, variant)) return BCM_ERR_ARG;
  m->sor_variant = variant;
  return BCM_OK;
}

extern "C" int bcm_sor_get_variant(bcm_mesh_t m) { return m ? m->sor_variant : -BCM_ERR_ARG; }

extern "C" int bcm_sor_autotune(bcm_field_t f, bcm_field_t source, double omega, const double* bvalues, int nbvalues,
                                int iters, int* best, double* ms_per_iter) {
  if (!f || !best || iters < 2 || (source && !same_mesh(f, source))) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  for (const auto& bi : m->blocks)
    for (int k = 0; k < 6; ++k)
      if (bi.bc_type[k] == BCM_BC_NEUMANN) return BCM_ERR_ARG;  // as bcm_dirichlet_mirror
  BCM_TRY(set_bvalues(m, bvalues, nbvalues));  // upload only: f itself stays untouched
  // Scratch copy of f: the candidates run on it, f keeps its values.
  bcm_field scratch = *f;
  void* buf = nullptr;
  BCM_TRY(dev_alloc(&buf, f->bytes));
  scratch.data = buf;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BCM_OK;
  double best_ms = 0.0;
  *best = BCM_SOR_AUTO;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = BCM_ERR_HIP;
  for (int v = BCM_SOR_FUSED_TASKS; v < BCM_SOR_NUM_VARIANTS && rc == BCM_OK; ++v) {
    if (ms_per_iter) ms_per_iter[v] = -1.0;
    if (!variant_supported(m, v)) continue;
    float ms = 0.f;
    for (int rep = 0; rep < 2 && rc == BCM_OK; ++rep) {  // the first run warms up (workspaces, caches)
      if (hipMemcpyAsync(scratch.data, f->data, f->bytes, hipMemcpyDeviceToDevice, m->stream) != hipSuccess ||
          hipEventRecord(e0, m->stream) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      rc = f->dtype == BCM_F64 ? sor_iterate_t<double>(&scratch, source, omega, bvalues, nbvalues, iters, v)
                               : sor_iterate_t<float>(&scratch, source, omega, bvalues, nbvalues, iters, v);
      if (rc == BCM_OK && (hipEventRecord(e1, m->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = BCM_ERR_HIP;
    }
    if (rc != BCM_OK) break;
    const double per = double(ms) / iters;
    if (ms_per_iter) ms_per_iter[v] = per;
    if (*best == BCM_SOR_AUTO || per < best_ms) {
      *best = v;
      best_ms = per;
    }
  }
  if (ms_per_iter) ms_per_iter[BCM_SOR_AUTO] = -1.0;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(m->stream);
  dev_free(scratch.data);
  if (rc == BCM_OK) m->sor_variant = *best;
  return rc;
}

extern "C" int bcm_sor_residual(bcm_field_t f, bcm_field_t source, double* relres) {
  if (!f || !source || !relres || !same_mesh(f, source) || f->dtype != BCM_F64) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  BCM_TRY(sor_coef(m, m->sor_form == 2 ? m->sor_omega : 1.0, 2));
  const int grid = cell_grid(m);
  if (size_t(2 * grid) > m->n_partials) return BCM_ERR_STATE;
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_RESID, &ev));
  hipLaunchKernelGGL(k_sor_residual<false>, dim3(grid), dim3(kWG), 0, m->stream,
                     static_cast<const double*>(f->data), static_cast<const double*>(source->data), m->d_sor_coef,
                     (const BlockMeta*)nullptr, cell_geo(m), m->d_partials);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  BCM_TRY(reduce_partials(m, grid, 0));
  BCM_TRY(reduce_partials_at(m, m->d_partials + grid, grid, 1));
  double v[2];
  BCM_TRY(fetch_result(m, 2, v));
  *relres = v[1] > 0.0 ? std::sqrt(v[0] / v[1]) : std::sqrt(v[0]);
  return BCM_OK;
}

// ---------------------------------------------------------------------
// bcm_sor_residual without allocation.  The partial sums and the two
// totals live in `scratch`, a file-scope device array every GPU holds once
// for the life of the process: kNoAllocParts partials per sum (one per
// work-group of kCellsPerWG cells, so meshes up to
// kNoAllocParts * kCellsPerWG = 268,435,456 interior cells) plus the two
// totals.  Every partial a call reads is written by the same call, so
// nothing needs clearing; calls that overlap on one GPU share the array
// and race.
// ---------------------------------------------------------------------
constexpr int kNoAllocParts = 1 << 18;
constexpr int kNoAllocScratch = 2 * kNoAllocParts + 64;  // 4 MiB + 512 B
static __device__ __align__(256) double scratch[kNoAllocScratch];

// Both totals in one launch: work-group w sums parts[w*n .. w*n+n) into out[w].
__global__ __launch_bounds__(256) void k_reduce_pair(const double* __restrict__ parts, int n,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  const double s = sum_partials_wg256(parts + (long long)blockIdx.x * n, n, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
static int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return BCM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    BCM_HIP_TRY(hipGetDevice(&cur));
    if (cur != device) BCM_HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) BCM_HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "bcm: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return BCM_ERR_HIP;
    }
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return BCM_OK;
}

extern "C" int bcm_sor_residual_noalloc(bcm_field_t f, bcm_field_t source, double* relres) {
  if (!f || !source || !relres || !same_mesh(f, source) || f->dtype != BCM_F64) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  const long long grid = ((long long)m->ncell + kCellsPerWG - 1) / kCellsPerWG;
  if (grid > kNoAllocParts) {
    std::fprintf(stderr,
                 "bcm: bcm_sor_residual_noalloc: %lld interior cells exceed its fixed scratch (at most %lld)\n",
                 (long long)m->ncell, (long long)kNoAllocParts * kCellsPerWG);
    return BCM_ERR_ARG;
  }
  double* ptr = nullptr;
  BCM_TRY(scratch_on(m->device, &ptr));
  double* parts = ptr;
  double* totals = ptr + 2 * kNoAllocParts;
  // No per-kernel timing here: the timer's events come from a pool that
  // grows on demand.
  hipLaunchKernelGGL(k_sor_residual<true>, dim3(int(grid)), dim3(kWG), 0, m->stream,
                     static_cast<const double*>(f->data), static_cast<const double*>(source->data),
                     (const double*)nullptr, m->d_meta, cell_geo(m), parts);
  BCM_HIP_TRY(hipGetLastError());
  // k_sor_residual writes the r^2 partials at parts[0..grid) and the s^2
  // partials at parts[grid..2 grid).
  hipLaunchKernelGGL(k_reduce_pair, dim3(2), dim3(256), 0, m->stream, (const double*)parts, int(grid), totals);
  BCM_HIP_TRY(hipGetLastError());
  double v[2] = {0.0, 0.0};
  BCM_HIP_TRY(hipMemcpyAsync(v, totals, sizeof(v), hipMemcpyDeviceToHost, m->stream));
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  *relres = v[1] > 0.0 ? std::sqrt(v[0] / v[1]) : std::sqrt(v[0]);
  return BCM_OK;
}

// A mask field (nullable) has the operands' mesh and type; a cell is
// visited when its value is 1, as blas.f90's `mask(ijk) == 1`.
static bool mask_ok(bcm_field_t x, bcm_field_t mask) { return !mask || same_mesh(x, mask); }

template <typename T>
static const T* mask_ptr(bcm_field_t mask) {
  return mask ? static_cast<const T*>(mask->data) : nullptr;
}

static int jacobi7_impl(bcm_field_t x1, bcm_field_t x0, const bcm_field_t A[7], bcm_field_t b, double omega,
                        bcm_field_t mask) {
  if (!same_mesh(x1, x0) || !same_mesh(x1, b) || !check_A(x1, A) || x1 == x0 || !mask_ok(x1, mask))
    return BCM_ERR_ARG;
  bcm_mesh* m = x1->mesh;
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_JACOBI, &ev));
  if (x1->dtype == BCM_F64)
    hipLaunchKernelGGL((k_jacobi<double, true>), dim3(cell_grid(m)), dim3(kWG), 0, m->stream,
                       static_cast<double*>(x1->data), static_cast<const double*>(x0->data), stored<double>(A),
                       m->d_meta, static_cast<const double*>(b->data), omega, cell_geo(m), mask_ptr<double>(mask));
  else
    hipLaunchKernelGGL((k_jacobi<float, true>), dim3(cell_grid(m)), dim3(kWG), 0, m->stream,
                       static_cast<float*>(x1->data), static_cast<const float*>(x0->data), stored<float>(A),
                       m->d_meta, static_cast<const float*>(b->data), omega, cell_geo(m), mask_ptr<float>(mask));
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_jacobi7(bcm_field_t x1, bcm_field_t x0, const bcm_field_t A[7], bcm_field_t b, double omega) {
  return jacobi7_impl(x1, x0, A, b, omega, nullptr);
}

extern "C" int bcm_jacobi7_mask(bcm_field_t x1, bcm_field_t x0, const bcm_field_t A[7], bcm_field_t b,
                                bcm_field_t mask, double omega) {
  if (!mask) return BCM_ERR_ARG;
  return jacobi7_impl(x1, x0, A, b, omega, mask);
}

template <typename T, int MODE>
static int apply_stored(bcm_field_t out, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, bcm_field_t mask) {
  bcm_mesh* m = x->mesh;
  EventPair ev;
  BCM_TRY(timer_begin(m, MODE == 0 ? BCM_K_SPMV : BCM_K_RESID, &ev));
  const int grid = cell_grid(m);
  hipLaunchKernelGGL((k_apply<T, true, MODE>), dim3(grid), dim3(kWG), 0, m->stream,
                     out ? static_cast<T*>(out->data) : nullptr, static_cast<const T*>(x->data), stored<T>(A),
                     m->d_meta, b ? static_cast<const T*>(b->data) : nullptr, cell_geo(m), m->d_partials,
                     mask_ptr<T>(mask));
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  if (MODE == 2) BCM_TRY(reduce_partials(m, grid, 0));
  return BCM_OK;
}

static int calc_ax_impl(bcm_field_t ax, const bcm_field_t A[7], bcm_field_t x, bcm_field_t mask) {
  if (!same_mesh(ax, x) || !check_A(x, A) || ax == x || !mask_ok(x, mask)) return BCM_ERR_ARG;
  return x->dtype == BCM_F64 ? apply_stored<double, 0>(ax, A, x, nullptr, mask)
                             : apply_stored<float, 0>(ax, A, x, nullptr, mask);
}

static int calc_r_impl(bcm_field_t r, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, bcm_field_t mask) {
  if (!same_mesh(r, x) || !same_mesh(x, b) || !check_A(x, A) || r == x || !mask_ok(x, mask)) return BCM_ERR_ARG;
  return x->dtype == BCM_F64 ? apply_stored<double, 1>(r, A, x, b, mask) : apply_stored<float, 1>(r, A, x, b, mask);
}

static int calc_r2_impl(const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, double* rr, bcm_field_t mask) {
  if (!same_mesh(x, b) || !check_A(x, A) || !rr || !mask_ok(x, mask)) return BCM_ERR_ARG;
  const int rc = x->dtype == BCM_F64 ? apply_stored<double, 2>(nullptr, A, x, b, mask)
                                     : apply_stored<float, 2>(nullptr, A, x, b, mask);
  if (rc != BCM_OK) return rc;
  return fetch_result(x->mesh, 1, rr);
}

static int dot_impl(bcm_field_t x, bcm_field_t y, double* xy, bcm_field_t mask) {
  if (!same_mesh(x, y) || !xy || !mask_ok(x, mask)) return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  if (x->dtype == BCM_F64)
    BCM_TRY(dot_raw<double>(m, static_cast<const double*>(x->data), static_cast<const double*>(y->data), 0,
                            mask_ptr<double>(mask)));
  else
    BCM_TRY(dot_raw<float>(m, static_cast<const float*>(x->data), static_cast<const float*>(y->data), 0,
                           mask_ptr<float>(mask)));
  return fetch_result(m, 1, xy);
}

extern "C" int bcm_calc_ax(bcm_field_t ax, const bcm_field_t A[7], bcm_field_t x) {
  return calc_ax_impl(ax, A, x, nullptr);
}
extern "C" int bcm_calc_r(bcm_field_t r, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b) {
  return calc_r_impl(r, A, x, b, nullptr);
}
extern "C" int bcm_calc_r2(const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, double* rr) {
  return calc_r2_impl(A, x, b, rr, nullptr);
}
extern "C" int bcm_dot(bcm_field_t x, bcm_field_t y, double* xy) { return dot_impl(x, y, xy, nullptr); }

extern "C" int bcm_calc_ax_mask(bcm_field_t ax, const bcm_field_t A[7], bcm_field_t x, bcm_field_t mask) {
  return mask ? calc_ax_impl(ax, A, x, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_calc_r_mask(bcm_field_t r, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b,
                               bcm_field_t mask) {
  return mask ? calc_r_impl(r, A, x, b, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_calc_r2_mask(const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, bcm_field_t mask, double* rr) {
  return mask ? calc_r2_impl(A, x, b, rr, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_dot_mask(bcm_field_t x, bcm_field_t y, bcm_field_t mask, double* xy) {
  return mask ? dot_impl(x, y, xy, mask) : BCM_ERR_ARG;
}

static int blas_c(int op, bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, double b,
                  bcm_field_t mask = nullptr) {
  if (!same_mesh(z, x) || (y && !same_mesh(z, y)) || !mask_ok(z, mask)) return BCM_ERR_ARG;
  bcm_mesh* m = z->mesh;
  if (z->dtype == BCM_F64)
    return blas_raw<double>(m, op, static_cast<double*>(z->data), static_cast<const double*>(x->data),
                            y ? static_cast<const double*>(y->data) : nullptr, a, b, mask_ptr<double>(mask));
  return blas_raw<float>(m, op, static_cast<float*>(z->data), static_cast<const float*>(x->data),
                         y ? static_cast<const float*>(y->data) : nullptr, a, b, mask_ptr<float>(mask));
}

extern "C" int bcm_axpy(bcm_field_t y, bcm_field_t x, double a) { return blas_c(OP_AXPY, y, x, nullptr, a, 0.0); }
extern "C" int bcm_xpay(bcm_field_t y, bcm_field_t x, double a) { return blas_c(OP_XPAY, y, x, nullptr, a, 0.0); }
extern "C" int bcm_axpyz(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a) {
  if (!y) return BCM_ERR_ARG;
  return blas_c(OP_AXPYZ, z, x, y, a, 0.0);
}
extern "C" int bcm_axpbypz(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, double b) {
  if (!y) return BCM_ERR_ARG;
  return blas_c(OP_AXPBYPZ, z, x, y, a, b);
}
extern "C" int bcm_axpy_mask(bcm_field_t y, bcm_field_t x, double a, bcm_field_t mask) {
  return mask ? blas_c(OP_AXPY, y, x, nullptr, a, 0.0, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_xpay_mask(bcm_field_t y, bcm_field_t x, double a, bcm_field_t mask) {
  return mask ? blas_c(OP_XPAY, y, x, nullptr, a, 0.0, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_axpyz_mask(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, bcm_field_t mask) {
  return (mask && y) ? blas_c(OP_AXPYZ, z, x, y, a, 0.0, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_axpbypz_mask(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, double b, bcm_field_t mask) {
  return (mask && y) ? blas_c(OP_AXPBYPZ, z, x, y, a, b, mask) : BCM_ERR_ARG;
}
extern "C" int bcm_copy_mask(bcm_field_t y, bcm_field_t x, bcm_field_t mask) {
  return mask ? blas_c(OP_COPY_INTERIOR, y, x, nullptr, 0.0, 0.0, mask) : BCM_ERR_ARG;
}

extern "C" int bcm_fill(bcm_field_t x, double a) {
  if (!x) return BCM_ERR_ARG;
  if (x->dtype == BCM_F64) return fill_raw<double>(x->mesh, static_cast<double*>(x->data), a);
  return fill_raw<float>(x->mesh, static_cast<float*>(x->data), a);
}

extern "C" int bcm_setup_mask(bcm_field_t mask) {
  if (!mask) return BCM_ERR_ARG;
  bcm_mesh* m = mask->mesh;
  const long long count = (long long)m->nblocks * (long long)m->bvol;
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_BLAS1, &ev));
  const unsigned grid = unsigned((count + 255) / 256);
  if (mask->dtype == BCM_F64)
    hipLaunchKernelGGL(k_setup_mask<double>, dim3(grid), dim3(256), 0, m->stream, static_cast<double*>(mask->data),
                       m->n, m->vc, count);
  else
    hipLaunchKernelGGL(k_setup_mask<float>, dim3(grid), dim3(256), 0, m->stream, static_cast<float*>(mask->data),
                       m->n, m->vc, count);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_copy(bcm_field_t y, bcm_field_t x) {
  if (!same_mesh(y, x)) return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  const long long count = (long long)m->nblocks * (long long)m->bvol;
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_BLAS1, &ev));
  const unsigned grid = unsigned((count + 255) / 256);
  if (x->dtype == BCM_F64)
    hipLaunchKernelGGL(k_copy_all<double>, dim3(grid), dim3(256), 0, m->stream, static_cast<double*>(y->data),
                       static_cast<const double*>(x->data), count);
  else
    hipLaunchKernelGGL(k_copy_all<float>, dim3(grid), dim3(256), 0, m->stream, static_cast<float*>(y->data),
                       static_cast<const float*>(x->data), count);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_poisson_assemble(const bcm_field_t A[7], bcm_field_t b, bcm_field_t rhs, const double* bvalues,
                                    int nbvalues) {
  if (!same_mesh(b, rhs) || b->dtype != BCM_F64) return BCM_ERR_ARG;
  if (A && !check_A(b, A)) return BCM_ERR_ARG;
  bcm_mesh* m = b->mesh;
  BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_ASSEMBLE, &ev));
  auto d = [&](int i) { return A ? static_cast<double*>(A[i]->data) : nullptr; };
  hipLaunchKernelGGL(k_assemble, dim3(cell_grid(m)), dim3(kWG), 0, m->stream, d(0), d(1), d(2), d(3), d(4), d(5),
                     d(6), static_cast<double*>(b->data), static_cast<const double*>(rhs->data), m->d_meta,
                     m->d_bvalues, cell_geo(m));
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_jacobi7_poisson(bcm_field_t x1, bcm_field_t x0, bcm_field_t b, double omega) {
  if (!same_mesh(x1, x0) || !same_mesh(x1, b) || x1 == x0) return BCM_ERR_ARG;
  if (x1->dtype == BCM_F64)
    return jacobi_poisson_raw<double>(x1->mesh, static_cast<double*>(x1->data), static_cast<const double*>(x0->data),
                                      static_cast<const double*>(b->data), omega);
  return jacobi_poisson_raw<float>(x1->mesh, static_cast<float*>(x1->data), static_cast<const float*>(x0->data),
                                   static_cast<const float*>(b->data), omega);
}

extern "C" int bcm_calc_ax_poisson(bcm_field_t ax, bcm_field_t x) {
  if (!same_mesh(ax, x) || ax == x) return BCM_ERR_ARG;
  if (x->dtype == BCM_F64)
    return spmv_poisson_raw<double>(x->mesh, static_cast<double*>(ax->data), static_cast<const double*>(x->data));
  return spmv_poisson_raw<float>(x->mesh, static_cast<float*>(ax->data), static_cast<const float*>(x->data));
}

extern "C" int bcm_calc_r_poisson(bcm_field_t r, bcm_field_t x, bcm_field_t b) {
  if (!same_mesh(r, x) || !same_mesh(x, b) || r == x) return BCM_ERR_ARG;
  if (x->dtype == BCM_F64)
    return residual_poisson_raw<double>(x->mesh, static_cast<double*>(r->data), static_cast<const double*>(x->data),
                                        static_cast<const double*>(b->data), false, 0);
  return residual_poisson_raw<float>(x->mesh, static_cast<float*>(r->data), static_cast<const float*>(x->data),
                                     static_cast<const float*>(b->data), false, 0);
}
