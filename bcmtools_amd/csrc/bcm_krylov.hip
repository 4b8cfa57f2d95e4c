// Fused, device-resident BiCGSTAB over one mesh or a batch of meshes.
//
// Algorithm: ILS3D::BiCGSTAB (examples/SolverTest/ILS3D.h:663-844) with
// the Jacobi preconditioner (ILS3D.h:452-475, blas.f90:452-502) on the
// Poisson operator with the Dirichlet fold (bcut.f90:2848-2866,
// bca.f90:12-202), restated for HBM:
//   * every scalar (rho, alpha, gamma, beta, norms) lives in a per-segment
//     state record on the device, so an iteration never waits for the
//     host and a block of iterations can be captured in one hipGraph;
//   * vector updates are fused with the first preconditioner sweep and
//     the dot products with the kernels that produce their operands;
//   * Krylov vectors that no stencil reads are stored compact (interior
//     cells only, n^3 per block); the three stencil inputs (the two
//     preconditioner iterates and the sweep scratch) keep the ghosted
//     block layout and get their faces from a block-wise ghost gather;
//   * a segment is one independent mesh of a batch: kernels run over all
//     blocks of all segments, reductions and scalars stay per segment and
//     a converged segment's blocks drop out.
// Precision: FP64 runs the whole iteration in FP64.  MIXED stores the
// Krylov vectors r, r0, p, q, s, t and the sweep scratch in FP32 and runs
// the Jacobi sweeps and vector updates in FP32, but keeps in FP64 what an
// FP32 rounding would spoil: a smooth FP32 vector v carries a relative
// error eps32 that A multiplies by ~(12/h^2)/lambda_min (1e-2..1e-1 on a
// level-6/7 shell mesh), so the SpMV inputs p_ = M^-1 p and s_ = M^-1 s,
// the SpMV arithmetic, the solution correction dx and every dot product are
// FP64.  An FP64 update (x += dx; r = b - A x, rounded to FP32) replaces
// the recurrence residual when it has fallen by refresh_delta, every
// refresh_max iterations and when the recurrence claims convergence;
// convergence is only ever declared on the true FP64 residual
// ||b - A x|| / ||b|| <= tol.
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "bcm_internal.h"
#include "bcm_interp.h"
#include "bcm_poisson.h"

namespace bcm {

struct SegState {
  double bb;        // (b, b)
  double rho;       // (r, r0) of the current r
  double rho_prev;  // rho of the previous iteration
  double alpha, gamma;
  double rr;        // recurrence (r, r)
  double rr_ref;    // (r, r) right after the last reliable update
  double rr_true;   // true FP64 ||b - A x||^2 at the last reliable update
  double tol2;      // tol^2
  double delta2;    // refresh_delta^2 (0: off)
  double rr_min;    // smallest recurrence (r, r) since the last reliable update
  double rr_best;   // smallest true (r, r) so far
  int32_t status;   // BCM_SOLVE_*
  int32_t first;    // next iteration starts with p = r
  int32_t refresh;  // reliable update after this iteration
  int32_t restart;  // that update also restarts the Krylov space (r0 = r)
  int32_t reset_r0; // the next iteration's p update also sets r0 = r
  int32_t trig;     // why the pending reliable update was requested (TRIG_*)
  int32_t iters, since, max_iter, refresh_max, nref, zero_x;
  int32_t restart_every;  // restart at every k-th reliable update (0: never)
  int32_t stall;          // iterations without a new rr_min that force an update (0: off)
  int32_t since_min;      // iterations since rr_min last fell by 1 %
  int32_t stag_max;       // consecutive true checks without halving ||r|| that end the solve (0: off)
  int32_t n_stag;         // current run of such checks
};

enum { TRIG_CONV = 1, TRIG_SCHEDULE = 2, TRIG_LIMIT = 4, TRIG_BREAKDOWN = 8 };

namespace {

constexpr int KWG = 256;

template <int N>
struct Blk {
  static constexpr int NX = N + 2;
  static constexpr int N3 = N * N * N;
  static constexpr long long GV = (long long)NX * NX * NX;
  static constexpr int PER = N3 / KWG;  // cells per thread of the iteration kernels (N >= 8)
  __device__ static __forceinline__ void ijk(int c, int& i, int& j, int& k) {
    i = c % N;
    j = (c / N) % N;
    k = c / (N * N);
  }
  __device__ static __forceinline__ int g(int i, int j, int k) { return (i + 1) + NX * ((j + 1) + NX * (k + 1)); }
};

// Deterministic work-group sums of two doubles (wave64 butterfly, then a
// fixed-order sum of the four wave totals); the result is in thread 0.
__device__ __forceinline__ void wg_sum2(double& a, double& b, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[4 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = ((red[0] + red[1]) + red[2]) + red[3];
    b = ((red[4] + red[5]) + red[6]) + red[7];
  }
}

// The same for a work-group of NW waves (wave totals summed in wave order).
template <int NW>
__device__ __forceinline__ void wg_sum2n(double& a, double& b, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[NW + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = red[0];
    b = red[NW];
    for (int i = 1; i < NW; ++i) {
      a += red[i];
      b += red[NW + i];
    }
  }
}

// Gate of a block-wise kernel: 0 every block, 1 blocks of active
// segments, 2 blocks of active segments with a reliable update pending.
template <int GATE>
__device__ __forceinline__ bool gated_out(const int* __restrict__ blk_seg, const SegState* __restrict__ st, int b) {
  if (GATE == 0) return false;
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE) return true;
  return GATE == 2 && !sp->refresh;
}

// ---------------------------------------------------------------------
// Block-wise ghost gather: one work-group per destination block fills its
// six faces from the neighbours' interiors (copy, F2C, C2F as
// include/Scalar3DUpdater.h:190-266, 321-347, 519-708), the same values as
// the per-face task kernel.  Dirichlet / unconnected faces are left as
// they are.  `a` holds blocks first, first+1, ... (first = 0: the whole
// mesh; one segment's blocks otherwise: a segment's faces connect only
// blocks of that segment).
// ---------------------------------------------------------------------
template <typename T, int N, int GATE>
__global__ __launch_bounds__(KWG) void k_gather(T* __restrict__ a, const int* __restrict__ desc,
                                                const int* __restrict__ blk_seg, const SegState* __restrict__ st,
    const int* __restrict__ blist, int first) {
  using B = Blk<N>;
  const int b = blist ? blist[blockIdx.x] : first + int(blockIdx.x);
  if (gated_out<GATE>(blk_seg, st, b)) return;
  __shared__ int d[48];
  if (threadIdx.x < 48) d[threadIdx.x] = desc[b * 48 + threadIdx.x];
  __syncthreads();
  T* __restrict__ dst = a + (long long)(b - first) * B::GV;
  constexpr int h = N / 2;
  for (int e = threadIdx.x; e < 6 * N * N; e += KWG) {
    const int f = e / (N * N), uv = e % (N * N), u = uv % N, v = uv / N;
    const int* df = d + f * 8;
    const int kind = df[0];
    if (kind != FD_COPY && kind != FD_F2C && kind != FD_C2F) continue;
    const int ax = f >> 1;
    const bool plus = f & 1;
    const int t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
    int c[3];
    c[ax] = plus ? N : -1;
    c[t0] = u;
    c[t1] = v;
    T val;
    if (kind == FD_COPY) {
      const T* __restrict__ src = a + (long long)(df[1] - first) * B::GV;
      int sc[3] = {c[0], c[1], c[2]};
      sc[ax] = plus ? 0 : N - 1;
      val = src[B::g(sc[0], sc[1], sc[2])];
    } else if (kind == FD_F2C) {
      const int s0 = u >= h, s1 = v >= h;
      const T* __restrict__ src = a + (long long)(df[1 + s0 + 2 * s1] - first) * B::GV;
      int I[3];
      I[ax] = plus ? 0 : h - 1;
      I[t0] = u - h * s0;
      I[t1] = v - h * s1;
      val = f2c_sum<T>([&](int i, int j, int k) { return src[B::g(i, j, k)]; }, I[0], I[1], I[2]);
    } else {
      const int sf = df[5];
      const T* __restrict__ src = a + (long long)(df[1] - first) * B::GV;
      int p[3];
      p[ax] = plus ? 0 : 2 * N - 1;
      p[t0] = u + N * (sf & 1);
      p[t1] = v + N * (sf >> 1);
      val = c2f_interp<T>([&](int i, int j, int k) { return src[B::g(i, j, k)]; }, N, p[0], p[1], p[2]);
    }
    dst[B::g(c[0], c[1], c[2])] = val;
  }
}

// ---------------------------------------------------------------------
// Stencil vectors of the iteration (the preconditioned p_, s_ and the
// sweep scratch) are compact (n^3 per block, x fastest) with an export of
// columns 0, 1, n-2, n-1 of every block ([blk][4][k][j]) that the kernel
// producing the vector writes.  A kernel applying the stencil builds its
// block's ghost frame in LDS from the neighbours' rows (Y/Z faces) and
// exports (X faces) -- coalesced reads only -- with the exchange's own
// arithmetic (copy, F2C, C2F as include/Scalar3DUpdater.h:190-266,
// 321-347, 519-708), so it sees exactly the values a ghost fill would
// have written.  Frame cells of Dirichlet / unconnected faces are zero.
// ---------------------------------------------------------------------
template <typename T, int N>
struct Compact {
  const T* __restrict__ v;   // compact vector
  const T* __restrict__ xe;  // its column export
  __device__ __forceinline__ T at(int blk, int i, int j, int k) const {
    return v[(long long)blk * (N * N * N) + i + N * (j + N * k)];
  }
  // i must be 0, 1, N-2 or N-1.
  __device__ __forceinline__ T x_at(int blk, int i, int j, int k) const {
    const int col = i < 2 ? i : i - (N - 4);
    return xe[((long long)(blk * 4 + col) * N + k) * N + j];
  }
};

// Ghost value of face f at in-face position (p, q) of the block whose face
// descriptors are df = desc + 8 f, computed from src with the exchange's
// arithmetic; c receives the ghost cell (i, j, k).  X faces p = j, q = k;
// Y faces p = i, q = k; Z faces p = i, q = j.  Faces without a neighbour
// (outer boundaries) give 0: their folded coefficient is 0.
// src is any accessor with at(blk, i, j, k) and, for the X faces, x_at()
// returning T (Compact, or the on-the-fly sweep values of SweepOne).
template <typename T, int N, typename S>
__device__ __forceinline__ T frame_value(const S& src, const int* df, int f, int p, int q, int (&c)[3]) {
  constexpr int h = N / 2;
  const int ax = f >> 1;
  const bool plus = f & 1;
  const int kind = df[0];
  c[ax] = plus ? N : -1;
  c[ax == 0 ? 1 : 0] = p;
  c[ax == 2 ? 1 : 2] = q;
  const int t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
  T val = T(0);
  if (kind == FD_COPY) {
    int sc[3] = {c[0], c[1], c[2]};
    sc[ax] = plus ? 0 : N - 1;
    val = ax == 0 ? src.x_at(df[1], sc[0], sc[1], sc[2]) : src.at(df[1], sc[0], sc[1], sc[2]);
  } else if (kind == FD_F2C) {
    const int s0 = c[t0] >= h, s1 = c[t1] >= h;
    const int fb = df[1 + s0 + 2 * s1];
    int I[3];
    I[ax] = plus ? 0 : h - 1;
    I[t0] = c[t0] - h * s0;
    I[t1] = c[t1] - h * s1;
    if (ax == 0)
      val = f2c_sum<T>([&](int i, int j, int k) { return src.x_at(fb, i, j, k); }, I[0], I[1], I[2]);
    else
      val = f2c_sum<T>([&](int i, int j, int k) { return src.at(fb, i, j, k); }, I[0], I[1], I[2]);
  } else if (kind == FD_C2F) {
    const int sf = df[5], cb = df[1];
    int pp[3];
    pp[ax] = plus ? 0 : 2 * N - 1;
    pp[t0] = c[t0] + N * (sf & 1);
    pp[t1] = c[t1] + N * (sf >> 1);
    if (ax == 0)
      val = c2f_interp<T>([&](int i, int j, int k) { return src.x_at(cb, i, j, k); }, N, pp[0], pp[1], pp[2]);
    else
      val = c2f_interp<T>([&](int i, int j, int k) { return src.at(cb, i, j, k); }, N, pp[0], pp[1], pp[2]);
  }
  return val;
}

// fr[f][p + N q], the six faces of the ghost frame.
template <typename T, int N>
__device__ __forceinline__ void build_frame(const Compact<T, N>& src, const int* d, T (*fr)[N * N]) {
  for (int e = threadIdx.x; e < 6 * N * N; e += KWG) {
    const int f = e / (N * N), pq = e % (N * N);
    int c[3];
    fr[f][pq] = frame_value<T, N>(src, d + f * 8, f, pq % N, pq / N, c);
  }
}

// A block with its face ghosts in LDS, (N+2)^3 values in Index3DS order.
template <int N>
struct Gh {
  static constexpr int NX = N + 2;
  static constexpr int V = NX * NX * NX;
  __device__ static __forceinline__ int at(int i, int j, int k) { return (i + 1) + NX * ((j + 1) + NX * (k + 1)); }
};

// Stage block b of src in LDS: the interior from the compact vector (loads
// issued first, one coalesced pass), then the face ghosts from the
// neighbours (descriptors read from global memory, so no barrier sits
// between the two sets of loads).  The caller synchronises.
template <typename T, int N, int W>
__device__ __forceinline__ void stage_block(const Compact<T, N>& src, const int* __restrict__ d, int b, T* lds) {
  using B = Blk<N>;
  constexpr int PER = B::N3 / W;
  T v[PER];
  const long long cb = (long long)b * B::N3;
#pragma unroll
  for (int m = 0; m < PER; ++m) v[m] = src.v[cb + threadIdx.x + W * m];
  for (int e = threadIdx.x; e < 6 * N * N; e += W) {
    const int f = e / (N * N), pq = e % (N * N);
    int c[3];
    const T val = frame_value<T, N>(src, d + f * 8, f, pq % N, pq / N, c);
    lds[Gh<N>::at(c[0], c[1], c[2])] = val;
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    lds[Gh<N>::at(i, j, k)] = v[m];
  }
}

// The six neighbours of cell (i, j, k) of block b: interior ones from the
// compact vector, the others from the frame.
template <typename T, int N, typename TV>
__device__ __forceinline__ void neighbours(const TV* __restrict__ g, long long o, int i, int j, int k,
                                           const TV (*fr)[N * N], TV& w, TV& e, TV& s, TV& n, TV& bo, TV& t) {
  w = i > 0 ? g[o - 1] : fr[0][j + N * k];
  e = i < N - 1 ? g[o + 1] : fr[1][j + N * k];
  s = j > 0 ? g[o - N] : fr[2][i + N * k];
  n = j < N - 1 ? g[o + N] : fr[3][i + N * k];
  bo = k > 0 ? g[o - N * N] : fr[4][i + N * j];
  t = k < N - 1 ? g[o + N * N] : fr[5][i + N * j];
}

// Export staging: a producer puts columns 0, 1, N-2, N-1 of its block into
// LDS as it computes them and writes them out in one coalesced pass.
template <int N>
__device__ __forceinline__ int export_col(int i) {
  return i < 2 ? i : (i >= N - 2 ? i - (N - 4) : -1);
}
template <typename T, int N>
__device__ __forceinline__ void export_flush(const T* xs, T* __restrict__ xe, int b) {
  __syncthreads();
  T* dst = xe + (long long)b * 4 * N * N;
  for (int e = threadIdx.x; e < 4 * N * N; e += KWG) dst[e] = xs[e];
}

// ---------------------------------------------------------------------
// Iteration kernels (one work-group per block, cell c = tid + 256 m)
// ---------------------------------------------------------------------
// Sweep 1 of the Jacobi preconditioner from a zero start, written as
// jacobi_smoother does it: r = b (the neighbours are zero), x = 0 - w (0 - r/Ap).
template <typename T>
__device__ __forceinline__ T first_sweep(T bv, T ap, T om) {
  const T xt = bv / ap;
  const T xc = T(0);
  return xc - om * (xc - xt);
}

// p = r (first) or p = r + beta (p - gamma q)  [AXPY(p, q, -gamma); XPAY(p, r, beta)],
// then the first preconditioner sweep of p into g (PRE) or g = p, with g's
// column export.  After a restart the same pass sets the shadow residual r0 = r.
template <typename T, int N, int PRE, typename TG>
__global__ __launch_bounds__(KWG) void k_kry_p(T* __restrict__ p, const T* __restrict__ r, const T* __restrict__ q,
                                               T* __restrict__ r0, TG* __restrict__ g, TG* __restrict__ gx,
                                               const BlockMeta* __restrict__ meta, const int* __restrict__ blk_seg,
                                               const SegState* __restrict__ st, double omega,
                                               const int* __restrict__ blist) {
  using B = Blk<N>;
  __shared__ TG xs[4 * N * N];
  const int b = blist[blockIdx.x];
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE) return;
  const bool first = sp->first != 0;
  const bool reset = sp->reset_r0 != 0;
  const double beta = first ? 0.0 : sp->rho / sp->rho_prev * sp->alpha / sp->gamma;
  const T tb = T(beta), tg = T(-sp->gamma), om = T(omega);
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    const T rv = r[o];
    const T pv = first ? rv : rv + tb * (p[o] + tg * q[o]);
    p[o] = pv;
    if (reset) r0[o] = rv;
    int i, j, k;
    B::ijk(c, i, j, k);
    TG gv;
    if (PRE) {
      const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
      gv = TG(first_sweep(pv, a.p, om));
    } else {
      gv = TG(pv);
    }
    g[o] = gv;
    const int col = export_col<N>(i);
    if (col >= 0) xs[(col * N + k) * N + j] = gv;
  }
  export_flush<TG, N>(xs, gx, b);
}

// s = r - alpha q  [AXPYZ(s, q, r, -alpha)], then the first sweep of s into g.
template <typename T, int N, int PRE, typename TG>
__global__ __launch_bounds__(KWG) void k_kry_s(T* __restrict__ s, const T* __restrict__ r, const T* __restrict__ q,
                                               TG* __restrict__ g, TG* __restrict__ gx,
                                               const BlockMeta* __restrict__ meta, const int* __restrict__ blk_seg,
                                               const SegState* __restrict__ st, double omega,
                                               const int* __restrict__ blist) {
  using B = Blk<N>;
  __shared__ TG xs[4 * N * N];
  const int b = blist[blockIdx.x];
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE) return;
  const T ta = T(-sp->alpha), om = T(omega);
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    const T sv = ta * q[o] + r[o];
    s[o] = sv;
    int i, j, k;
    B::ijk(c, i, j, k);
    TG gv;
    if (PRE) {
      const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
      gv = TG(first_sweep(sv, a.p, om));
    } else {
      gv = TG(sv);
    }
    g[o] = gv;
    const int col = export_col<N>(i);
    if (col >= 0) xs[(col * N + k) * N + j] = gv;
  }
  export_flush<TG, N>(xs, gx, b);
}

// Later preconditioner sweeps: gout = jacobi_smoother(gin; b = rhs),
// arithmetic in T (the sweep precision), input / output stored as TI / TO.
template <typename T, int N, typename TI, typename TO>
__global__ __launch_bounds__(KWG) void k_kry_jacobi(TO* __restrict__ gout, TO* __restrict__ gox,
                                                    const TI* __restrict__ gin, const TI* __restrict__ gix,
                                                    const T* __restrict__ rhs, const BlockMeta* __restrict__ meta,
                                                    const int* __restrict__ blk_seg,
                                                    const SegState* __restrict__ st, double omega,
                                                    const int* __restrict__ blist, const int* __restrict__ desc) {
  using B = Blk<N>;
  __shared__ TI fr[6][N * N];
  __shared__ TO xs[4 * N * N];
  __shared__ int d[48];
  const int b = blist[blockIdx.x];
  if (st[blk_seg[b]].status != BCM_SOLVE_ACTIVE) return;
  if (threadIdx.x < 48) d[threadIdx.x] = desc[b * 48 + threadIdx.x];
  __syncthreads();
  build_frame<TI, N>(Compact<TI, N>{gin, gix}, d, fr);
  __syncthreads();
  const T om = T(omega);
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    int i, j, k;
    B::ijk(c, i, j, k);
    TI w, e, s, n, bo, t;
    neighbours<T, N, TI>(gin, o, i, j, k, fr, w, e, s, n, bo, t);
    const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
    const TO v = TO(jacobi_vals(a, T(w), T(e), T(s), T(n), T(bo), T(t), T(gin[o]), rhs[o], om));
    gout[o] = v;
    const int col = export_col<N>(i);
    if (col >= 0) xs[(col * N + k) * N + j] = v;
  }
  export_flush<TO, N>(xs, gox, b);
}

// The same sweep for N <= 16 with the whole ghosted block staged in LDS,
// W threads per block: the interior and frame loads are all in flight
// together, and the seven stencil values of every cell come from LDS.  The
// outputs are then staged in the same LDS space so the column export is
// written coalesced.
// Threads per block of the LDS-staged kernels (at most one per cell),
// the faster of 256/512/1024 measured on the level-6 shell mesh: the
// sweep at 512 (176 us vs 180 us at 1024), the SpMV at 1024 (143 us vs
// 159 us at 512).
template <int N, int W>
constexpr int kwg_cap() {
  return N * N * N < W ? N * N * N : W;
}
template <int N>
constexpr int kwg_jac() {
  return kwg_cap<N, 512>();
}
template <int N>
constexpr int kwg_spmv() {
  return kwg_cap<N, 1024>();
}

template <typename T, int N, typename TI, typename TO, int W>
__global__ __launch_bounds__(W) __attribute__((amdgpu_waves_per_eu(W / 128))) void k_kry_jacobi_l(TO* __restrict__ gout, TO* __restrict__ gox,
                                                    const TI* __restrict__ gin, const TI* __restrict__ gix,
                                                    const T* __restrict__ rhs, const BlockMeta* __restrict__ meta,
                                                    const int* __restrict__ blk_seg,
                                                    const SegState* __restrict__ st, double omega,
                                                    const int* __restrict__ blist, const int* __restrict__ desc) {
  using B = Blk<N>;
  using G = Gh<N>;
  constexpr int PER = B::N3 / W;
  constexpr int ES = sizeof(TI) > sizeof(TO) ? sizeof(TI) : sizeof(TO);
  __shared__ __align__(16) char raw[G::V * ES];
  TI* lin = reinterpret_cast<TI*>(raw);
  TO* lout = reinterpret_cast<TO*>(raw);
  const int b = blist[blockIdx.x];
  if (st[blk_seg[b]].status != BCM_SOLVE_ACTIVE) return;
  const long long cb = (long long)b * B::N3;
  T bv[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) bv[m] = rhs[cb + threadIdx.x + W * m];
  stage_block<TI, N, W>(Compact<TI, N>{gin, gix}, desc + b * 48, b, lin);
  __syncthreads();
  const T om = T(omega);
  const BlockMeta me = meta[b];
  TO out[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    const int a0 = G::at(i, j, k);
    const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
    out[m] = TO(jacobi_vals(a, T(lin[a0 - 1]), T(lin[a0 + 1]), T(lin[a0 - G::NX]), T(lin[a0 + G::NX]),
                            T(lin[a0 - G::NX * G::NX]), T(lin[a0 + G::NX * G::NX]), T(lin[a0]), bv[m], om));
    gout[cb + threadIdx.x + W * m] = out[m];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    if (export_col<N>(i) >= 0) lout[G::at(i, j, k)] = out[m];
  }
  __syncthreads();
  TO* dst = gox + (long long)b * 4 * N * N;
  for (int e = threadIdx.x; e < 4 * N * N; e += W) {
    const int col = e / (N * N), k = (e / N) % N, j = e % N;
    const int i = col < 2 ? col : col + N - 4;
    dst[e] = lout[G::at(i, j, k)];
  }
}

// k_kry_spmv for N <= 16 with the ghosted block staged in LDS; with OX the
// column export of `out` too (q, read by the fused update kernels).
template <typename T, int N, int NDOT, int W, bool OX = false>
__global__ __launch_bounds__(W) __attribute__((amdgpu_waves_per_eu(W / 128))) void k_kry_spmv_l(T* __restrict__ out, const double* __restrict__ g,
                                                  const double* __restrict__ gx, const T* __restrict__ y,
                                                  const BlockMeta* __restrict__ meta,
                                                  const int* __restrict__ blk_seg, const SegState* __restrict__ st,
                                                  double* __restrict__ parts, const int* __restrict__ blist,
                                                  const int* __restrict__ desc, T* __restrict__ ox = nullptr) {
  using B = Blk<N>;
  using G = Gh<N>;
  constexpr int PER = B::N3 / W;
  __shared__ __align__(16) double lin[G::V];
  __shared__ double red[2 * (W / 64)];
  const int b = blist[blockIdx.x];
  if (st[blk_seg[b]].status != BCM_SOLVE_ACTIVE) return;
  const long long cb = (long long)b * B::N3;
  stage_block<double, N, W>(Compact<double, N>{g, gx}, desc + b * 48, b, lin);
  __syncthreads();
  const BlockMeta me = meta[b];
  double d0 = 0.0, d1 = 0.0;
  T ov[OX ? PER : 1];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    const int a0 = G::at(i, j, k);
    const Coef7<double> a = poisson_coef<double>(me, i, j, k, N);
    const T v = T(ax_vals(a, lin[a0 - 1], lin[a0 + 1], lin[a0 - G::NX], lin[a0 + G::NX], lin[a0 - G::NX * G::NX],
                          lin[a0 + G::NX * G::NX], lin[a0]));
    out[cb + threadIdx.x + W * m] = v;
    if (OX) ov[m] = v;
    d0 += double(v) * double(y[cb + threadIdx.x + W * m]);
    if (NDOT == 2) d1 += double(v) * double(v);
  }
  if (OX) {
    T* xs = reinterpret_cast<T*>(lin);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      int i, j, k;
      B::ijk(threadIdx.x + W * m, i, j, k);
      const int col = export_col<N>(i);
      if (col >= 0) xs[(col * N + k) * N + j] = ov[m];
    }
    __syncthreads();
    T* dst = ox + (long long)b * 4 * N * N;
    for (int e = threadIdx.x; e < 4 * N * N; e += W) dst[e] = xs[e];
  }
  wg_sum2n<W / 64>(d0, d1, red);
  if (threadIdx.x == 0) {
    parts[2 * b] = d0;
    parts[2 * b + 1] = d1;
  }
}

// ---------------------------------------------------------------------
// Two-sweep preconditioner fused into the vector update (N <= 16,
// count_pre == 2).  The update kernel of p (or s) computes the new vector,
// its first Jacobi sweep (pointwise: the start is zero) and, in the same
// pass, the second sweep, which needs the first sweep's values on the
// block's ghost frame.  Those are functions of the neighbours' r, p, q
// (s: r, q), read from the stored vectors (Y/Z faces) and their column
// exports (X faces), so each frame value is recomputed exactly as the
// neighbour's own kernel computes it and then combined with the
// exchange's arithmetic.  This removes the sweep scratch vector and the
// separate sweep kernel.  p is read by the neighbours while it is
// replaced, so it alternates between two buffers (the iteration count's
// parity picks the one holding the current p); s is not read by them.
// ---------------------------------------------------------------------
// The update of the vector at one cell, as k_kry_p / k_kry_s write it:
// p = r (first) or r + c1 (p + c2 q) with c1 = beta, c2 = -gamma;
// s = c1 q + r with c1 = -alpha.
template <typename T, bool IS_P>
__device__ __forceinline__ T updated(bool first, T c1, T c2, T rv, T pv, T qv) {
  if (IS_P) return first ? rv : rv + c1 * (pv + c2 * qv);
  return c1 * qv + rv;
}

// First-sweep value of the updated vector at any cell of any block.
template <typename T, int N, bool IS_P>
struct SweepOne {
  const T* __restrict__ r;
  const T* __restrict__ rx;
  const T* __restrict__ p;  // IS_P only
  const T* __restrict__ px;
  const T* __restrict__ q;
  const T* __restrict__ qx;
  const BlockMeta* __restrict__ meta;
  T c1, c2, om;
  bool first;
  __device__ __forceinline__ T sweep(int blk, int i, int j, int k, T v) const {
    const Coef7<T> a = poisson_coef<T>(meta[blk], i, j, k, N);
    return first_sweep(v, a.p, om);
  }
  __device__ __forceinline__ T at(int blk, int i, int j, int k) const {
    const long long o = (long long)blk * (N * N * N) + i + N * (j + N * k);
    return sweep(blk, i, j, k, updated<T, IS_P>(first, c1, c2, r[o], IS_P ? p[o] : T(0), q[o]));
  }
  // i must be 0, 1, N-2 or N-1.
  __device__ __forceinline__ T x_at(int blk, int i, int j, int k) const {
    const int col = i < 2 ? i : i - (N - 4);
    const long long e = ((long long)(blk * 4 + col) * N + k) * N + j;
    return sweep(blk, i, j, k, updated<T, IS_P>(first, c1, c2, rx[e], IS_P ? px[e] : T(0), qx[e]));
  }
};

// v0/v1: p's two buffers (IS_P; read v[par], write v[1 - par]) or s (v0);
// vx0/vx1 their column exports.  g, gx: the preconditioned vector (FP64)
// and its export, as k_kry_jacobi_l writes them.
template <typename T, int N, int W, bool IS_P>
__global__ __launch_bounds__(W) __attribute__((amdgpu_waves_per_eu(W / 128))) void k_kry_upre2(
    T* __restrict__ v0, T* __restrict__ v1, T* __restrict__ vx0, T* __restrict__ vx1, const T* __restrict__ r,
    const T* __restrict__ rx, const T* __restrict__ q, const T* __restrict__ qx, T* __restrict__ r0,
    double* __restrict__ g, double* __restrict__ gx, const BlockMeta* __restrict__ meta,
    const int* __restrict__ blk_seg, const SegState* __restrict__ st, double omega, const int* __restrict__ blist,
    const int* __restrict__ desc) {
  using B = Blk<N>;
  using G = Gh<N>;
  constexpr int PER = B::N3 / W;
  constexpr int LB = G::V * int(sizeof(T));
  constexpr int XB = 4 * N * N * int(sizeof(double) + sizeof(T));
  __shared__ __align__(16) char raw[LB > XB ? LB : XB];
  T* lin = reinterpret_cast<T*>(raw);
  const int b = blist[blockIdx.x];
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE) return;
  const int par = IS_P ? (sp->iters & 1) : 0;
  const T* vin = par ? v1 : v0;
  T* vout = IS_P ? (par ? v0 : v1) : v0;
  const T* vxin = par ? vx1 : vx0;
  T* vxout = par ? vx0 : vx1;
  const bool first = IS_P && sp->first != 0;
  const bool reset = IS_P && sp->reset_r0 != 0;
  T c1, c2 = T(0);
  if (IS_P) {
    const double beta = first ? 0.0 : sp->rho / sp->rho_prev * sp->alpha / sp->gamma;
    c1 = T(beta);
    c2 = T(-sp->gamma);
  } else {
    c1 = T(-sp->alpha);
  }
  const T om = T(omega);
  const SweepOne<T, N, IS_P> src{r, rx, vin, vxin, q, qx, meta, c1, c2, om, first};
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3;
  // The frame first, then the own cells: the update (kept for the second
  // sweep's right-hand side) and its first sweep into LDS.
  const int* d = desc + b * 48;
  for (int e = threadIdx.x; e < 6 * N * N; e += W) {
    const int f = e / (N * N), pq = e % (N * N);
    int c[3];
    const T val = frame_value<T, N>(src, d + f * 8, f, pq % N, pq / N, c);
    lin[G::at(c[0], c[1], c[2])] = val;
  }
  T vv[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const long long o = cb + threadIdx.x + W * m;
    const T rv = r[o];
    vv[m] = updated<T, IS_P>(first, c1, c2, rv, IS_P ? vin[o] : T(0), q[o]);
    vout[o] = vv[m];
    if (reset) r0[o] = rv;
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
    lin[G::at(i, j, k)] = first_sweep(vv[m], a.p, om);
  }
  __syncthreads();
  // Second sweep (jacobi_smoother with b = the updated vector).
  double out[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    const int a0 = G::at(i, j, k);
    const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
    out[m] = double(jacobi_vals(a, lin[a0 - 1], lin[a0 + 1], lin[a0 - G::NX], lin[a0 + G::NX],
                                lin[a0 - G::NX * G::NX], lin[a0 + G::NX * G::NX], lin[a0], vv[m], om));
    g[cb + threadIdx.x + W * m] = out[m];
  }
  __syncthreads();
  // Column exports: the preconditioned vector and, for p, the new p.
  double* xg = reinterpret_cast<double*>(raw);
  T* xv = reinterpret_cast<T*>(raw + 4 * N * N * sizeof(double));
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    int i, j, k;
    B::ijk(threadIdx.x + W * m, i, j, k);
    const int col = export_col<N>(i);
    if (col >= 0) {
      xg[(col * N + k) * N + j] = out[m];
      if (IS_P) xv[(col * N + k) * N + j] = vv[m];
    }
  }
  __syncthreads();
  const long long xb = (long long)b * 4 * N * N;
  for (int e = threadIdx.x; e < 4 * N * N; e += W) {
    gx[xb + e] = xg[e];
    if (IS_P) vxout[xb + e] = xv[e];
  }
}

// out = A g (calc_ax) in FP64 from the FP64 preconditioned vector g, stored
// in T, with the per-block partial dots (out, y) and, for NDOT == 2,
// (out, out), formed and summed in FP64.
template <typename T, int N, int NDOT>
__global__ __launch_bounds__(KWG) void k_kry_spmv(T* __restrict__ out, const double* __restrict__ g,
                                                  const double* __restrict__ gx, const T* __restrict__ y,
                                                  const BlockMeta* __restrict__ meta,
                                                  const int* __restrict__ blk_seg, const SegState* __restrict__ st,
                                                  double* __restrict__ parts, const int* __restrict__ blist,
                                                  const int* __restrict__ desc) {
  using B = Blk<N>;
  __shared__ double fr[6][N * N];
  __shared__ double red[8];
  __shared__ int d[48];
  const int b = blist[blockIdx.x];
  if (st[blk_seg[b]].status != BCM_SOLVE_ACTIVE) return;
  if (threadIdx.x < 48) d[threadIdx.x] = desc[b * 48 + threadIdx.x];
  __syncthreads();
  build_frame<double, N>(Compact<double, N>{g, gx}, d, fr);
  __syncthreads();
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3;
  double d0 = 0.0, d1 = 0.0;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    int i, j, k;
    B::ijk(c, i, j, k);
    double w, e, s, n, bo, t;
    neighbours<double, N, double>(g, o, i, j, k, fr, w, e, s, n, bo, t);
    const Coef7<double> a = poisson_coef<double>(me, i, j, k, N);
    const T v = T(ax_vals(a, w, e, s, n, bo, t, g[o]));
    out[o] = v;
    d0 += double(v) * double(y[o]);
    if (NDOT == 2) d1 += double(v) * double(v);
  }
  wg_sum2(d0, d1, red);
  if (threadIdx.x == 0) {
    parts[2 * b] = d0;
    parts[2 * b + 1] = d1;
  }
}

// dx += alpha p_ + gamma s_  [AXPBYPZ(x, p_, s_, alpha, gamma)] in FP64;
// r = s - gamma t  [AXPYZ(r, t_, s, -gamma)] in T; partial (r, r), (r, r0).
// With RX, also r's column export (read by the fused update kernels).
template <typename T, int N, bool RX = false>
__global__ __launch_bounds__(KWG) void k_kry_xr(double* __restrict__ dx, const double* __restrict__ gp,
                                                const double* __restrict__ gs, T* __restrict__ r,
                                                const T* __restrict__ s, const T* __restrict__ t,
                                                const T* __restrict__ r0, const int* __restrict__ blk_seg,
                                                const SegState* __restrict__ st, double* __restrict__ parts,
    const int* __restrict__ blist, T* __restrict__ rx = nullptr) {
  using B = Blk<N>;
  __shared__ double red[8];
  __shared__ T xs[RX ? 4 * N * N : 1];
  const int b = blist[blockIdx.x];
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE) return;
  const double al = sp->alpha, ga = sp->gamma;
  const T tmg = T(-sp->gamma);
  const long long cb = (long long)b * B::N3;
  double d0 = 0.0, d1 = 0.0;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    dx[o] = (al * gp[o] + ga * gs[o]) + dx[o];
    const T rv = tmg * t[o] + s[o];
    r[o] = rv;
    if (RX) {
      int i, j, k;
      B::ijk(c, i, j, k);
      const int col = export_col<N>(i);
      if (col >= 0) xs[(col * N + k) * N + j] = rv;
    }
    d0 += double(rv) * double(rv);
    d1 += double(rv) * double(r0[o]);
  }
  if (RX) export_flush<T, N>(xs, rx, b);
  wg_sum2(d0, d1, red);
  if (threadIdx.x == 0) {
    parts[2 * b] = d0;
    parts[2 * b + 1] = d1;
  }
}

// ---------------------------------------------------------------------
// Reliable update (FP64) and set-up kernels
// ---------------------------------------------------------------------
// x += dx (interior), dx = 0.
template <int N>
__global__ __launch_bounds__(KWG) void k_ref_x(double* __restrict__ x, double* __restrict__ dx,
                                               const int* __restrict__ blk_seg, const SegState* __restrict__ st,
    const int* __restrict__ blist) {
  using B = Blk<N>;
  const int b = blist[blockIdx.x];
  if (gated_out<2>(blk_seg, st, b)) return;
  const long long cb = (long long)b * B::N3, gb = (long long)b * B::GV;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    int i, j, k;
    B::ijk(c, i, j, k);
    const long long go = gb + B::g(i, j, k);
    x[go] = x[go] + dx[cb + c];
    dx[cb + c] = 0.0;
  }
}

// r = T(b - A x) in FP64 (calc_r).  Partials: FP64 (r, r) of the
// unrounded residual and (r, r0) of the rounded one.
// With RX, also r's column export.
template <typename T, int N, bool RX = false>
__global__ __launch_bounds__(KWG) void k_ref_r(T* __restrict__ r, const T* __restrict__ r0,
                                               const double* __restrict__ x, const double* __restrict__ b64,
                                               const BlockMeta* __restrict__ meta, const int* __restrict__ blk_seg,
                                               const SegState* __restrict__ st, double* __restrict__ parts,
    const int* __restrict__ blist, T* __restrict__ rx = nullptr) {
  using B = Blk<N>;
  __shared__ double red[8];
  __shared__ T xs[RX ? 4 * N * N : 1];
  const int b = blist[blockIdx.x];
  const SegState* sp = st + blk_seg[b];
  if (sp->status != BCM_SOLVE_ACTIVE || !sp->refresh) return;
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3, gb = (long long)b * B::GV;
  double acc = 0.0, d1 = 0.0;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    const long long o = cb + c;
    int i, j, k;
    B::ijk(c, i, j, k);
    const Coef7<double> a = poisson_coef<double>(me, i, j, k, N);
    const double rv = r_cell(a, x, gb + B::g(i, j, k), B::NX, B::NX * B::NX, b64[o]);
    const T rt = T(rv);
    r[o] = rt;
    if (RX) {
      const int col = export_col<N>(i);
      if (col >= 0) xs[(col * N + k) * N + j] = rt;
    }
    acc += rv * rv;
    d1 += double(rt) * double(r0[o]);
  }
  if (RX) export_flush<T, N>(xs, rx, b);
  wg_sum2(acc, d1, red);
  if (threadIdx.x == 0) {
    parts[2 * b] = acc;
    parts[2 * b + 1] = d1;
  }
}

// Folded right-hand side, compact (k_assemble's b, bca.f90 folds),
// and the per-block (b, b).
template <int N>
__global__ __launch_bounds__(KWG) void k_kry_rhs(double* __restrict__ b64, const double* __restrict__ rhs,
                                                 const BlockMeta* __restrict__ meta, const double* __restrict__ bval,
                                                 double* __restrict__ parts) {
  using B = Blk<N>;
  __shared__ double red[8];
  const int b = blockIdx.x;
  const BlockMeta me = meta[b];
  const long long cb = (long long)b * B::N3, gb = (long long)b * B::GV;
  double acc = 0.0;
  for (int m = 0; m < B::PER; ++m) {
    const int c = threadIdx.x + KWG * m;
    int i, j, k;
    B::ijk(c, i, j, k);
    const double bv = fold_rhs(me, i, j, k, N, rhs[gb + B::g(i, j, k)], bval);
    b64[cb + c] = bv;
    acc += bv * bv;
  }
  double z = 0.0;
  wg_sum2(acc, z, red);
  if (threadIdx.x == 0) {
    parts[2 * b] = acc;
    parts[2 * b + 1] = 0.0;
  }
}

// A segment whose right-hand side is zero has the solution zero.
__global__ __launch_bounds__(KWG) void k_zero_x(double* __restrict__ x, long long bvol,
                                                const int* __restrict__ blk_seg, const SegState* __restrict__ st) {
  const int b = blockIdx.x;
  if (!st[blk_seg[b]].zero_x) return;
  for (long long e = threadIdx.x; e < bvol; e += KWG) x[b * bvol + e] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_kry_zero(T* __restrict__ x, long long count) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) x[i] = T(0);
}

// ---------------------------------------------------------------------
// Per-segment reductions and scalar updates (one work-group per segment,
// blocks of the segment summed in a fixed order, so a segment's scalars
// do not depend on the other segments of the batch).
// ---------------------------------------------------------------------
enum Phase { PH_BB = 0, PH_Q, PH_T, PH_R, PH_F };

__device__ __forceinline__ bool conv_test(double rr, const SegState& s) {
  return s.bb > double(FLT_MIN) ? rr <= s.tol2 * s.bb : rr <= s.tol2;
}

template <int PH>
__global__ __launch_bounds__(KWG) void k_seg(SegState* __restrict__ st, const double* __restrict__ parts,
                                             const int* __restrict__ seg_first) {
  __shared__ double red[8];
  const int sg = blockIdx.x;
  SegState* sp = st + sg;
  if (sp->status != BCM_SOLVE_ACTIVE) return;
  if (PH == PH_F && !sp->refresh) return;
  // Each thread sums its blocks in block order; the loads of UNR blocks are
  // issued together and then added one after the other, so the sum is the
  // same as a plain loop's while the loads overlap.
  constexpr int UNR = 8;
  const double2* pp = reinterpret_cast<const double2*>(parts);
  const int last = seg_first[sg + 1];
  double a = 0.0, c = 0.0;
  int b = seg_first[sg] + threadIdx.x;
  for (; b + (UNR - 1) * KWG < last; b += UNR * KWG) {
    double2 v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) v[u] = pp[b + u * KWG];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      a += v[u].x;
      c += v[u].y;
    }
  }
  for (; b < last; b += KWG) {
    const double2 v = pp[b];
    a += v.x;
    c += v.y;
  }
  wg_sum2(a, c, red);
  if (threadIdx.x != 0) return;
  SegState s = *sp;
  if (PH == PH_BB) {
    s.bb = a;
    if (a == 0.0) {
      s.status = BCM_SOLVE_CONVERGED;
      s.zero_x = 1;
    }
  } else if (PH == PH_Q) {
    // alpha = rho / (q_, r0) (ILS3D.h:756); a breakdown restarts.
    s.reset_r0 = 0;
    double al = s.rho / a;
    if (!std::isfinite(al)) {
      al = 0.0;
      s.restart = 1;
      s.refresh = 1;
      s.trig |= TRIG_BREAKDOWN;
    }
    s.alpha = al;
  } else if (PH == PH_T) {
    // gamma = (t_, s) / (t_, t_) (ILS3D.h:790-793).
    double ga = a / c;
    if (!std::isfinite(ga) || ga == 0.0) {
      ga = std::isfinite(ga) ? ga : 0.0;
      s.restart = 1;
      s.refresh = 1;
      s.trig |= TRIG_BREAKDOWN;
    }
    s.gamma = ga;
  } else if (PH == PH_R) {
    s.iters += 1;
    s.since += 1;
    if (!std::isfinite(a) || !std::isfinite(c)) {
      s.status = BCM_SOLVE_FAILED;
      s.refresh = 0;
    } else {
      s.rr = a;
      s.rho_prev = s.rho;
      s.rho = c;
      s.first = 0;
      if (std::fabs(c) < double(FLT_MIN)) {
        s.restart = 1;
        s.refresh = 1;
        s.trig |= TRIG_BREAKDOWN;
      }
      // The recurrence is asked for half the tolerance before the true
      // residual is checked, as the host-driven FP64 path does.
      if (conv_test(4.0 * a, s)) s.trig |= TRIG_CONV;
      // A stalled low-precision recurrence (no new minimum for `stall`
      // iterations) gains nothing more before an FP64 update.
      if (a < 0.98 * s.rr_min) {
        s.rr_min = a;
        s.since_min = 0;
      } else {
        s.since_min += 1;
      }
      if (s.since >= s.refresh_max || a <= s.delta2 * s.rr_ref || (s.stall > 0 && s.since_min >= s.stall))
        s.trig |= TRIG_SCHEDULE;
      if (s.iters >= s.max_iter) s.trig |= TRIG_LIMIT;
      if (s.trig) s.refresh = 1;
    }
  } else {  // PH_F: after the reliable update
    if (!std::isfinite(c)) s.restart = 1;  // (r, r0) unusable: start a new Krylov space
    if (!std::isfinite(a)) {
      s.status = BCM_SOLVE_FAILED;
    } else {
      s.rr_true = a;
      s.rr = a;
      s.rr_ref = a;
      s.rr_min = a;
      s.since_min = 0;
      s.since = 0;
      s.nref += 1;
      // Stagnation: the true residual has reached the accuracy the
      // arithmetic allows (about eps * cond(A) of ||b||) when it has not
      // halved over stag_max consecutive checks.
      if (a < 0.25 * s.rr_best) {
        s.rr_best = a;
        s.n_stag = 0;
      } else {
        s.n_stag += 1;
      }
      if (conv_test(a, s)) {
        s.status = BCM_SOLVE_CONVERGED;
      } else if (s.iters >= s.max_iter || (s.stag_max > 0 && s.n_stag >= s.stag_max)) {
        s.status = BCM_SOLVE_NOCONV;
      } else if (s.restart || (s.trig & TRIG_CONV) || (s.restart_every > 0 && s.nref % s.restart_every == 0)) {
        // Breakdown, or the recurrence met the tolerance while the true
        // residual did not: restart from the current solution (r0 = r,
        // p = r on the next pass), as the host-driven solver restarts.
        s.first = 1;
        s.reset_r0 = 1;
        s.rho = a;
      } else {
        // Reliable update: keep the Krylov space, replace the residual.
        s.rho = c;
      }
    }
    s.restart = 0;
    s.trig = 0;
    s.refresh = 0;
  }
  *sp = s;
}

}  // namespace

// ---------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------
// Position of an ongoing solve (between the phases).
struct Run {
  double* x = nullptr;
  int chunk = 8, pre = 2;
  double omega = 0.8;
  bool use_graph = true;
  long long max_blocks = 0, blocks = 0;
};

struct Krylov {
  Run run;
  hipEvent_t ev = nullptr;  // the state mirror is valid once this completes
  int esize = 0;            // bytes per value of the sweep-precision vectors (4 or 8)
  bool fused = false;       // vectors allocated for the fused two-sweep path
  void* v[6] = {};          // compact, sweep precision: r, r0, p, q, s, t
  double* dx = nullptr;     // compact FP64 correction since the last reliable update
  // Unfused path: compact sweep scratch and its column export.
  void* g1 = nullptr;
  void* x1 = nullptr;
  // Fused two-sweep path (N <= 16, count_pre == 2): p's second buffer and
  // the column exports of r, q and both p buffers.  All sweep precision.
  void* p2 = nullptr;
  void* rxe = nullptr;
  void* qxe = nullptr;
  void* pxe[2] = {};
  double* g[2] = {};        // compact FP64 preconditioned vectors p_, s_ (SpMV inputs)
  double* xg[2] = {};       // their column exports
  double* b64 = nullptr;    // folded right-hand side, compact
  double* parts = nullptr;  // 2 per block
  SegState* st = nullptr;
  SegState* h_st = nullptr;  // pinned mirror
  // Blocks the iteration kernels visit: those of the segments still active
  // (work-group i takes block list[i]); shrunk between graph replays.
  int* list = nullptr;
  int n_list = 0;
  std::vector<int> h_list;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // What the captured graph was built for.
  int key_esize = 0, key_pre = -1, key_iters = 0, key_nlist = -1;
  double key_omega = 0.0;
  const void* key_x = nullptr;
};

static void krylov_free_vectors(Krylov* k) {
  for (auto& p : k->v) { dev_free(p); p = nullptr; }
  dev_free(k->g1);
  dev_free(k->x1);
  k->g1 = k->x1 = nullptr;
  for (void** p : {&k->p2, &k->rxe, &k->qxe, &k->pxe[0], &k->pxe[1]}) {
    dev_free(*p);
    *p = nullptr;
  }
  k->esize = 0;
  k->fused = false;
}

static void krylov_drop_graph(Krylov* k) {
  if (k->exec) (void)hipGraphExecDestroy(k->exec);
  if (k->graph) (void)hipGraphDestroy(k->graph);
  k->exec = nullptr;
  k->graph = nullptr;
  k->key_esize = 0;
}

void krylov_release(bcm_mesh* m) {
  Krylov* k = m->kry;
  if (!k) return;
  krylov_drop_graph(k);
  krylov_free_vectors(k);
  dev_free(k->dx);
  for (auto& p : k->g) dev_free(p);
  for (auto& p : k->xg) dev_free(p);
  dev_free(k->b64);
  dev_free(k->parts);
  dev_free(k->list);
  dev_free(k->st);
  if (k->h_st) (void)hipHostFree(k->h_st);
  if (k->ev) (void)hipEventDestroy(k->ev);
  delete k;
  m->kry = nullptr;
}

template <typename T>
static int zero_async(bcm_mesh* m, T* p, size_t count) {
  hipLaunchKernelGGL(k_kry_zero<T>, dim3(unsigned((count + 255) / 256)), dim3(256), 0, m->stream, p,
                     (long long)count);
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}

static int krylov_get(bcm_mesh* m, int esize, bool fused, Krylov** out) {
  const size_t nc = m->ncell, nx = size_t(m->nblocks) * 4 * m->n * m->n;
  if (!m->kry) {
    Krylov* k = new (std::nothrow) Krylov();
    if (!k) return BCM_ERR_NOMEM;
    m->kry = k;
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->b64), sizeof(double) * nc));
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->dx), sizeof(double) * nc));
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->parts), sizeof(double) * 2 * size_t(m->nblocks)));
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->list), sizeof(int) * size_t(m->nblocks)));
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->st), sizeof(SegState) * size_t(m->nseg)));
    BCM_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&k->h_st), sizeof(SegState) * size_t(m->nseg)));
    for (int i = 0; i < 2; ++i) {
      BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->g[i]), sizeof(double) * nc));
      BCM_TRY(dev_alloc(reinterpret_cast<void**>(&k->xg[i]), sizeof(double) * nx));
      BCM_TRY(zero_async(m, k->g[i], nc));
      BCM_TRY(zero_async(m, k->xg[i], nx));
    }
  }
  Krylov* k = m->kry;
  if (k->esize != esize || k->fused != fused) {
    BCM_HIP_TRY(hipStreamSynchronize(m->stream));
    krylov_drop_graph(k);
    krylov_free_vectors(k);
    for (auto& p : k->v) {
      BCM_TRY(dev_alloc(&p, nc * esize));
      if (esize == 8) BCM_TRY(zero_async(m, static_cast<double*>(p), nc));
      else BCM_TRY(zero_async(m, static_cast<float*>(p), nc));
    }
    std::vector<void**> full, cols;
    if (fused) {
      full = {&k->p2};
      cols = {&k->rxe, &k->qxe, &k->pxe[0], &k->pxe[1]};
    } else {
      full = {&k->g1};
      cols = {&k->x1};
    }
    for (void** p : full) {
      BCM_TRY(dev_alloc(p, nc * esize));
      if (esize == 8) BCM_TRY(zero_async(m, static_cast<double*>(*p), nc));
      else BCM_TRY(zero_async(m, static_cast<float*>(*p), nc));
    }
    for (void** p : cols) {
      BCM_TRY(dev_alloc(p, nx * esize));
      if (esize == 8) BCM_TRY(zero_async(m, static_cast<double*>(*p), nx));
      else BCM_TRY(zero_async(m, static_cast<float*>(*p), nx));
    }
    k->esize = esize;
    k->fused = fused;
  }
  *out = k;
  return BCM_OK;
}

// Block-wise gather.  gate 0: every block of the mesh; 1 / 2: the blocks of
// the active list whose segment is active (2: with an update pending).
template <typename T, int N>
static int gather(bcm_mesh* m, T* a, int gate) {
  const dim3 blk(KWG);
  const SegState* st = m->kry ? m->kry->st : nullptr;
  if (gate == 0) {
    hipLaunchKernelGGL((k_gather<T, N, 0>), dim3(m->nblocks), blk, 0, m->stream, a, m->d_desc, m->d_blk_seg, st,
                       (const int*)nullptr, 0);
  } else {
    const dim3 grid(m->kry->n_list);
    const int* bl = m->kry->list;
    if (gate == 1)
      hipLaunchKernelGGL((k_gather<T, N, 1>), grid, blk, 0, m->stream, a, m->d_desc, m->d_blk_seg, st, bl, 0);
    else
      hipLaunchKernelGGL((k_gather<T, N, 2>), grid, blk, 0, m->stream, a, m->d_desc, m->d_blk_seg, st, bl, 0);
  }
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}

// Rebuild the active list from the segment states in the host mirror
// (blocks of active segments, in block order).
static int set_list(bcm_mesh* m, bool all) {
  Krylov* K = m->kry;
  K->h_list.clear();
  for (int s = 0; s < m->nseg; ++s) {
    if (!all && K->h_st[s].status != BCM_SOLVE_ACTIVE) continue;
    for (int b = m->seg_first[size_t(s)]; b < m->seg_first[size_t(s) + 1]; ++b) K->h_list.push_back(b);
  }
  K->n_list = int(K->h_list.size());
  if (K->n_list > 0)
    BCM_HIP_TRY(hipMemcpyAsync(K->list, K->h_list.data(), sizeof(int) * K->h_list.size(), hipMemcpyHostToDevice,
                               m->stream));
  // The copy reads the host vector: finish it before the vector can change.
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  return BCM_OK;
}

template <typename T>
int ghost_gather_raw(bcm_mesh* m, T* data) {
  if (m->vc != 1) return BCM_ERR_ARG;
  switch (m->n) {
    case 4: return gather<T, 4>(m, data, 0);
    case 8: return gather<T, 8>(m, data, 0);
    case 16: return gather<T, 16>(m, data, 0);
    case 32: return gather<T, 32>(m, data, 0);
    default: return BCM_ERR_ARG;
  }
}
template int ghost_gather_raw<double>(bcm_mesh*, double*);
template int ghost_gather_raw<float>(bcm_mesh*, float*);

template <int PH>
static int seg_reduce(bcm_mesh* m) {
  hipLaunchKernelGGL(k_seg<PH>, dim3(m->nseg), dim3(KWG), 0, m->stream, m->kry->st, m->kry->parts, m->d_seg_first);
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}

// Jacobi preconditioner chain: the first sweep is fused into the update
// kernel; sweeps 2..pre alternate between the sweep scratch (precision T)
// and the FP64 result so that the last one lands in `result`.
template <typename T, int N>
static int precondition(bcm_mesh* m, T* scratch, T* sx, double* result, double* rx, const T* rhs, int pre,
                        double omega) {
  bool in_result = (pre % 2 == 1);
  const dim3 grid(m->kry->n_list), blk(KWG);
  const int* bs = m->d_blk_seg;
  const int* bl = m->kry->list;
  const SegState* st = m->kry->st;
  for (int sweep = 2; sweep <= pre; ++sweep) {
    if (in_result) {
      if constexpr (N <= 16)
        hipLaunchKernelGGL((k_kry_jacobi_l<T, N, double, T, kwg_jac<N>()>), grid, dim3(kwg_jac<N>()), 0, m->stream, scratch, sx, result, rx, rhs,
                           m->d_meta, bs, st, omega, bl, m->d_desc);
      else
        hipLaunchKernelGGL((k_kry_jacobi<T, N, double, T>), grid, blk, 0, m->stream, scratch, sx, result, rx, rhs,
                           m->d_meta, bs, st, omega, bl, m->d_desc);
    } else {
      if constexpr (N <= 16)
        hipLaunchKernelGGL((k_kry_jacobi_l<T, N, T, double, kwg_jac<N>()>), grid, dim3(kwg_jac<N>()), 0, m->stream, result, rx, scratch, sx, rhs,
                           m->d_meta, bs, st, omega, bl, m->d_desc);
      else
        hipLaunchKernelGGL((k_kry_jacobi<T, N, T, double>), grid, blk, 0, m->stream, result, rx, scratch, sx, rhs,
                           m->d_meta, bs, st, omega, bl, m->d_desc);
    }
    BCM_HIP_TRY(hipGetLastError());
    in_result = !in_result;
  }
  return BCM_OK;
}

// Update kernel (p or s) with the fused first sweep into the right target.
template <typename T, int N, bool IS_P>
static int update_and_first_sweep(bcm_mesh* m, T* v, const T* r, const T* q, T* r0, T* scratch, T* sx,
                                  double* result, double* rx, int pre, double omega) {
  const dim3 grid(m->kry->n_list), blk(KWG);
  const int* bs = m->d_blk_seg;
  const int* bl = m->kry->list;
  const SegState* st = m->kry->st;
  const bool to_result = pre == 0 || pre % 2 == 1;
  if (IS_P) {
    if (pre == 0)
      hipLaunchKernelGGL((k_kry_p<T, N, 0, double>), grid, blk, 0, m->stream, v, r, q, r0, result, rx, m->d_meta,
                         bs, st, omega, bl);
    else if (to_result)
      hipLaunchKernelGGL((k_kry_p<T, N, 1, double>), grid, blk, 0, m->stream, v, r, q, r0, result, rx, m->d_meta,
                         bs, st, omega, bl);
    else
      hipLaunchKernelGGL((k_kry_p<T, N, 1, T>), grid, blk, 0, m->stream, v, r, q, r0, scratch, sx, m->d_meta, bs,
                         st, omega, bl);
  } else {
    if (pre == 0)
      hipLaunchKernelGGL((k_kry_s<T, N, 0, double>), grid, blk, 0, m->stream, v, r, q, result, rx, m->d_meta, bs, st,
                         omega, bl);
    else if (to_result)
      hipLaunchKernelGGL((k_kry_s<T, N, 1, double>), grid, blk, 0, m->stream, v, r, q, result, rx, m->d_meta, bs, st,
                         omega, bl);
    else
      hipLaunchKernelGGL((k_kry_s<T, N, 1, T>), grid, blk, 0, m->stream, v, r, q, scratch, sx, m->d_meta, bs, st,
                         omega, bl);
  }
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}

// Whether the iteration uses the fused two-sweep update kernels.
template <int N>
static constexpr bool fused_pre2_ok() {
  return N <= 16;
}
static bool fused_pre2(int n, int pre) {
  return n <= 16 && pre == 2;
}

// One iteration of the fused path: p and s updates carry both
// preconditioner sweeps; q and r carry their column exports.
template <typename T, int N>
static int iteration_fused(bcm_mesh* m, double* x, double omega) {
  Krylov* K = m->kry;
  T *r = static_cast<T*>(K->v[0]), *r0 = static_cast<T*>(K->v[1]), *p = static_cast<T*>(K->v[2]),
    *q = static_cast<T*>(K->v[3]), *s = static_cast<T*>(K->v[4]), *t = static_cast<T*>(K->v[5]);
  T *p2 = static_cast<T*>(K->p2), *rx = static_cast<T*>(K->rxe), *qx = static_cast<T*>(K->qxe);
  T *px0 = static_cast<T*>(K->pxe[0]), *px1 = static_cast<T*>(K->pxe[1]);
  double *gp = K->g[0], *gs = K->g[1], *xp = K->xg[0], *xs = K->xg[1];
  const dim3 grid(K->n_list), blk(KWG);
  const int* bs = m->d_blk_seg;
  const int* bl = K->list;
  SegState* st = K->st;
  constexpr int WJ = kwg_jac<N>(), WS = kwg_spmv<N>();
  hipLaunchKernelGGL((k_kry_upre2<T, N, WJ, true>), grid, dim3(WJ), 0, m->stream, p, p2, px0, px1, r, rx, q, qx, r0,
                     gp, xp, m->d_meta, bs, st, omega, bl, m->d_desc);
  BCM_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((k_kry_spmv_l<T, N, 1, WS, true>), grid, dim3(WS), 0, m->stream, q, gp, xp, r0, m->d_meta, bs,
                     st, K->parts, bl, m->d_desc, qx);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_Q>(m));
  hipLaunchKernelGGL((k_kry_upre2<T, N, WJ, false>), grid, dim3(WJ), 0, m->stream, s, static_cast<T*>(nullptr),
                     static_cast<T*>(nullptr), static_cast<T*>(nullptr), r, rx, q, qx, static_cast<T*>(nullptr), gs,
                     xs, m->d_meta, bs, st, omega, bl, m->d_desc);
  BCM_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((k_kry_spmv_l<T, N, 2, WS>), grid, dim3(WS), 0, m->stream, t, gs, xs, s, m->d_meta, bs, st,
                     K->parts, bl, m->d_desc, static_cast<T*>(nullptr));
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_T>(m));
  hipLaunchKernelGGL((k_kry_xr<T, N, true>), grid, blk, 0, m->stream, K->dx, gp, gs, r, s, t, r0, bs, st, K->parts,
                     bl, rx);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_R>(m));
  hipLaunchKernelGGL((k_ref_x<N>), grid, blk, 0, m->stream, x, K->dx, bs, st, bl);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY((gather<double, N>(m, x, 2)));
  hipLaunchKernelGGL((k_ref_r<T, N, true>), grid, blk, 0, m->stream, r, r0, x, K->b64, m->d_meta, bs, st, K->parts,
                     bl, rx);
  BCM_HIP_TRY(hipGetLastError());
  return seg_reduce<PH_F>(m);
}

// One BiCGSTAB iteration plus the (gated) reliable update.
template <typename T, int N>
static int iteration(bcm_mesh* m, double* x, int pre, double omega) {
  if constexpr (fused_pre2_ok<N>())
    if (pre == 2) return iteration_fused<T, N>(m, x, omega);
  Krylov* K = m->kry;
  T *r = static_cast<T*>(K->v[0]), *r0 = static_cast<T*>(K->v[1]), *p = static_cast<T*>(K->v[2]),
    *q = static_cast<T*>(K->v[3]), *s = static_cast<T*>(K->v[4]), *t = static_cast<T*>(K->v[5]);
  T *g1 = static_cast<T*>(K->g1), *x1 = static_cast<T*>(K->x1);
  double *gp = K->g[0], *gs = K->g[1], *xp = K->xg[0], *xs = K->xg[1];
  const dim3 grid(K->n_list), blk(KWG);
  const int* bs = m->d_blk_seg;
  const int* bl = K->list;
  SegState* st = K->st;
  BCM_TRY((update_and_first_sweep<T, N, true>(m, p, r, q, r0, g1, x1, gp, xp, pre, omega)));
  BCM_TRY((precondition<T, N>(m, g1, x1, gp, xp, p, pre, omega)));
  if constexpr (N <= 16)
    hipLaunchKernelGGL((k_kry_spmv_l<T, N, 1, kwg_spmv<N>()>), grid, dim3(kwg_spmv<N>()), 0, m->stream, q, gp, xp, r0, m->d_meta, bs, st, K->parts,
                       bl, m->d_desc);
  else
    hipLaunchKernelGGL((k_kry_spmv<T, N, 1>), grid, blk, 0, m->stream, q, gp, xp, r0, m->d_meta, bs, st, K->parts,
                       bl, m->d_desc);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_Q>(m));
  BCM_TRY((update_and_first_sweep<T, N, false>(m, s, r, q, nullptr, g1, x1, gs, xs, pre, omega)));
  BCM_TRY((precondition<T, N>(m, g1, x1, gs, xs, s, pre, omega)));
  if constexpr (N <= 16)
    hipLaunchKernelGGL((k_kry_spmv_l<T, N, 2, kwg_spmv<N>()>), grid, dim3(kwg_spmv<N>()), 0, m->stream, t, gs, xs, s, m->d_meta, bs, st, K->parts,
                       bl, m->d_desc);
  else
    hipLaunchKernelGGL((k_kry_spmv<T, N, 2>), grid, blk, 0, m->stream, t, gs, xs, s, m->d_meta, bs, st, K->parts,
                       bl, m->d_desc);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_T>(m));
  hipLaunchKernelGGL((k_kry_xr<T, N>), grid, blk, 0, m->stream, K->dx, gp, gs, r, s, t, r0, bs, st, K->parts, bl);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_R>(m));
  // Reliable update where requested.
  hipLaunchKernelGGL((k_ref_x<N>), grid, blk, 0, m->stream, x, K->dx, bs, st, bl);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY((gather<double, N>(m, x, 2)));
  hipLaunchKernelGGL((k_ref_r<T, N>), grid, blk, 0, m->stream, r, r0, x, K->b64, m->d_meta, bs, st, K->parts, bl);
  BCM_HIP_TRY(hipGetLastError());
  return seg_reduce<PH_F>(m);
}

static bool any_active(const bcm_mesh* m) {
  for (int s = 0; s < m->nseg; ++s)
    if (m->kry->h_st[s].status == BCM_SOLVE_ACTIVE) return true;
  return false;
}

// Copy the segment states to the pinned mirror and mark the point in the
// stream after which the mirror is valid.
static int post_state(bcm_mesh* m) {
  Krylov* K = m->kry;
  BCM_HIP_TRY(hipMemcpyAsync(K->h_st, K->st, sizeof(SegState) * size_t(m->nseg), hipMemcpyDeviceToHost, m->stream));
  if (!K->ev) BCM_HIP_TRY(hipEventCreateWithFlags(&K->ev, hipEventDisableTiming));
  BCM_HIP_TRY(hipEventRecord(K->ev, m->stream));
  return BCM_OK;
}

// Phase 1: segment states, right-hand side, initial residual and (with
// use_graph) the captured block of `chunk` iterations; all asynchronous.
template <typename T, int N>
static int begin_impl(bcm_mesh* m, double* x, const double* rhs, const bcm_solve_params& P, const double* seg_tol) {
  static_assert(Blk<N>::N3 % KWG == 0, "block volume must be a multiple of the work-group");
  Krylov* K = nullptr;
  BCM_TRY(krylov_get(m, int(sizeof(T)), fused_pre2(N, P.count_pre), &K));
  const dim3 grid(m->nblocks), blk(KWG);
  // Defaults measured on the shell meshes of levels 4-7 (16^3 blocks): an
  // FP64 update each time the FP32 recurrence has gained three digits or
  // stalled for 40 iterations, restarting the Krylov space there.
  const double delta = P.precision == BCM_PREC_MIXED ? (P.refresh_delta > 0.0 ? P.refresh_delta : 1e-3) : 0.0;
  const int rmax = P.refresh_max > 0 ? P.refresh_max : (P.precision == BCM_PREC_MIXED ? 1000 : (1 << 30));
  const int rs_every = P.precision == BCM_PREC_FP64 ? 0 : (P.restart_every == 0 ? 1 : std::max(P.restart_every, 0));
  const int stall = P.precision == BCM_PREC_FP64 ? 0 : (P.stall_window > 0 ? P.stall_window : (P.stall_window == 0 ? 40 : 0));
  // The pinned mirror may still be the target of an earlier copy.
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  for (int s = 0; s < m->nseg; ++s) {
    SegState z;
    std::memset(&z, 0, sizeof(z));
    const double tol = seg_tol ? seg_tol[s] : P.tol;
    z.tol2 = tol * tol;
    z.delta2 = delta * delta;
    z.status = BCM_SOLVE_ACTIVE;
    z.first = 1;
    z.refresh = 1;
    z.restart = 1;  // the first pass sets r0 = r
    z.max_iter = P.max_iter;
    z.refresh_max = rmax;
    z.restart_every = rs_every;
    z.stall = stall;
    // FP64: 8 checks (each follows a recurrence that met the tolerance).
    // MIXED/FP32 check after every inner solve, and an inner solve can set x
    // back before later ones recover (measured on the level-7 shell mesh),
    // so they get 30.
    const int stag_default = P.precision == BCM_PREC_FP64 ? 8 : 30;
    z.stag_max = P.stagnation > 0 ? P.stagnation : (P.stagnation == 0 ? stag_default : 0);
    z.rr_best = DBL_MAX;
    z.rho_prev = 1.0;
    z.gamma = 1.0;
    K->h_st[s] = z;
  }
  BCM_HIP_TRY(hipMemcpyAsync(K->st, K->h_st, sizeof(SegState) * size_t(m->nseg), hipMemcpyHostToDevice, m->stream));
  // b, (b, b), zero solutions of zero right-hand sides.
  hipLaunchKernelGGL(k_kry_rhs<N>, grid, blk, 0, m->stream, K->b64, rhs, m->d_meta, m->d_bvalues, K->parts);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_BB>(m));
  hipLaunchKernelGGL(k_zero_x, grid, blk, 0, m->stream, x, (long long)m->bvol, m->d_blk_seg, K->st);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(zero_async(m, K->dx, m->ncell));
  BCM_TRY(set_list(m, true));
  // Initial residual r = b - A x; the first pass sets r0 = r.
  BCM_TRY((gather<double, N>(m, x, 2)));
  if (fused_pre2(N, P.count_pre))
    hipLaunchKernelGGL((k_ref_r<T, N, fused_pre2_ok<N>()>), dim3(K->n_list), blk, 0, m->stream,
                       static_cast<T*>(K->v[0]), static_cast<const T*>(K->v[1]), x, K->b64, m->d_meta, m->d_blk_seg,
                       K->st, K->parts, K->list, static_cast<T*>(K->rxe));
  else
    hipLaunchKernelGGL((k_ref_r<T, N>), dim3(K->n_list), blk, 0, m->stream, static_cast<T*>(K->v[0]),
                       static_cast<const T*>(K->v[1]), x, K->b64, m->d_meta, m->d_blk_seg, K->st, K->parts, K->list);
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(seg_reduce<PH_F>(m));
  Run& R = K->run;
  R.x = x;
  R.chunk = P.check_every > 0 ? P.check_every : 8;
  R.pre = P.count_pre;
  R.omega = P.omega;
  R.use_graph = P.use_graph != 0;
  R.max_blocks = (long long)P.max_iter / R.chunk + 4;
  R.blocks = 0;
  return post_state(m);
}

// The captured block of `chunk` iterations for the current list.
template <typename T, int N>
static int ensure_graph(bcm_mesh* m) {
  Krylov* K = m->kry;
  Run& R = K->run;
  if (K->exec && K->key_esize == int(sizeof(T)) && K->key_pre == R.pre && K->key_iters == R.chunk &&
      K->key_omega == R.omega && K->key_x == R.x && K->key_nlist == K->n_list)
    return BCM_OK;
  krylov_drop_graph(K);
  BCM_HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = BCM_OK;
  for (int it = 0; it < R.chunk && rc == BCM_OK; ++it) rc = iteration<T, N>(m, R.x, R.pre, R.omega);
  hipGraph_t gph = nullptr;
  const hipError_t ec = hipStreamEndCapture(m->stream, &gph);
  if (rc != BCM_OK) {
    if (gph) (void)hipGraphDestroy(gph);
    return rc;
  }
  BCM_HIP_TRY(ec);
  K->graph = gph;
  BCM_HIP_TRY(hipGraphInstantiate(&K->exec, K->graph, nullptr, nullptr, 0));
  K->key_esize = int(sizeof(T));
  K->key_pre = R.pre;
  K->key_iters = R.chunk;
  K->key_omega = R.omega;
  K->key_x = R.x;
  K->key_nlist = K->n_list;
  return BCM_OK;
}

// Phase 2: one block of `chunk` iterations.
template <typename T, int N>
static int chunk_impl(bcm_mesh* m) {
  Krylov* K = m->kry;
  Run& R = K->run;
  // Drop the blocks of finished segments once they are 30 % of the list:
  // their work-groups would only return at once, but at 17 launches an
  // iteration that costs milliseconds on a large batch.
  long long active_blocks = 0;
  for (int s = 0; s < m->nseg; ++s)
    if (K->h_st[s].status == BCM_SOLVE_ACTIVE) active_blocks += m->seg_first[size_t(s) + 1] - m->seg_first[size_t(s)];
  if (active_blocks > 0 && double(active_blocks) < 0.7 * K->n_list) BCM_TRY(set_list(m, false));
  if (R.use_graph) {
    BCM_TRY((ensure_graph<T, N>(m)));
    BCM_HIP_TRY(hipGraphLaunch(K->exec, m->stream));
  } else {
    for (int it = 0; it < R.chunk; ++it) BCM_TRY((iteration<T, N>(m, R.x, R.pre, R.omega)));
  }
  R.blocks += 1;
  return post_state(m);
}

// Phase 3: ghosts of the solution (as ImposeBoundaryCondition leaves them)
// and the per-segment outcome.
template <int N>
static int finish_impl(bcm_mesh* m, bcm_solve_stats* stats) {
  Krylov* K = m->kry;
  BCM_TRY((gather<double, N>(m, K->run.x, 0)));
  BCM_TRY(post_state(m));
  BCM_HIP_TRY(hipEventSynchronize(K->ev));
  for (int s = 0; s < m->nseg; ++s) {
    const SegState& z = K->h_st[s];
    bcm_solve_stats& o = stats[s];
    o.status = z.status == BCM_SOLVE_ACTIVE ? BCM_SOLVE_NOCONV : z.status;
    o.iters = z.iters;
    o.refreshes = z.nref > 0 ? z.nref - 1 : 0;
    o.pad = 0;
    o.relres = z.bb > double(FLT_MIN) ? std::sqrt(z.rr_true / z.bb) : std::sqrt(z.rr_true);
  }
  return BCM_OK;
}

int krylov_begin(bcm_mesh* m, double* x, const double* rhs, const bcm_solve_params& P, const double* seg_tol) {
  const bool f64 = P.precision == BCM_PREC_FP64;
  switch (m->n) {
    case 8: return f64 ? begin_impl<double, 8>(m, x, rhs, P, seg_tol) : begin_impl<float, 8>(m, x, rhs, P, seg_tol);
    case 16: return f64 ? begin_impl<double, 16>(m, x, rhs, P, seg_tol) : begin_impl<float, 16>(m, x, rhs, P, seg_tol);
    case 32: return f64 ? begin_impl<double, 32>(m, x, rhs, P, seg_tol) : begin_impl<float, 32>(m, x, rhs, P, seg_tol);
    default: return BCM_ERR_ARG;
  }
}

int krylov_advance(bcm_mesh* m, bool wait, bool* active) {
  Krylov* K = m->kry;
  if (!K || !K->ev) return BCM_ERR_STATE;
  *active = true;
  if (wait) {
    BCM_HIP_TRY(hipEventSynchronize(K->ev));
  } else {
    const hipError_t q = hipEventQuery(K->ev);
    if (q == hipErrorNotReady) return BCM_OK;
    BCM_HIP_TRY(q);
  }
  if (!any_active(m) || K->run.blocks >= K->run.max_blocks) {
    *active = false;
    return BCM_OK;
  }
  const bool f64 = K->esize == 8;
  switch (m->n) {
    case 8: return f64 ? chunk_impl<double, 8>(m) : chunk_impl<float, 8>(m);
    case 16: return f64 ? chunk_impl<double, 16>(m) : chunk_impl<float, 16>(m);
    case 32: return f64 ? chunk_impl<double, 32>(m) : chunk_impl<float, 32>(m);
    default: return BCM_ERR_ARG;
  }
}

int krylov_state(bcm_mesh* m, int* state) {
  Krylov* K = m->kry;
  if (!K || !K->ev) return BCM_ERR_STATE;
  const hipError_t q = hipEventQuery(K->ev);
  if (q == hipErrorNotReady) {
    *state = 0;
    return BCM_OK;
  }
  BCM_HIP_TRY(q);
  *state = (!any_active(m) || K->run.blocks >= K->run.max_blocks) ? 2 : 1;
  return BCM_OK;
}

int krylov_finish(bcm_mesh* m, bcm_solve_stats* stats) {
  switch (m->n) {
    case 8: return finish_impl<8>(m, stats);
    case 16: return finish_impl<16>(m, stats);
    case 32: return finish_impl<32>(m, stats);
    default: return BCM_ERR_ARG;
  }
}

// The current iterate of a running solve, one segment: y = x + dx on the
// interior of its blocks (the sum the next reliable update forms), x's
// stored ghost cells elsewhere; then the segment's connected faces' ghosts
// are gathered, as the end of a solve leaves them.  y holds the segment's
// blocks only.
template <int N>
__global__ __launch_bounds__(KWG) void k_cur_x(double* __restrict__ y, const double* __restrict__ x,
                                               const double* __restrict__ dx, int first) {
  using B = Blk<N>;
  const long long b = first + (long long)blockIdx.x;
  const double* xb = x + b * B::GV;
  const double* db = dx + b * B::N3;
  double* yb = y + (long long)blockIdx.x * B::GV;
  for (int e = threadIdx.x; e < B::GV; e += KWG) {
    const int i = e % B::NX - 1, j = (e / B::NX) % B::NX - 1, k = e / (B::NX * B::NX) - 1;
    const bool in = i >= 0 && i < N && j >= 0 && j < N && k >= 0 && k < N;
    yb[e] = in ? xb[e] + db[i + N * (j + N * k)] : xb[e];
  }
}

template <int N>
static int snapshot_impl(bcm_mesh* m, double* y, int first, int nb) {
  Krylov* K = m->kry;
  hipLaunchKernelGGL(k_cur_x<N>, dim3(unsigned(nb)), dim3(KWG), 0, m->stream, y, K->run.x, K->dx, first);
  BCM_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((k_gather<double, N, 0>), dim3(unsigned(nb)), dim3(KWG), 0, m->stream, y, m->d_desc,
                     m->d_blk_seg, K->st, (const int*)nullptr, first);
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}

int krylov_snapshot(bcm_mesh* m, int seg, double* y, int* iters) {
  Krylov* K = m->kry;
  if (!K || !K->ev || !K->dx || !K->run.x || seg < 0 || seg >= m->nseg) return BCM_ERR_STATE;
  const int first = m->seg_first[size_t(seg)], nb = m->seg_first[size_t(seg) + 1] - first;
  // Iterations already queued finish first, so x and dx belong together.
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  switch (m->n) {
    case 8: BCM_TRY(snapshot_impl<8>(m, y, first, nb)); break;
    case 16: BCM_TRY(snapshot_impl<16>(m, y, first, nb)); break;
    case 32: BCM_TRY(snapshot_impl<32>(m, y, first, nb)); break;
    default: return BCM_ERR_ARG;
  }
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  if (iters) *iters = K->h_st[seg].iters;
  return BCM_OK;
}

int krylov_check(const bcm_mesh* m, const bcm_solve_params* p, const double* seg_tol) {
  if (m->vc != 1 || (m->n != 8 && m->n != 16 && m->n != 32)) return BCM_ERR_ARG;
  if (p->precision < BCM_PREC_FP64 || p->precision > BCM_PREC_FP32 || !(p->tol > 0.0) || p->max_iter < 1 ||
      p->count_pre < 0 || p->count_pre > 16 || p->check_every < 0 || p->refresh_max < 0 || p->refresh_delta < 0.0)
    return BCM_ERR_ARG;
  if (seg_tol)
    for (int s = 0; s < m->nseg; ++s)
      if (!(seg_tol[s] > 0.0)) return BCM_ERR_ARG;
  return BCM_OK;
}

size_t krylov_bytes(int nblocks, int n, int precision, int count_pre) {
  const size_t nc = size_t(nblocks) * n * n * n, nx = size_t(nblocks) * 4 * n * n;
  const size_t es = precision == BCM_PREC_FP64 ? 8 : 4;
  // b, dx, p_, s_ (FP64 compact) + the exports of p_, s_; r, r0, p, q, s, t
  // (compact); then either p's second buffer and the exports of r, q and
  // both p buffers (fused two-sweep path) or the sweep scratch and its
  // export; partials, list.
  const size_t path = fused_pre2(n, count_pre) ? nc * es + 4 * nx * es : nc * es + nx * es;
  return 4 * nc * 8 + 2 * nx * 8 + 6 * nc * es + path + size_t(nblocks) * 20;
}

}  // namespace bcm

using namespace bcm;

extern "C" int bcm_mesh_set_progress(bcm_mesh_t m, double seconds) {
  if (!m || !(seconds >= 0.0)) return BCM_ERR_ARG;
  m->progress_s = seconds;
  return BCM_OK;
}

extern "C" int bcm_solve(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
                         const bcm_solve_params* p, const double* seg_tol, bcm_solve_stats* stats) {
  if (!x || !rhs || !p || !stats || x->mesh != rhs->mesh || x->dtype != BCM_F64 || rhs->dtype != BCM_F64)
    return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  BCM_TRY(krylov_check(m, p, seg_tol));
  BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  BCM_TRY(krylov_begin(m, static_cast<double*>(x->data), static_cast<const double*>(rhs->data), *p, seg_tol));
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  auto t_last = t_start;
  bool active = true;
  while (active) {
    BCM_TRY(krylov_advance(m, true, &active));
    if (m->progress_s > 0.0) {
      const auto now = clock::now();
      if (std::chrono::duration<double>(now - t_last).count() >= m->progress_s) {
        t_last = now;
        int done = 0, most = 0;
        for (int s = 0; s < m->nseg; ++s) {
          done += m->kry->h_st[s].status != BCM_SOLVE_ACTIVE;
          most = std::max(most, int(m->kry->h_st[s].iters));
        }
        std::fprintf(stderr, "[bcm_solve] %.1f s: %d/%d segments done, %d iterations (most)\n",
                     std::chrono::duration<double>(now - t_start).count(), done, m->nseg, most);
        std::fflush(stderr);
      }
    }
  }
  return krylov_finish(m, stats);
}
