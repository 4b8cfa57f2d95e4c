// Halo-gathering red-black SOR sweep.
//
// The reference's outer iteration is: SOR on every block with frozen
// ghosts, then the ghost exchange (copy / F2C / C2F) and the Dirichlet
// mirror (examples/sor2/Solver.cpp:77-133).  The ghosts seen by sweep k+1
// are therefore a pure function of the interiors after sweep k
// (include/Scalar3DUpdater.h:190-266, 321-347; examples/sor1/Solver.cpp:232-294).
// This kernel runs sweep k+1 reading the iterate of sweep k from buffer A
// and writing buffer B, and builds each block's ghost frame on the fly from
// A with the same arithmetic as the exchange, so no separate ghost pass and
// no ghost writes are needed between sweeps.  A, B and the source term are
// compact (n^3 values per block, x fastest, no ghost cells), so every row
// is read and written whole and aligned.  The X-face sources (columns 0,
// 1, n-2, n-1 of every block) come from a compact export written by the
// previous sweep, which keeps every read coalesced.
#include <type_traits>

#include "bcm_internal.h"
#include "bcm_interp.h"
#include "bcm_poisson.h"

namespace bcm {

namespace {

template <typename T, int N>
struct HaloSrc {
  const T* __restrict__ A;   // previous iterate, compact
  const T* __restrict__ xe;  // its X-face export [blk][4][k][j]
  const int* desc;   // face descriptors of this work-group's block (LDS copy)
  const double* __restrict__ bval;
  const double* bface;  // Dirichlet value per face of this block (LDS copy)
  // Ghosted field whose ghost cells stand for the faces without a
  // neighbour (FD_NONE), read as stored; null where no such face occurs.
  const T* G = nullptr;
  static constexpr int N3 = N * N * N;
  static constexpr int NXG = N + 2;
  __device__ __forceinline__ const T* g_ptr(int blk, int i, int j, int k) const {
    return G + (long long)blk * NXG * NXG * NXG + (i + 1) + NXG * ((j + 1) + NXG * (k + 1));
  }
  __device__ __forceinline__ T a_at(int blk, int i, int j, int k) const {
    return A[(long long)blk * N3 + i + N * (j + N * k)];
  }
  // i must be 0, 1, N-2 or N-1.
  __device__ __forceinline__ T x_at(int blk, int i, int j, int k) const {
    const int col = i < 2 ? i : i - (N - 4);
    return xe[((long long)(blk * 4 + col) * N + k) * N + j];
  }
  __device__ __forceinline__ T at(int a, int blk, int i, int j, int k) const {
    return a == 0 ? x_at(blk, i, j, k) : a_at(blk, i, j, k);
  }
  __device__ __forceinline__ const T* a_ptr(int blk, int i, int j, int k) const {
    return A + (long long)blk * N3 + i + N * (j + N * k);
  }
  __device__ __forceinline__ const T* x_ptr(int blk, int i, int j, int k) const {
    const int col = i < 2 ? i : i - (N - 4);
    return xe + ((long long)(blk * 4 + col) * N + k) * N + j;
  }
  __device__ __forceinline__ const T* ptr(int a, int blk, int i, int j, int k) const {
    return a == 0 ? x_ptr(blk, i, j, k) : a_ptr(blk, i, j, k);
  }
  // Phase 1 of a ghost value of block b at frame cell c (one coordinate -1
  // or N) on face F: the source addresses of the face kind (a single
  // source repeated for copy / Dirichlet faces), then eight loads in
  // straight-line code so the outstanding-load count stays exact.
  template <int F>
  __device__ __forceinline__ int gather(int b, int ci, int cj, int ck, T* v) const {
    constexpr int a = F >> 1;
    constexpr bool plus = F & 1;
    constexpr int t0 = (a + 1) % 3, t1 = (a + 2) % 3;
    const int* d = desc + F * 8;
    const int kind = d[0];
    int c[3] = {ci, cj, ck};
    const T* base;
    int o[8];
    if (kind == FD_F2C) {
      constexpr int h = N / 2;
      const int s0 = c[t0] >= h, s1 = c[t1] >= h;
      const int fb = d[1 + s0 + 2 * s1];
      int I[3];
      I[t0] = c[t0] - h * s0;
      I[t1] = c[t1] - h * s1;
      I[a] = plus ? 0 : h - 1;
      base = ptr(a, fb, 2 * I[0], 2 * I[1], 2 * I[2]);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        o[q] = int(ptr(a, fb, 2 * I[0] + (q & 1), 2 * I[1] + ((q >> 1) & 1), 2 * I[2] + (q >> 2)) - base);
    } else if (kind == FD_C2F) {
      const int cb = d[1], sf = d[5];
      int p[3], I[3];
      double w;
      p[t0] = c[t0] + N * (sf & 1);
      p[t1] = c[t1] + N * (sf >> 1);
      p[a] = plus ? 0 : 2 * N - 1;
#pragma unroll
      for (int e = 0; e < 3; ++e) c2f_weight(p[e], N, I[e], w);
      base = ptr(a, cb, I[0], I[1], I[2]);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        o[q] = int(ptr(a, cb, I[0] + (q & 1), I[1] + ((q >> 1) & 1), I[2] + (q >> 2)) - base);
    } else {
      if (kind == FD_COPY) {
        c[a] = plus ? 0 : N - 1;
        base = ptr(a, d[1], c[0], c[1], c[2]);
      } else if (kind == FD_DIR) {
        c[a] = plus ? N - 1 : 0;
        base = ptr(a, b, c[0], c[1], c[2]);
      } else if (G) {
        // No neighbour: the stored ghost.
        base = g_ptr(b, ci, cj, ck);
      } else {
        // No neighbour and no condition: the SOR schedule is only used on
        // meshes without such faces (the stored ghost is not in the compact
        // iterate); an in-block cell keeps the address valid.
        base = a_ptr(b, min(max(ci, 0), N - 1), min(max(cj, 0), N - 1), min(max(ck, 0), N - 1));
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) o[q] = 0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = base[o[q]];
    return kind;
  }
  // Phase 2: the exchange / mirror arithmetic on the gathered values.
  template <int F>
  __device__ __forceinline__ T combine(int b, int ci, int cj, int ck, int kind, const T* v) const {
    constexpr int a = F >> 1;
    constexpr bool plus = F & 1;
    constexpr int t0 = (a + 1) % 3, t1 = (a + 2) % 3;
    if (kind == FD_DIR) return T(2.0 * bface[F] - double(v[0]));
    if (kind == FD_F2C) return f2c_combine<T>(v);
    if (kind == FD_C2F) {
      const int sf = desc[F * 8 + 5];
      const int c[3] = {ci, cj, ck};
      int p[3], I[3];
      double w[3], vv[8];
      p[t0] = c[t0] + N * (sf & 1);
      p[t1] = c[t1] + N * (sf >> 1);
      p[a] = plus ? 0 : 2 * N - 1;
#pragma unroll
      for (int e = 0; e < 3; ++e) c2f_weight(p[e], N, I[e], w[e]);
#pragma unroll
      for (int q = 0; q < 8; ++q) vv[q] = double(v[q]);
      return T(c2f_combine(vv, w[0], w[1], w[2]));
    }
    return v[0];
  }
  // X/Y frame edges.  A frame cell's sources move along z in a fixed
  // pattern, so each thread resolves its face once (block, base cell,
  // strides) and every plane then costs an index select and one address
  // add per value.
  //
  // z-index of plane p inside the source block: the plane itself for copy /
  // Dirichlet / stored ghosts, 2(p - N/2 [p >= N/2]) in the fine block of an
  // F2C face, the clamped coarse index of fine row p + N*sfz for C2F.
  static __device__ __forceinline__ int z_index(int kind, int zoff, int p) {
    constexpr int h = N / 2;
    const int zf = 2 * (p - h * (p >= h));
    const int zc0 = min(max((p + 1) / 2 - 1, 0), N - 2);
    const int zc1 = min(max((p + N + 1) / 2 - 1, 0), N - 2);
    return kind == FD_F2C ? zf : (kind == FD_C2F ? (zoff ? zc1 : zc0) : p);
  }
  template <int VPT>
  struct FrameLoad {
    const T* base[2];  // source of z-index 0, planes p < N/2 and p >= N/2
    int off[VPT];      // offsets of this thread's values from the base cell
    int zs;            // stride of one z-index
    int kind;
    int zoff;          // C2F: N * (subface bit along z)
  };
  template <int VPT>
  __device__ __forceinline__ FrameLoad<VPT> frame_setup(int face, int b, int ci, int cj, int q0) const {
    constexpr int h = N / 2;
    FrameLoad<VPT> fl;
    const int a = face >> 1;  // 0: X face (tangents y, z), 1: Y face (tangents z, x)
    const bool plus = face & 1;
    const int* d = desc + face * 8;
    const int kind = d[0];
    const int cin = a == 0 ? cj : ci;  // in-plane tangential coordinate
    const int sf = d[5];
    const int sin_ = cin >= h;
    const int sfin = a == 0 ? (sf & 1) : (sf >> 1);
    const int sfz = a == 0 ? (sf >> 1) : (sf & 1);
    int nrm, inp, blk0, blk1;
    double w;
    if (kind == FD_F2C) {
      nrm = plus ? 0 : N - 2;
      inp = 2 * (cin - h * sin_);
      // subface (s0, s1): X faces s0 along y, s1 along z; Y faces s0 along z, s1 along x
      blk0 = a == 0 ? d[1 + sin_] : d[1 + 2 * sin_];
      blk1 = a == 0 ? d[3 + sin_] : d[2 + 2 * sin_];
    } else if (kind == FD_C2F) {
      c2f_weight(plus ? 0 : 2 * N - 1, N, nrm, w);
      c2f_weight(cin + N * sfin, N, inp, w);
      blk0 = blk1 = d[1];
    } else if (kind == FD_COPY) {
      nrm = plus ? 0 : N - 1;
      inp = cin;
      blk0 = blk1 = d[1];
    } else if (kind == FD_DIR) {
      nrm = plus ? N - 1 : 0;
      inp = cin;
      blk0 = blk1 = b;
    } else {
      nrm = a == 0 ? min(max(ci, 0), N - 1) : min(max(cj, 0), N - 1);  // not reached (see gather)
      inp = cin;
      blk0 = blk1 = b;
    }
    const int x = a == 0 ? nrm : inp, y = a == 0 ? inp : nrm;
    const bool viax = a == 0 && kind != FD_NONE;  // X faces read the column export
    fl.base[0] = viax ? x_ptr(blk0, x, y, 0) : a_ptr(blk0, x, y, 0);
    fl.base[1] = viax ? x_ptr(blk1, x, y, 0) : a_ptr(blk1, x, y, 0);
    fl.zs = viax ? N : N * N;
    if (kind == FD_NONE && G) {
      // No neighbour: the stored ghost cells of this block, plane by plane.
      const int gx = a == 0 ? (plus ? N : -1) : inp, gy = a == 0 ? inp : (plus ? N : -1);
      fl.base[0] = fl.base[1] = g_ptr(b, gx, gy, 0);
      fl.zs = NXG * NXG;
    }
    const int xs = viax ? N * N : 1, ys = viax ? 1 : N;
    const bool many = kind == FD_F2C || kind == FD_C2F;
#pragma unroll
    for (int e = 0; e < VPT; ++e) {
      const int q = q0 + e;
      fl.off[e] = many ? (q & 1) * xs + ((q >> 1) & 1) * ys + (q >> 2) * fl.zs : 0;
    }
    fl.kind = kind;
    fl.zoff = kind == FD_C2F ? N * sfz : 0;
    return fl;
  }
  template <int VPT>
  static __device__ __forceinline__ void frame_load(const FrameLoad<VPT>& fl, int p, T* v) {
    const T* bp = fl.base[p >= N / 2];
    const int zo = z_index(fl.kind, fl.zoff, p) * fl.zs;
#pragma unroll
    for (int e = 0; e < VPT; ++e) v[e] = bp[fl.off[e] + zo];
  }
  // Combiner of one frame cell: kind, the C2F weights along x and y, the
  // Dirichlet term 2 * value.
  struct FrameComb {
    int kind, zoff;
    double wx, wy, dir2;
  };
  __device__ __forceinline__ FrameComb comb_setup(int face, int ci, int cj) const {
    FrameComb fc;
    const int a = face >> 1;
    const bool plus = face & 1;
    const int* d = desc + face * 8;
    const int sf = d[5];
    const int cin = a == 0 ? cj : ci;
    const int sfin = a == 0 ? (sf & 1) : (sf >> 1);
    const int sfz = a == 0 ? (sf >> 1) : (sf & 1);
    int I;
    double wn, win;
    c2f_weight(plus ? 0 : 2 * N - 1, N, I, wn);
    c2f_weight(cin + N * sfin, N, I, win);
    fc.kind = d[0];
    fc.zoff = N * sfz;
    fc.wx = a == 0 ? wn : win;
    fc.wy = a == 0 ? win : wn;
    fc.dir2 = 2.0 * bface[face];
    return fc;
  }
  static __device__ __forceinline__ T frame_combine(const FrameComb& fc, int p, const T* v) {
    if (fc.kind == FD_F2C) return f2c_combine<T>(v);
    if (fc.kind == FD_C2F) {
      int I;
      double w0, w1, vv[8];
      c2f_weight(p, N, I, w0);
      c2f_weight(p + N, N, I, w1);
#pragma unroll
      for (int q = 0; q < 8; ++q) vv[q] = double(v[q]);
      return T(c2f_combine(vv, fc.wx, fc.wy, fc.zoff ? w1 : w0));
    }
    if (fc.kind == FD_DIR) return T(fc.dir2 - double(v[0]));
    return v[0];
  }
};

// One aligned two-cell store (16 B for FP64, 8 B for FP32), marked
// non-temporal: the next sweep reads the iterate back long after this
// block's data has left the caches, so the store streams past them.
__device__ __forceinline__ void store_pair(double* p, double a, double b) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = a;
  v.y = b;
  __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}
__device__ __forceinline__ void store_pair(float* p, float a, float b) {
  typedef float v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = a;
  v.y = b;
  __builtin_nontemporal_store(v, reinterpret_cast<v2*>(p));
}

template <typename T, bool SRC>
__device__ __forceinline__ T sor_cell(T fc, T fw, T fe, T fs, T fn, T fb, T ft, T sv, T c0, T cx, T cy, T cz, T om,
                                      T om1) {
  if (SRC) {
    const T fnew = c0 * (((cx * (fw + fe) + cy * (fs + fn)) + cz * (fb + ft)) - sv);
    return om * fnew + om1 * fc;
  } else {
    return ((c0 * fc + cx * (fw + fe)) + cy * (fs + fn)) + cz * (fb + ft);
  }
}

}  // namespace

// One block per work-group, N^2/2 threads; thread t owns the cell pair
// (2pi..2pi+1, pj) of every z-plane (one red, one black).  The thread's own
// cells of planes k-2..k+3 sit in registers (vertical neighbours never go
// through LDS); LDS holds three ghosted planes for the in-plane neighbours
// and two staging planes for the X/Y ghost frame; one barrier per z-step:
//   step k: commit plane k (own old values); park the frame sources of plane
//   k+1; issue the loads of plane k+3 (own cells and frame sources) and of
//   the source term of plane k+2; barrier; combine the frame of plane k+1;
//   red cells of plane k (in-plane from LDS, vertical from registers; red
//   results also to LDS); black cells of plane k-1 (in-plane red values from
//   LDS, black results too), then plane k-1's final pair straight from
//   registers to B (one aligned 16-byte store) and, where a neighbour reads
//   it, to the X-face export.
// Every LDS cell is written by one thread; slot k%3 is rewritten at step
// k+3, after the barriers of steps k+1 and k+2; staging slot p&1 is
// rewritten two steps after it was read.
template <typename T, int N, bool SRC, bool GOUT>
__global__ __launch_bounds__(N* N / 2) __attribute__((amdgpu_waves_per_eu(4))) void k_rbsor_halo(
    const T* __restrict__ A, T* __restrict__ B, const T* __restrict__ s, const double* __restrict__ coef,
    double omega, const int* __restrict__ desc, const T* __restrict__ xin, T* __restrict__ xout,
    const double* __restrict__ bval, T* __restrict__ gout) {
  constexpr int NX = N + 2;
  constexpr int PL = NX * NX;
  constexpr int HALF = N / 2;
  constexpr int NN = N * N;
  __shared__ __align__(16) T lds[3][PL];
  __shared__ T stage[2][8][4 * N];
  __shared__ T xstage[3][4][N];  // X-export values of a plane, [slot][column][j]
  __shared__ int bdesc[48];
  __shared__ double bface[6];
  const int b = blockIdx.x;
  if (threadIdx.x < 48) bdesc[threadIdx.x] = desc[b * 48 + threadIdx.x];
  if (threadIdx.x < 6) {
    const int id = desc[b * 48 + threadIdx.x * 8 + 6];
    bface[threadIdx.x] = desc[b * 48 + threadIdx.x * 8] == FD_DIR ? bval[id] : 0.0;
  }
  __syncthreads();
  const HaloSrc<T, N> src{A, xin, bdesc, bval, bface};
  const long long base = (long long)b * NN * N;
  const T* __restrict__ Ab = A + base;
  T* __restrict__ Bb = B + base;
  const T* __restrict__ sb = SRC ? s + base : nullptr;
  const T c0 = T(coef[4 * b + 0]), cx = T(coef[4 * b + 1]), cy = T(coef[4 * b + 2]), cz = T(coef[4 * b + 3]);
  const T om = T(omega), om1 = T(1.0 - omega);
  const int t = threadIdx.x;
  const int i0 = 2 * (t % HALF), j0 = t / HALF;
  const int own = j0 * N + i0;  // offset of the thread's pair in a compact plane (even: 16-byte aligned)
  const bool framer = t < 4 * N;
  const int edge = t / N, u = t % N;
  const int fci = edge == 0 ? -1 : (edge == 1 ? N : u);
  const int fcj = edge < 2 ? u : (edge == 2 ? -1 : N);
  auto L = [](int i, int j) __attribute__((always_inline)) { return (j + 1) * NX + (i + 1); };
  auto LQ = [&](int q) __attribute__((always_inline)) { return (j0 + 1) * NX + (i0 + q + 1); };

  // Own cells of plane p: interior rows of A, or the gathered z ghost plane.
  auto load_own = [&](int p, T* o) __attribute__((always_inline)) {
    if (p >= 0 && p < N) {
      const T* row = Ab + (long long)p * NN + own;
#pragma unroll
      for (int q = 0; q < 2; ++q) o[q] = row[q];
    } else {
#pragma unroll 1
      for (int q = 0; q < 2; ++q) {
        T v[8], w;
        const int i = i0 + q, j = j0;
        if (p < 0) {
          const int kd = src.template gather<BCM_Z_M>(b, i, j, p, v);
          w = src.template combine<BCM_Z_M>(b, i, j, p, kd, v);
        } else {
          const int kd = src.template gather<BCM_Z_P>(b, i, j, p, v);
          w = src.template combine<BCM_Z_P>(b, i, j, p, kd, v);
        }
        o[0] = q == 0 ? w : o[0];
        o[1] = q == 1 ? w : o[1];
      }
    }
  };
  // The source term is read once per sweep and by no other work-group:
  // non-temporal loads.  The iterate's rows stay cacheable (the neighbours
  // read this block's faces for their ghost frames).
  auto load_s = [&](int p, T* sv) __attribute__((always_inline)) {
    if (!SRC) return;
    const T* row = sb + (long long)p * NN + own;
#pragma unroll
    for (int q = 0; q < 2; ++q) sv[q] = __builtin_nontemporal_load(row + q);
  };
  // X/Y ghost frame: 4N cells per plane, eight source values each, spread
  // over all threads: thread t loads values VPT*gq .. VPT*gq+VPT-1 of frame
  // cell gc (lanes of a wave take consecutive cells of one edge).  A plane's
  // loads are issued three steps before its red cells need it, parked in
  // LDS one step before, and combined by threads t < 4N.
  constexpr int VPT = 64 / N;
  const int gq = t / (4 * N), gc = t % (4 * N);
  const int gedge = gc / N, gu = gc % N;
  const int gci = gedge == 0 ? -1 : (gedge == 1 ? N : gu);
  const int gcj = gedge < 2 ? gu : (gedge == 2 ? -1 : N);
  const auto fload = src.template frame_setup<VPT>(gedge, b, gci, gcj, VPT * gq);
  const auto fcomb = src.comb_setup(framer ? edge : 0, fci, fcj);
  auto gather_frame = [&](int p, T* v) __attribute__((always_inline)) { src.frame_load(fload, p, v); };
  auto park_frame = [&](int p, const T* v) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < VPT; ++e) stage[p & 1][VPT * gq + e][gc] = v[e];
  };
  auto put_frame = [&](int p) __attribute__((always_inline)) {
    if (framer) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = stage[p & 1][q][t];
      lds[p % 3][L(fci, fcj)] = src.frame_combine(fcomb, p, v);
    }
  };
  auto commit = [&](int p, const T* o) __attribute__((always_inline)) {
    const int S = p % 3;
#pragma unroll
    for (int q = 0; q < 2; ++q) lds[S][LQ(q)] = o[q];
  };
  // Red cells of plane p: in-plane black values from LDS, vertical from the
  // own cells of planes p-1 (dn) and p+1 (up).
  auto red = [&](auto par, int p, T* o, const T* dn, const T* up, const T* sv) __attribute__((always_inline)) {
    constexpr int PAR = decltype(par)::value;
    const int S = p % 3;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (((q + j0 + PAR) & 1) != 0) continue;
      const int c = LQ(q);
      o[q] = sor_cell<T, SRC>(o[q], lds[S][c - 1], lds[S][c + 1], lds[S][c - NX], lds[S][c + NX], dn[q], up[q],
                              SRC ? sv[q] : T(0), c0, cx, cy, cz, om, om1);
      lds[S][c] = o[q];
    }
  };
  auto black = [&](auto par, int p, T* o, const T* dn, const T* up, const T* sv) __attribute__((always_inline)) {
    constexpr int PAR = decltype(par)::value;
    const int S = p % 3;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (((q + j0 + PAR) & 1) != 1) continue;
      const int c = LQ(q);
      o[q] = sor_cell<T, SRC>(o[q], lds[S][c - 1], lds[S][c + 1], lds[S][c - NX], lds[S][c + NX], dn[q], up[q],
                              SRC ? sv[q] : T(0), c0, cx, cy, cz, om, om1);
      lds[S][c] = o[q];
    }
  };
  // Export columns the next sweep reads: the outer column (0 or N-1) on a
  // side with a neighbour or a Dirichlet condition (copy, mirror), the inner
  // one (1 or N-2) only across a level change (F2C and C2F read two
  // columns).  The threads whose pair is cells 0, 1 or N-2, N-1 write them.
  const bool xlo = i0 == 0, xhi = i0 == N - 2;
  const int xkind = bdesc[(xlo ? BCM_X_M : BCM_X_P) * 8];
  const bool lvl = xkind == FD_F2C || xkind == FD_C2F;
  const bool xo0 = (xlo && (xkind != FD_NONE)) || (xhi && lvl);   // cell i0 (column 0 or N-2)
  const bool xo1 = (xlo && lvl) || (xhi && (xkind != FD_NONE));   // cell i0+1 (column 1 or N-1)
  const int xcol = xlo ? 0 : 2;
  // Exported columns 0, 1, N-2, N-1 (bits 0..3), as xo0 / xo1 above.
  const int kxm = bdesc[BCM_X_M * 8], kxp = bdesc[BCM_X_P * 8];
  const int xcols = (kxm != FD_NONE ? 1 : 0) | (kxm == FD_F2C || kxm == FD_C2F ? 2 : 0) |
                    (kxp == FD_F2C || kxp == FD_C2F ? 4 : 0) | (kxp != FD_NONE ? 8 : 0);
  // The export of plane p is staged in LDS by the threads owning the cells
  // (step p+1) and written at step p+2 by the last wave, two values per
  // lane, so each column row of N values leaves as whole cache lines.
  constexpr int XT0 = N * N / 2 - 64;
  auto export_plane = [&](int p) __attribute__((always_inline)) {
    const int l = t - XT0, col = l / (N / 2), j2 = 2 * (l % (N / 2));
    if (l >= 0 && l < 2 * N && ((xcols >> col) & 1))
      store_pair(xout + ((long long)(b * 4 + col) * N + p) * N + j2, xstage[p % 3][col][j2],
                 xstage[p % 3][col][j2 + 1]);
  };
  // Plane p's final pair, once black(p) has run: B and the export, or
  // (GOUT: the last sweep of a call) the ghosted field's interior.  GOUT is
  // a template parameter so that no store sits under a run-time branch:
  // the compiler then counts the outstanding memory operations exactly
  // (a store under a branch counts as possibly not issued, which makes the
  // waits for earlier loads drain more than they need).
  auto store_own = [&](int p, const T* o) __attribute__((always_inline)) {
    if constexpr (GOUT) {
      // The row's two X ghost cells are written as well (with the row's end
      // values), so the rows j = 0..N-1 of a plane form one contiguous span
      // of whole cache lines instead of rows with 8-byte holes at both ends
      // (partial lines cost a read-modify-write).  Those cells are not part
      // of the result: this sweep only runs on meshes whose every face is
      // connected or Dirichlet, and the call ends with the ghost fill and
      // the Dirichlet mirror, which overwrite them.
      T* g = gout + (long long)b * PL * NX + (long long)(p + 1) * PL + LQ(0);
      if (xlo) g[-1] = o[0];
      g[0] = o[0];
      g[1] = o[1];
      if (xhi) g[2] = o[1];
      return;
    }
    store_pair(Bb + (long long)p * NN + own, o[0], o[1]);
    if (xo0) xstage[p % 3][xcol][j0] = o[0];
    if (xo1) xstage[p % 3][xcol + 1][j0] = o[1];
  };

  // Register window: own cells of planes k-2 .. k+3 (loads of plane k+3
  // issued at step k), source of planes k-1 .. k+2 (k+2 issued at step k),
  // frame sources of planes k+1 .. k+3 (fr1 .. fr3).  The loop is unrolled
  // so each cell's colour and each plane's frame index are compile-time.
  T um2[2] = {T(0), T(0)}, um1[2], u0[2], up1[2], up2[2], up3[2];
  T sp[2] = {T(0), T(0)}, sc[2] = {T(0), T(0)}, sn[2] = {T(0), T(0)}, snn[2] = {T(0), T(0)};
  T zp[2];
  load_own(N, zp);  // z ghost planes (own positions) up front
  load_own(-1, um1);
  load_own(0, u0);
  load_own(1, up1);
  load_own(2, up2);
  load_s(0, sc);
  load_s(1, sn);
  T fr1[VPT], fr2[VPT], fr3[VPT];
  gather_frame(0, fr1);
  park_frame(0, fr1);
  gather_frame(1, fr1);
  gather_frame(2, fr2);
  __syncthreads();
  put_frame(0);
  auto step = [&](auto par, int k) __attribute__((always_inline)) {
    constexpr int PAR = decltype(par)::value;  // k % 2
    commit(k, u0);
    if (k + 1 < N) park_frame(k + 1, fr1);
    if (k + 3 < N) {
      const T* row = Ab + (long long)(k + 3) * NN + own;
#pragma unroll
      for (int q = 0; q < 2; ++q) up3[q] = row[q];
    } else if (k + 3 == N) {
      up3[0] = zp[0];
      up3[1] = zp[1];
    }
    if (k + 2 < N) load_s(k + 2, snn);
    if (k + 3 < N) gather_frame(k + 3, fr3);
    __syncthreads();
    if (k + 1 < N) put_frame(k + 1);
    if (!GOUT && k >= 2) export_plane(k - 2);
    red(std::integral_constant<int, PAR>(), k, u0, um1, up1, sc);
    if (k >= 1) {
      black(std::integral_constant<int, 1 - PAR>(), k - 1, um1, um2, u0, sp);
      store_own(k - 1, um1);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      um2[q] = um1[q];
      um1[q] = u0[q];
      u0[q] = up1[q];
      up1[q] = up2[q];
      up2[q] = up3[q];
      sp[q] = sc[q];
      sc[q] = sn[q];
      sn[q] = snn[q];
    }
#pragma unroll
    for (int e = 0; e < VPT; ++e) {
      fr1[e] = fr2[e];
      fr2[e] = fr3[e];
    }
  };
  // Fully unrolled: the window rotation becomes register renaming, so no
  // copy waits for a load that is still in flight.
#pragma unroll
  for (int k = 0; k < N; k += 2) {
    step(std::integral_constant<int, 0>(), k);
    step(std::integral_constant<int, 1>(), k + 1);
  }
  // um1 = plane N-1, um2 = plane N-2, u0 = plane N (ghost); sp = s(N-1).
  __syncthreads();
  if (!GOUT) export_plane(N - 2);
  black(std::integral_constant<int, (N - 1) & 1>(), N - 1, um1, um2, u0, sp);
  store_own(N - 1, um1);
  if (!GOUT) {
    __syncthreads();
    export_plane(N - 1);
  }
}

template <typename T, int N>
static void launch_halo(bcm_mesh* m, const T* A, T* B, const T* s, double omega, const T* xin, T* xout, T* gout) {
  const dim3 grid(m->nblocks), blk(N * N / 2);
  if (s && gout)
    hipLaunchKernelGGL((k_rbsor_halo<T, N, true, true>), grid, blk, 0, m->stream, A, B, s, m->d_sor_coef, omega, m->d_desc, xin, xout, m->d_bvalues, gout);
  else if (s)
    hipLaunchKernelGGL((k_rbsor_halo<T, N, true, false>), grid, blk, 0, m->stream, A, B, s, m->d_sor_coef, omega, m->d_desc, xin, xout, m->d_bvalues, gout);
  else if (gout)
    hipLaunchKernelGGL((k_rbsor_halo<T, N, false, true>), grid, blk, 0, m->stream, A, B, s, m->d_sor_coef, omega, m->d_desc, xin, xout, m->d_bvalues, gout);
  else
    hipLaunchKernelGGL((k_rbsor_halo<T, N, false, false>), grid, blk, 0, m->stream, A, B, s, m->d_sor_coef, omega, m->d_desc, xin, xout, m->d_bvalues, gout);
}

template <typename T>
int halo_sweep_raw(bcm_mesh* m, const T* A, T* B, const T* s, double omega, const T* xin, T* xout, T* gout) {
  EventPair ev;
  BCM_TRY(timer_begin(m, gout ? BCM_K_SOR : BCM_K_SOR_HALO, &ev));
  if (m->n == 32) {
    launch_halo<T, 32>(m, A, B, s, omega, xin, xout, gout);
  } else if (m->n == 16) {
    launch_halo<T, 16>(m, A, B, s, omega, xin, xout, gout);
  } else {
    return BCM_ERR_ARG;
  }
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}
template int halo_sweep_raw<double>(bcm_mesh*, const double*, double*, const double*, double, const double*, double*,
                                    double*);
template int halo_sweep_raw<float>(bcm_mesh*, const float*, float*, const float*, double, const float*, float*, float*);

// ---------------------------------------------------------------------
// Halo-gathering Jacobi iteration of the matrix-free Poisson operator.
//
// The reference's step is jacobi_smoother on every block (blas.f90:452-502,
// with the Poisson coefficients of bcut.f90:2848-2866 folded at outer
// faces as bca.f90 folds them) followed by the ghost exchange
// (include/Scalar3DUpdater.h:190-266, 321-347).  As for SOR above, the
// ghosts of step k+1 are a function of the interiors after step k, so
// step k+1 builds each block's ghost frame from the compact iterate of
// step k (copy / F2C / C2F with the exchange's arithmetic) and no ghost
// pass runs between steps.  Faces without a neighbour keep the ghost
// values stored in the field (the exchange never writes them; a folded
// face multiplies them by a zero coefficient), read from the field as
// they are.
// ---------------------------------------------------------------------

// Step 1 reads the field with its stored ghosts, as the reference's first
// step does, and writes the compact iterate, its four-column X export and
// the compact right-hand side.  One work-group per block.
template <typename T, int N>
__global__ __launch_bounds__(256) void k_jacobi_first(const T* __restrict__ x, const T* __restrict__ rhs,
                                                      const BlockMeta* __restrict__ meta, double omega,
                                                      T* __restrict__ out, T* __restrict__ xout,
                                                      T* __restrict__ rout) {
  constexpr int NX = N + 2, GV = NX * NX * NX, N3 = N * N * N;
  const int b = blockIdx.x;
  const BlockMeta me = meta[b];
  const T* __restrict__ xb = x + (long long)b * GV;
  const T* __restrict__ bb = rhs + (long long)b * GV;
  const T om = T(omega);
  for (int e = threadIdx.x; e < N3; e += 256) {
    const int i = e % N, j = (e / N) % N, k = e / (N * N);
    const int o = (i + 1) + NX * ((j + 1) + NX * (k + 1));
    const Coef7<T> a = poisson_coef<T>(me, i, j, k, N);
    const T bv = bb[o];
    const T v = jacobi_cell(a, xb, o, NX, NX * NX, bv, om);
    out[(long long)b * N3 + e] = v;
    rout[(long long)b * N3 + e] = bv;
    if (i < 2 || i >= N - 2) {
      const int col = i < 2 ? i : i - (N - 4);
      xout[((long long)(b * 4 + col) * N + k) * N + j] = v;
    }
  }
}

// Steps 2..k: one block per work-group, N^2/2 threads, thread t owning the
// cell pair (2pi..2pi+1, pj) of every z-plane, as the SOR sweep above.
// Own cells of planes k-1..k+3 sit in registers (vertical neighbours),
// LDS holds three ghosted planes (in-plane neighbours and the X/Y frame)
// and two frame staging planes; one barrier per z-step:
//   step k: commit plane k; park the frame sources of plane k+1; issue the
//   loads of plane k+3 (own cells, frame sources) and of the right-hand
//   side of plane k+2; barrier; combine the frame of plane k+1; Jacobi on
//   plane k's pair; store it (compact iterate and X export, or -- GOUT, the
//   last step -- the field's rows).
// Slot k%3 is rewritten at step k+2 (frame) and k+3 (own cells), after the
// barriers of steps k+1 and k+2.  For 16^3 blocks (four frame values per
// thread and plane) the kernel needs 132 VGPRs: three waves per SIMD
// instead of four, without spills (0.550 against 0.556 ms per step on
// 26,888 blocks).
template <typename T, int N, bool GOUT>
__global__ __launch_bounds__(N* N / 2) __attribute__((amdgpu_waves_per_eu(N == 16 ? 3 : 4))) void k_jacobi_halo(
    const T* __restrict__ A, T* __restrict__ B, const T* __restrict__ rhs, const BlockMeta* __restrict__ meta,
    double omega, const int* __restrict__ desc, const T* __restrict__ xin, T* __restrict__ xout, T* x) {
  constexpr int NX = N + 2;
  constexpr int PL = NX * NX;
  constexpr int HALF = N / 2;
  constexpr int NN = N * N;
  __shared__ __align__(16) T lds[3][PL];
  __shared__ T stage[2][8][4 * N];
  __shared__ T xstage[3][4][N];  // X-export values of a plane, [slot][column][j]
  __shared__ int bdesc[48];
  __shared__ double bnone[6];  // no Dirichlet mirror here: zero values
  const int b = blockIdx.x;
  // Faces without a neighbour read the stored ghosts: a Dirichlet face is
  // one of them here (its condition is folded into the coefficients).
  if (threadIdx.x < 48) {
    const int v = desc[b * 48 + threadIdx.x];
    bdesc[threadIdx.x] = (threadIdx.x % 8 == 0 && v == FD_DIR) ? int(FD_NONE) : v;
  }
  if (threadIdx.x < 6) bnone[threadIdx.x] = 0.0;
  __syncthreads();
  HaloSrc<T, N> src{A, xin, bdesc, nullptr, bnone};
  src.G = x;
  const long long base = (long long)b * NN * N;
  const T* __restrict__ Ab = A + base;
  T* __restrict__ Bb = B + base;
  const T* __restrict__ rb = rhs + base;
  const BlockMeta me = meta[b];
  const T om = T(omega);
  const int t = threadIdx.x;
  const int i0 = 2 * (t % HALF), j0 = t / HALF;
  const int own = j0 * N + i0;
  const bool framer = t < 4 * N;
  const int edge = t / N, u = t % N;
  const int fci = edge == 0 ? -1 : (edge == 1 ? N : u);
  const int fcj = edge < 2 ? u : (edge == 2 ? -1 : N);
  auto L = [](int i, int j) __attribute__((always_inline)) { return (j + 1) * NX + (i + 1); };
  auto LQ = [&](int q) __attribute__((always_inline)) { return (j0 + 1) * NX + (i0 + q + 1); };

  auto load_own = [&](int p, T* o) __attribute__((always_inline)) {
    if (p >= 0 && p < N) {
      const T* row = Ab + (long long)p * NN + own;
#pragma unroll
      for (int q = 0; q < 2; ++q) o[q] = row[q];
    } else {
#pragma unroll 1
      for (int q = 0; q < 2; ++q) {
        T v[8], w;
        const int i = i0 + q, j = j0;
        if (p < 0) {
          const int kd = src.template gather<BCM_Z_M>(b, i, j, p, v);
          w = src.template combine<BCM_Z_M>(b, i, j, p, kd, v);
        } else {
          const int kd = src.template gather<BCM_Z_P>(b, i, j, p, v);
          w = src.template combine<BCM_Z_P>(b, i, j, p, kd, v);
        }
        o[0] = q == 0 ? w : o[0];
        o[1] = q == 1 ? w : o[1];
      }
    }
  };
  // The right-hand side is read once per step and by no other work-group.
  auto load_r = [&](int p, T* rv) __attribute__((always_inline)) {
    const T* row = rb + (long long)p * NN + own;
#pragma unroll
    for (int q = 0; q < 2; ++q) rv[q] = __builtin_nontemporal_load(row + q);
  };
  constexpr int VPT = 64 / N;
  const int gq = t / (4 * N), gc = t % (4 * N);
  const int gedge = gc / N, gu = gc % N;
  const int gci = gedge == 0 ? -1 : (gedge == 1 ? N : gu);
  const int gcj = gedge < 2 ? gu : (gedge == 2 ? -1 : N);
  const auto fload = src.template frame_setup<VPT>(gedge, b, gci, gcj, VPT * gq);
  const auto fcomb = src.comb_setup(framer ? edge : 0, fci, fcj);
  auto gather_frame = [&](int p, T* v) __attribute__((always_inline)) { src.frame_load(fload, p, v); };
  auto park_frame = [&](int p, const T* v) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < VPT; ++e) stage[p & 1][VPT * gq + e][gc] = v[e];
  };
  auto put_frame = [&](int p) __attribute__((always_inline)) {
    if (framer) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = stage[p & 1][q][t];
      lds[p % 3][L(fci, fcj)] = src.frame_combine(fcomb, p, v);
    }
  };
  auto commit = [&](int p, const T* o) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < 2; ++q) lds[p % 3][LQ(q)] = o[q];
  };
  // Export columns the next step reads (as the SOR sweep's export).
  const bool xlo = i0 == 0, xhi = i0 == N - 2;
  const int xkind = bdesc[(xlo ? BCM_X_M : BCM_X_P) * 8];
  const bool lvl = xkind == FD_F2C || xkind == FD_C2F;
  const bool xo0 = (xlo && (xkind != FD_NONE)) || (xhi && lvl);
  const bool xo1 = (xlo && lvl) || (xhi && (xkind != FD_NONE));
  const int xcol = xlo ? 0 : 2;
  // The export of plane p is staged in LDS by the threads owning the cells
  // (step p) and written at step p+1 by the last wave, two values per
  // lane, as whole cache lines (as the SOR sweep does).
  const int kxm = bdesc[BCM_X_M * 8], kxp = bdesc[BCM_X_P * 8];
  const int xcols = (kxm != FD_NONE ? 1 : 0) | (kxm == FD_F2C || kxm == FD_C2F ? 2 : 0) |
                    (kxp == FD_F2C || kxp == FD_C2F ? 4 : 0) | (kxp != FD_NONE ? 8 : 0);
  constexpr int XT0 = N * N / 2 - 64;
  auto export_plane = [&](int p) __attribute__((always_inline)) {
    const int l = t - XT0, col = l / (N / 2), j2 = 2 * (l % (N / 2));
    if (l >= 0 && l < 2 * N && ((xcols >> col) & 1))
      store_pair(xout + ((long long)(b * 4 + col) * N + p) * N + j2, xstage[p % 3][col][j2],
                 xstage[p % 3][col][j2 + 1]);
  };
  auto store_own = [&](int p, const T* o) __attribute__((always_inline)) {
    if constexpr (GOUT) {
      // Whole rows of the field: the interior pair, and at the row ends the
      // X ghost cells with this step's frame values -- the stored ghost
      // where the face has no neighbour (so it is written back unchanged),
      // the exchanged value elsewhere (rewritten by the closing exchange).
      T* g = x + (long long)b * PL * NX + (long long)(p + 1) * PL + LQ(0);
      if (xlo) g[-1] = lds[p % 3][L(-1, j0)];
      g[0] = o[0];
      g[1] = o[1];
      if (xhi) g[2] = lds[p % 3][L(N, j0)];
      return;
    }
    store_pair(Bb + (long long)p * NN + own, o[0], o[1]);
    if (xo0) xstage[p % 3][xcol][j0] = o[0];
    if (xo1) xstage[p % 3][xcol + 1][j0] = o[1];
  };
  // Jacobi on plane p's pair: in-plane neighbours from LDS, vertical ones
  // from the own cells of planes p-1 (dn) and p+1 (up).
  auto jacobi = [&](int p, const T* cur, const T* dn, const T* up, const T* rv, T* res)
      __attribute__((always_inline)) {
    const int S = p % 3;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = LQ(q);
      const Coef7<T> a = poisson_coef<T>(me, i0 + q, j0, p, N);
      res[q] = jacobi_values(a, lds[S][c - 1], lds[S][c + 1], lds[S][c - NX], lds[S][c + NX], dn[q], up[q], cur[q],
                             rv[q], om);
    }
  };

  T um1[2], u0[2], up1[2], up2[2], up3[2];
  T rc[2], rn[2], rnn[2] = {T(0), T(0)};
  T zp[2];
  load_own(N, zp);
  load_own(-1, um1);
  load_own(0, u0);
  load_own(1, up1);
  load_own(2, up2);
  load_r(0, rc);
  load_r(1, rn);
  T fr1[VPT], fr2[VPT], fr3[VPT];
  gather_frame(0, fr1);
  park_frame(0, fr1);
  gather_frame(1, fr1);
  gather_frame(2, fr2);
  __syncthreads();
  put_frame(0);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    commit(k, u0);
    if (k + 1 < N) park_frame(k + 1, fr1);
    if (k + 3 < N) {
      const T* row = Ab + (long long)(k + 3) * NN + own;
#pragma unroll
      for (int q = 0; q < 2; ++q) up3[q] = row[q];
    } else if (k + 3 == N) {
      up3[0] = zp[0];
      up3[1] = zp[1];
    }
    if (k + 2 < N) load_r(k + 2, rnn);
    if (k + 3 < N) gather_frame(k + 3, fr3);
    __syncthreads();
    if (k + 1 < N) put_frame(k + 1);
    if (!GOUT && k >= 1) export_plane(k - 1);
    T res[2];
    jacobi(k, u0, um1, up1, rc, res);
    store_own(k, res);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      um1[q] = u0[q];
      u0[q] = up1[q];
      up1[q] = up2[q];
      up2[q] = up3[q];
      rc[q] = rn[q];
      rn[q] = rnn[q];
    }
#pragma unroll
    for (int e = 0; e < VPT; ++e) {
      fr1[e] = fr2[e];
      fr2[e] = fr3[e];
    }
  }
  if (!GOUT) {
    __syncthreads();
    export_plane(N - 1);
  }
}

template <typename T>
int jacobi_first_raw(bcm_mesh* m, const T* x, const T* rhs, double omega, T* out, T* xout, T* rout) {
  const dim3 grid(m->nblocks), blk(256);
  if (m->n == 32)
    hipLaunchKernelGGL((k_jacobi_first<T, 32>), grid, blk, 0, m->stream, x, rhs, m->d_meta, omega, out, xout, rout);
  else if (m->n == 16)
    hipLaunchKernelGGL((k_jacobi_first<T, 16>), grid, blk, 0, m->stream, x, rhs, m->d_meta, omega, out, xout, rout);
  else
    return BCM_ERR_ARG;
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}
template int jacobi_first_raw<double>(bcm_mesh*, const double*, const double*, double, double*, double*, double*);

template <typename T, int N>
static void launch_jacobi_halo(bcm_mesh* m, const T* A, T* B, const T* rhs, double omega, const T* xin, T* xout,
                               T* x, bool gout) {
  const dim3 grid(m->nblocks), blk(N * N / 2);
  if (gout)
    hipLaunchKernelGGL((k_jacobi_halo<T, N, true>), grid, blk, 0, m->stream, A, B, rhs, m->d_meta, omega, m->d_desc,
                       xin, xout, x);
  else
    hipLaunchKernelGGL((k_jacobi_halo<T, N, false>), grid, blk, 0, m->stream, A, B, rhs, m->d_meta, omega,
                       m->d_desc, xin, xout, x);
}

template <typename T>
int jacobi_halo_raw(bcm_mesh* m, const T* A, T* B, const T* rhs, double omega, const T* xin, T* xout, T* x,
                    bool gout) {
  if (m->n == 32)
    launch_jacobi_halo<T, 32>(m, A, B, rhs, omega, xin, xout, x, gout);
  else if (m->n == 16)
    launch_jacobi_halo<T, 16>(m, A, B, rhs, omega, xin, xout, x, gout);
  else
    return BCM_ERR_ARG;
  BCM_HIP_TRY(hipGetLastError());
  return BCM_OK;
}
template int jacobi_halo_raw<double>(bcm_mesh*, const double*, double*, const double*, double, const double*, double*,
                                     double*, bool);


}  // namespace bcm
