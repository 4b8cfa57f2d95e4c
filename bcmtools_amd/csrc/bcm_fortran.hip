// The per-block kernels of examples/SolverTest with their Fortran
// by-reference convention (blas.h, sup.h, bca.h, bcx.h, bcax.h, comm.h;
// real = double), for drivers whose block loops pass one block's
// device-resident arrays to them.
//
// Every array is one block's (sz(1)+2g) x (sz(2)+2g) x (sz(3)+2g) values,
// x fastest, which blas.f90 views as x(1-g:sz(1)+g, 1-g:sz(2)+g,
// 1-g:sz(3)+g); Fortran index (i, j, k) is at flat offset
// (i-1+g) + cx (j-1+g) + cx cy (k-1+g) with cx = sz(1)+2g, cy = sz(2)+2g.
//
// Each call checks its sizes, checks that every array it touches lies
// inside one device allocation of the current device, runs on a HIP
// stream of the calling host thread (one per thread and device, created on
// first use, a blocking stream so that it is ordered with the legacy
// default stream), and returns when its results are in place; scalar
// results land in host memory.  Calls from different host threads run
// concurrently (the reference's block loops are OpenMP loops), on blocks
// the caller keeps apart.  A bad argument prints a message and exits the
// process, as the reference's Exit() does.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "bcm_fortran.h"
#include "bcm_internal.h"
#include "bcm_poisson.h"

namespace bcm {
namespace {

constexpr int kFWG = 256;
constexpr int kFMaxDevices = 64;

// One block's array geometry.
struct FGeo {
  int nx, ny, nz, g;
  int cx;            // nx + 2g: the y stride
  long long cxy;     // cx (ny + 2g): the z stride
  long long vol;     // values per array
  long long ncell;   // nx ny nz interior cells
  long long lo, hi;  // flat range of the masked kernels (blas.f90:868-869)
};

__device__ __forceinline__ long long f_at(const FGeo& g, int i, int j, int k) {
  return (long long)(i - 1 + g.g) + (long long)g.cx * (j - 1 + g.g) + g.cxy * (k - 1 + g.g);
}

// Interior cell c (x fastest) -> flat offset.
__device__ __forceinline__ long long f_interior(const FGeo& g, long long c) {
  const int i = int(c % g.nx);
  const long long t = c / g.nx;
  const int j = int(t % g.ny);
  const int k = int(t / g.ny);
  return f_at(g, i + 1, j + 1, k + 1);
}

__device__ __forceinline__ double f_wg_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

// The seven coefficient arrays, or one array of seven interleaved values
// per cell, A(0:6, i, j, k) in the order p, w, e, s, n, b, t
// (blas.f90:504-554).
struct FCoef {
  const double *p, *w, *e, *s, *n, *b, *t;
  const double* packed;
};

__device__ __forceinline__ Coef7<double> f_coef(const FCoef& A, long long o) {
  if (A.packed) {
    const double* q = A.packed + 7 * o;
    return Coef7<double>{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
  }
  return Coef7<double>{A.p[o], A.w[o], A.e[o], A.s[o], A.n[o], A.b[o], A.t[o]};
}

// ---------------------------------------------------------------------
// blas.f90: interior element-wise operations, in the statement's order.
// ---------------------------------------------------------------------
enum FBlas { FB_AXPY, FB_XPAY, FB_AXPYZ, FB_AXPBYPZ, FB_ADD, FB_TRIAD, FB_SCAL };

template <int OP>
__global__ __launch_bounds__(kFWG) void k_fb_blas(double* __restrict__ z, const double* __restrict__ x,
                                                  const double* __restrict__ y, double a, double b, FGeo g) {
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  if (c >= g.ncell) return;
  const long long o = f_interior(g, c);
  if (OP == FB_AXPY) z[o] = z[o] + a * x[o];                    // axpy (blas.f90:264)
  else if (OP == FB_XPAY) z[o] = x[o] + a * z[o];               // xpay (:298)
  else if (OP == FB_AXPYZ) z[o] = a * x[o] + y[o];              // axpyz (:333)
  else if (OP == FB_AXPBYPZ) z[o] = (a * x[o] + b * y[o]) + z[o];  // axpbypz (:367)
  else if (OP == FB_ADD) z[o] = x[o] + y[o];                    // add: C = A + B (:163)
  else if (OP == FB_TRIAD) z[o] = x[o] + a * y[o];              // triad: C = A + d B (:197)
  else z[o] = a * z[o];                                         // scal (:230)
}

// Whole-array fill and copy, ghosts included (blas.f90:18-140).
template <typename T>
__global__ __launch_bounds__(kFWG) void k_fb_fill(T* __restrict__ x, T v, long long count) {
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  if (c < count) x[c] = v;
}

template <typename T>
__global__ __launch_bounds__(kFWG) void k_fb_copy(T* __restrict__ y, const T* __restrict__ x, long long count) {
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  if (c < count) y[c] = x[c];
}

// setup_mask (blas.f90:700-751): 1 on the interior, 0 elsewhere.
__global__ __launch_bounds__(kFWG) void k_fb_setup_mask(double* __restrict__ m, int* __restrict__ mask, FGeo g) {
  const long long o = (long long)blockIdx.x * kFWG + threadIdx.x;
  if (o >= g.vol) return;
  const int i = int(o % g.cx) + 1 - g.g;
  const long long t = o / g.cx;
  const int j = int(t % (g.ny + 2 * g.g)) + 1 - g.g;
  const int k = int(t / (g.ny + 2 * g.g)) + 1 - g.g;
  const bool in = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
  m[o] = in ? 1.0 : 0.0;
  mask[o] = in ? 1 : 0;
}

// ---------------------------------------------------------------------
// Stencil operations on the interior (blas.f90:452-698), the cell
// operators of the whole-mesh kernels.
// ---------------------------------------------------------------------
enum FStencil { FS_JACOBI, FS_AX, FS_R, FS_R2 };

template <int MODE>
__global__ __launch_bounds__(kFWG) void k_fb_stencil(double* __restrict__ out, const double* __restrict__ x,
                                                     FCoef A, const double* __restrict__ b, double omega, FGeo g,
                                                     double* __restrict__ parts) {
  __shared__ double red[4];
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  const int sy = g.cx, sz = int(g.cxy);
  double acc = 0.0;
  if (c < g.ncell) {
    const long long o = f_interior(g, c);
    const Coef7<double> a = f_coef(A, o);
    if (MODE == FS_JACOBI) out[o] = jacobi_cell(a, x, o, sy, sz, b[o], omega);
    else if (MODE == FS_AX) out[o] = ax_cell(a, x, o, sy, sz);
    else {
      const double r = r_cell(a, x, o, sy, sz, b[o]);
      if (MODE == FS_R) out[o] = r;
      else acc = r * r;
    }
  }
  if (MODE == FS_R2) {
    const double s = f_wg_sum(acc, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
  }
}

// dot / dotx2 (blas.f90:379-450): x.y and, for N = 2, x.z; the partials
// of the second sum follow those of the first.
template <int N>
__global__ __launch_bounds__(kFWG) void k_fb_dot(const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, FGeo g, double* __restrict__ parts) {
  __shared__ double red[4];
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  double xy = 0.0, xz = 0.0;
  if (c < g.ncell) {
    const long long o = f_interior(g, c);
    xy = x[o] * y[o];
    if (N == 2) xz = x[o] * z[o];
  }
  const double s = f_wg_sum(xy, red);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
  if (N == 2) {
    __syncthreads();
    const double t = f_wg_sum(xz, red);
    if (threadIdx.x == 0) parts[gridDim.x + blockIdx.x] = t;
  }
}

// The masked forms (blas.f90:858-1030) visit the flat range from the first
// to the last interior cell, ghosts between them included, and act where
// mask == 1.
enum FMasked { FM_JACOBI, FM_AX, FM_COPY, FM_DOT };

template <int MODE>
__global__ __launch_bounds__(kFWG) void k_fb_masked(double* __restrict__ out, const double* __restrict__ x,
                                                    const double* __restrict__ y, FCoef A,
                                                    const double* __restrict__ b, const int* __restrict__ mask,
                                                    double omega, FGeo g, double* __restrict__ parts) {
  __shared__ double red[4];
  const long long o = g.lo + (long long)blockIdx.x * kFWG + threadIdx.x;
  const int sy = g.cx, sz = int(g.cxy);
  double acc = 0.0;
  if (o <= g.hi && mask[o] == 1) {
    if (MODE == FM_JACOBI) out[o] = jacobi_cell(f_coef(A, o), x, o, sy, sz, b[o], omega);
    else if (MODE == FM_AX) out[o] = ax_cell(f_coef(A, o), x, o, sy, sz);
    else if (MODE == FM_COPY) out[o] = x[o];
    else acc = x[o] * y[o];
  }
  if (MODE == FM_DOT) {
    const double s = f_wg_sum(acc, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
  }
}

// Sum of the partials: work-group w sums parts[w n .. w n + n) into out[w].
__global__ __launch_bounds__(256) void k_fb_reduce(const double* __restrict__ parts, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[4];
  const double s = sum_partials_wg256(parts + (long long)blockIdx.x * n, n, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------------------------------------------------------------
// sup.f90: box copies between two blocks' arrays or a packed buffer.
// ---------------------------------------------------------------------
enum FRegion { FR_COPY, FR_C2F, FR_F2C };

struct FBox {
  int i1d[3], i1s[3];  // first Fortran index of the box in dst and src
  int szc[3];          // box size
};

// sup_get_intp_index (sup.f90:18-33): coarse index and weight of fine
// position i (1-based) in a coarse block of n cells.
__device__ __forceinline__ void f_intp(int i, int n, int& ic, double& r) {
  ic = i / 2;
  r = 0.25 + 0.5 * double(i % 2);
  if (i < 2) {
    ic = 1;
    r = -0.25;
  } else if (i > 2 * n - 1) {
    ic = n - 1;
    r = 1.25;
  }
}

// SRCBUF / DSTBUF: the source / destination is a buffer of szc values,
// index i + szc(1) (j + szc(2) k), instead of a block array.
template <int OP, bool SRCBUF, bool DSTBUF>
__global__ __launch_bounds__(kFWG) void k_fb_region(double* __restrict__ dst, const double* __restrict__ src,
                                                    FBox box, FGeo g) {
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  const long long cnt = (long long)box.szc[0] * box.szc[1] * box.szc[2];
  if (c >= cnt) return;
  const int i = int(c % box.szc[0]);
  const long long t = c / box.szc[0];
  const int j = int(t % box.szc[1]);
  const int k = int(t / box.szc[1]);
  double v;
  if (OP == FR_COPY) {
    v = SRCBUF ? src[c] : src[f_at(g, i + box.i1s[0], j + box.i1s[1], k + box.i1s[2])];
  } else if (OP == FR_C2F) {
    int i1, j1, k1;
    double r, s, tt;
    f_intp(i + box.i1s[0], g.nx, i1, r);
    f_intp(j + box.i1s[1], g.ny, j1, s);
    f_intp(k + box.i1s[2], g.nz, k1, tt);
    const double nnn = src[f_at(g, i1, j1, k1)], pnn = src[f_at(g, i1 + 1, j1, k1)];
    const double npn = src[f_at(g, i1, j1 + 1, k1)], nnp = src[f_at(g, i1, j1, k1 + 1)];
    const double npp = src[f_at(g, i1, j1 + 1, k1 + 1)], pnp = src[f_at(g, i1 + 1, j1, k1 + 1)];
    const double ppn = src[f_at(g, i1 + 1, j1 + 1, k1)], ppp = src[f_at(g, i1 + 1, j1 + 1, k1 + 1)];
    // sup.f90:128-133
    v = (1.0 - tt) * ((1.0 - s) * ((1.0 - r) * nnn + r * pnn) + s * ((1.0 - r) * npn + r * ppn)) +
        tt * ((1.0 - s) * ((1.0 - r) * nnp + r * pnp) + s * ((1.0 - r) * npp + r * ppp));
  } else {
    const int i1 = 2 * (i + box.i1s[0]) - 1, j1 = 2 * (j + box.i1s[1]) - 1, k1 = 2 * (k + box.i1s[2]) - 1;
    const double nnn = src[f_at(g, i1, j1, k1)], pnn = src[f_at(g, i1 + 1, j1, k1)];
    const double npn = src[f_at(g, i1, j1 + 1, k1)], nnp = src[f_at(g, i1, j1, k1 + 1)];
    const double npp = src[f_at(g, i1, j1 + 1, k1 + 1)], pnp = src[f_at(g, i1 + 1, j1, k1 + 1)];
    const double ppn = src[f_at(g, i1 + 1, j1 + 1, k1)], ppp = src[f_at(g, i1 + 1, j1 + 1, k1 + 1)];
    // sup.f90:192-193, summed left to right
    v = 0.125 * (((((((nnn + pnn) + npn) + nnp) + npp) + pnp) + ppn) + ppp);
  }
  if (DSTBUF) dst[c] = v;
  else dst[f_at(g, i + box.i1d[0], j + box.i1d[1], k + box.i1d[2])] = v;
}

// ---------------------------------------------------------------------
// bcx.f90 / bca.f90 / bcax.f90: one thread per cell of a block face.
// Face f = 0..5 is X_M, X_P, Y_M, Y_P, Z_M, Z_P (the reference's x3, x1,
// x4, x2, x6, x5).  The boundary cell is index 1 on a minus face and sz
// on a plus face; ghost layer m is m cells beyond it.
// ---------------------------------------------------------------------
enum FFace { FF_D, FF_N, FF_DF, FF_PERIODIC, FF_AD, FF_AN, FF_POIS_U, FF_POIS_P, FF_POIS_AU, FF_POIS_AP };

struct FFaceArgs {
  int face;
  double xc;
  double center[3], org[3], csize[3];
};

// bc_get_u / bc_get_gp (bcax.f90:18-84), the reference's rx = x0(1) - x0(1)
// included.
__device__ __forceinline__ double f_pois_r(const FFaceArgs& f, int j, int k) {
  const double dx = f.csize[0];
  const double x0y = f.org[1] + (double(j) - 0.5) * dx;
  const double x0z = f.org[2] + (double(k) - 0.5) * dx;
  const double rx = 0.0, ry = x0y - f.center[1], rz = x0z - f.center[2];
  const double r2 = (rx * rx + ry * ry) + rz * rz;
  return sqrt(r2);
}

template <int OP>
__global__ __launch_bounds__(kFWG) void k_fb_face(double* __restrict__ x, double* __restrict__ Ap,
                                                  double* __restrict__ An, double* __restrict__ b, FFaceArgs f,
                                                  FGeo g) {
  const int axis = f.face >> 1;
  const bool plus = f.face & 1;
  const int n[3] = {g.nx, g.ny, g.nz};
  const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;  // tangential axes, in order
  // The periodic copies cover the whole tangential range (bcx.f90:575-582).
  const int lo = OP == FF_PERIODIC ? 1 - g.g : 1;
  const int wa = n[ta] + (OP == FF_PERIODIC ? 2 * g.g : 0);
  const int wb = n[tb] + (OP == FF_PERIODIC ? 2 * g.g : 0);
  const long long c = (long long)blockIdx.x * kFWG + threadIdx.x;
  if (c >= (long long)wa * wb) return;
  int idx[3];
  idx[ta] = lo + int(c % wa);
  idx[tb] = lo + int(c / wa);
  const long long step = axis == 0 ? 1 : axis == 1 ? (long long)g.cx : g.cxy;
  if (OP == FF_PERIODIC) {
    const int nn = n[axis];
    idx[axis] = 1;
    const long long o1 = f_at(g, idx[0], idx[1], idx[2]);  // index 1
    const long long on = o1 + (nn - 1) * step;             // index sz
    x[on + step] = x[o1];                                  // sz+1 <- 1
    x[on + 2 * step] = x[o1 + step];                       // sz+2 <- 2
    x[o1 - step] = x[on];                                  // 0 <- sz
    x[o1 - 2 * step] = x[on - step];                       // -1 <- sz-1
    return;
  }
  idx[axis] = plus ? n[axis] : 1;
  const long long o = f_at(g, idx[0], idx[1], idx[2]);
  const long long s1 = plus ? step : -step;
  const double xc = f.xc;
  if (OP == FF_D) {  // bcx.f90:18-196
    x[o + s1] = 2.0 * xc - x[o];
    x[o + 2 * s1] = xc;
  } else if (OP == FF_N) {  // bcx.f90:198-376
    if (plus) {
      x[o + s1] = x[o] + xc;
      x[o + 2 * s1] = x[o] + xc * 2.0;
    } else {
      x[o + s1] = x[o] - xc;
      x[o + 2 * s1] = x[o] - xc * 2.0;
    }
  } else if (OP == FF_DF) {  // bcx.f90:378-559: a plus face also sets its boundary cell
    if (plus) x[o] = xc;
    x[o + s1] = xc;
    x[o + 2 * s1] = xc;
  } else if (OP == FF_AD) {  // bca.f90:12-202
    Ap[o] = Ap[o] - An[o];
    b[o] = b[o] - An[o] * xc * 2.0;
    An[o] = 0.0;
  } else if (OP == FF_AN) {  // bca.f90:204-394: b += Anb xc on a minus face, -= on a plus face
    Ap[o] = Ap[o] + An[o];
    b[o] = plus ? b[o] - An[o] * xc : b[o] + An[o] * xc;
    An[o] = 0.0;
  } else {  // bcax.f90:86-280, X_M face (i = 1)
    const double r = f_pois_r(f, idx[1], idx[2]);
    if (OP == FF_POIS_U || OP == FF_POIS_AU) {
      double u = xc * 2.0 * (1.0 - r * r / (0.5 * 0.5));
      if (r > 0.5) u = 0.0;
      if (OP == FF_POIS_U) {
        x[o - 1] = 2.0 * u - x[o];
        x[o - 2] = u;
      } else {
        Ap[o] = Ap[o] - An[o];
        b[o] = b[o] - An[o] * u * 2.0;
        An[o] = 0.0;
      }
    } else {
      double gp = xc * f.csize[0];
      if (r > 0.5) gp = 0.0;
      if (OP == FF_POIS_P) {
        x[o - 1] = x[o] - gp;
        x[o - 2] = x[o] - gp * 2.0;
      } else {
        Ap[o] = Ap[o] + An[o];
        b[o] = b[o] + An[o] * gp;
        An[o] = 0.0;
      }
    }
  }
}

// ---------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------
struct ThreadState {
  hipStream_t stream = nullptr;
  double* parts = nullptr;
  size_t nparts = 0;
};
thread_local ThreadState t_state[kFMaxDevices];

[[noreturn]] void die(const char* fn, const char* why) {
  std::fprintf(stderr, "bcm: %s: %s\n", fn, why);
  std::fflush(stderr);
  std::exit(1);
}

FGeo geo_of(const char* fn, const int* sz, const int* g, int min_g) {
  if (!sz || !g) die(fn, "NULL size argument");
  if (sz[0] < 1 || sz[1] < 1 || sz[2] < 1) die(fn, "block size must be positive");
  if (*g < min_g) {
    char msg[96];
    std::snprintf(msg, sizeof(msg), "needs at least %d ghost layer%s (g = %d)", min_g, min_g == 1 ? "" : "s", *g);
    die(fn, msg);
  }
  FGeo o;
  o.nx = sz[0];
  o.ny = sz[1];
  o.nz = sz[2];
  o.g = *g;
  o.cx = sz[0] + 2 * *g;
  o.cxy = (long long)o.cx * (sz[1] + 2 * *g);
  o.vol = o.cxy * (sz[2] + 2 * *g);
  o.ncell = (long long)sz[0] * sz[1] * sz[2];
  if (o.vol > (1LL << 31) || o.cxy > (1LL << 30)) die(fn, "block too large");
  o.lo = o.cxy * *g + (long long)o.cx * *g + *g;
  o.hi = o.vol - 1 - o.lo;
  return o;
}

ThreadState& state(const char* fn, size_t nparts) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kFMaxDevices) die(fn, "no HIP device");
  ThreadState& s = t_state[dev];
  if (!s.stream && hipStreamCreate(&s.stream) != hipSuccess) die(fn, "hipStreamCreate failed");
  if (nparts > s.nparts) {
    if (s.parts) {
      (void)hipStreamSynchronize(s.stream);
      dev_free(s.parts);
      s.parts = nullptr;
      s.nparts = 0;
    }
    const size_t want = nparts < 1024 ? 1024 : nparts;
    if (dev_alloc(reinterpret_cast<void**>(&s.parts), want * sizeof(double)) != BCM_OK)
      die(fn, "device allocation failed");
    s.nparts = want;
  }
  return s;
}

// The `bytes` bytes at p must lie inside one device allocation.
void need(const char* fn, const char* what, const void* p, size_t bytes) {
  if (!p) {
    char msg[96];
    std::snprintf(msg, sizeof(msg), "argument %s is NULL", what);
    die(fn, msg);
  }
  int dev = -1;
  hipPointerAttribute_t attr{};
  const bool on_device = hipGetDevice(&dev) == hipSuccess && hipPointerGetAttributes(&attr, p) == hipSuccess &&
                         ((attr.type == hipMemoryTypeDevice && attr.device == dev) ||
                          attr.type == hipMemoryTypeManaged);
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  const hipError_t e = on_device ? hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) : hipErrorInvalidValue;
  if (e != hipSuccess) {
    (void)hipGetLastError();
    char msg[128];
    std::snprintf(msg, sizeof(msg), "argument %s is not device memory of the current device", what);
    die(fn, msg);
  }
  const char* b = static_cast<const char*>(base);
  const char* q = static_cast<const char*>(p);
  if (q < b || size_t(q - b) + bytes > size) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "argument %s: %zu bytes from it run past the end of its allocation", what,
                  bytes);
    die(fn, msg);
  }
}

void need_field(const char* fn, const char* what, const void* p, const FGeo& g) {
  need(fn, what, p, size_t(g.vol) * sizeof(double));
}

void need_coef(const char* fn, const FGeo& g, double* Ap, double* Aw, double* Ae, double* As, double* An,
               double* Ab, double* At) {
  need_field(fn, "Ap", Ap, g);
  need_field(fn, "Aw", Aw, g);
  need_field(fn, "Ae", Ae, g);
  need_field(fn, "As", As, g);
  need_field(fn, "An", An, g);
  need_field(fn, "Ab", Ab, g);
  need_field(fn, "At", At, g);
}

unsigned blocks_for(long long n) { return unsigned((n + kFWG - 1) / kFWG); }

void finish(const char* fn, ThreadState& s) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) die(fn, hipGetErrorString(e));
  const hipError_t w = hipStreamSynchronize(s.stream);
  if (w != hipSuccess) die(fn, hipGetErrorString(w));
}

// Reduces `nsum` runs of `grid` partials and returns them to host memory.
void reduce_to_host(const char* fn, ThreadState& s, int grid, int nsum, double* out) {
  double* res = s.parts + size_t(nsum) * grid;
  hipLaunchKernelGGL(k_fb_reduce, dim3(nsum), dim3(256), 0, s.stream, (const double*)s.parts, grid, res);
  double v[2] = {0.0, 0.0};
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(v, res, sizeof(double) * nsum, hipMemcpyDeviceToHost, s.stream) != hipSuccess)
    die(fn, "reduction failed");
  finish(fn, s);
  for (int q = 0; q < nsum; ++q) out[q] = v[q];
}

FCoef coef7(double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab, double* At) {
  return FCoef{Ap, Aw, Ae, As, An, Ab, At, nullptr};
}

template <int OP>
void run_blas(const char* fn, double* z, double* x, double* y, double a, double b, const int* sz, const int* g) {
  const FGeo geo = geo_of(fn, sz, g, 0);
  need_field(fn, "z", z, geo);
  if (OP != FB_SCAL) need_field(fn, "x", x, geo);
  if (OP == FB_AXPYZ || OP == FB_AXPBYPZ || OP == FB_ADD || OP == FB_TRIAD) need_field(fn, "y", y, geo);
  ThreadState& s = state(fn, 0);
  hipLaunchKernelGGL(k_fb_blas<OP>, dim3(blocks_for(geo.ncell)), dim3(kFWG), 0, s.stream, z, (const double*)x,
                     (const double*)y, a, b, geo);
  finish(fn, s);
}

template <int MODE>
void run_stencil(const char* fn, double* out, double* x, FCoef A, double* b, double omega, const int* sz,
                 const int* g, double* rr) {
  const FGeo geo = geo_of(fn, sz, g, 1);
  if (A.packed) need(fn, "A", A.packed, size_t(geo.vol) * 7 * sizeof(double));
  else need_coef(fn, geo, (double*)A.p, (double*)A.w, (double*)A.e, (double*)A.s, (double*)A.n, (double*)A.b,
                 (double*)A.t);
  need_field(fn, "x", x, geo);
  if (MODE != FS_AX) need_field(fn, "b", b, geo);
  if (MODE != FS_R2) need_field(fn, MODE == FS_JACOBI ? "x1" : "output", out, geo);
  if (MODE == FS_R2 && !rr) die(fn, "NULL result argument");
  const int grid = int(blocks_for(geo.ncell));
  ThreadState& s = state(fn, MODE == FS_R2 ? size_t(grid) + 1 : 0);
  hipLaunchKernelGGL(k_fb_stencil<MODE>, dim3(grid), dim3(kFWG), 0, s.stream, out, (const double*)x, A,
                     (const double*)b, omega, geo, s.parts);
  if (MODE == FS_R2) reduce_to_host(fn, s, grid, 1, rr);
  else finish(fn, s);
}

template <int MODE>
void run_masked(const char* fn, double* out, double* x, double* y, FCoef A, double* b, int* mask, double omega,
                const int* sz, const int* g, double* xy) {
  const FGeo geo = geo_of(fn, sz, g, MODE == FM_JACOBI || MODE == FM_AX ? 1 : 0);
  need(fn, "mask", mask, size_t(geo.vol) * sizeof(int));
  need_field(fn, "x", x, geo);
  if (MODE == FM_JACOBI || MODE == FM_AX)
    need_coef(fn, geo, (double*)A.p, (double*)A.w, (double*)A.e, (double*)A.s, (double*)A.n, (double*)A.b,
              (double*)A.t);
  if (MODE == FM_JACOBI) need_field(fn, "b", b, geo);
  if (MODE == FM_DOT) need_field(fn, "y", y, geo);
  else need_field(fn, "output", out, geo);
  if (MODE == FM_DOT && !xy) die(fn, "NULL result argument");
  const long long count = geo.hi >= geo.lo ? geo.hi - geo.lo + 1 : 0;
  const int grid = count > 0 ? int(blocks_for(count)) : 0;
  ThreadState& s = state(fn, MODE == FM_DOT ? size_t(grid) + 1 : 0);
  if (grid == 0) {
    if (MODE == FM_DOT) *xy = 0.0;
    return;
  }
  hipLaunchKernelGGL(k_fb_masked<MODE>, dim3(grid), dim3(kFWG), 0, s.stream, out, (const double*)x,
                     (const double*)y, A, (const double*)b, (const int*)mask, omega, geo, s.parts);
  if (MODE == FM_DOT) reduce_to_host(fn, s, grid, 1, xy);
  else finish(fn, s);
}

void run_dot(const char* fn, double* out, double* x, double* y, double* z, const int* sz, const int* g) {
  const FGeo geo = geo_of(fn, sz, g, 0);
  if (!out) die(fn, "NULL result argument");
  need_field(fn, "x", x, geo);
  need_field(fn, "y", y, geo);
  if (z) need_field(fn, "z", z, geo);
  const int grid = int(blocks_for(geo.ncell));
  const int nsum = z ? 2 : 1;
  ThreadState& s = state(fn, size_t(nsum) * grid + nsum);
  if (z)
    hipLaunchKernelGGL(k_fb_dot<2>, dim3(grid), dim3(kFWG), 0, s.stream, (const double*)x, (const double*)y,
                       (const double*)z, geo, s.parts);
  else
    hipLaunchKernelGGL(k_fb_dot<1>, dim3(grid), dim3(kFWG), 0, s.stream, (const double*)x, (const double*)y,
                       (const double*)nullptr, geo, s.parts);
  reduce_to_host(fn, s, grid, nsum, out);
}

// The index range [lo, hi] that a box of `n` positions from `first`
// touches along one axis of a source or destination array.
bool inside(int lo, int hi, const FGeo& g, int axis) {
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  return lo >= 1 - g.g && hi <= n + g.g;
}

template <int OP, bool SRCBUF, bool DSTBUF>
void run_region(const char* fn, double* dst, const int* i1_dst, double* src, const int* i1_src, const int* sz_c,
                const int* sz, const int* g) {
  const FGeo geo = geo_of(fn, sz, g, 0);
  if (!sz_c || (!DSTBUF && !i1_dst) || (!SRCBUF && !i1_src)) die(fn, "NULL index argument");
  if (sz_c[0] < 0 || sz_c[1] < 0 || sz_c[2] < 0) die(fn, "negative copy size");
  if (OP == FR_C2F && (geo.nx < 2 || geo.ny < 2 || geo.nz < 2))
    die(fn, "interpolation needs at least 2 cells along every axis");
  const long long cnt = (long long)sz_c[0] * sz_c[1] * sz_c[2];
  if (cnt == 0) return;
  FBox box{};
  for (int a = 0; a < 3; ++a) {
    box.szc[a] = sz_c[a];
    box.i1d[a] = DSTBUF ? 0 : i1_dst[a];
    box.i1s[a] = SRCBUF ? 0 : i1_src[a];
    if (!DSTBUF && !inside(i1_dst[a], i1_dst[a] + sz_c[a] - 1, geo, a))
      die(fn, "destination box outside the block array");
    if (!SRCBUF) {
      const int first = i1_src[a], last = i1_src[a] + sz_c[a] - 1;
      const int n = a == 0 ? geo.nx : a == 1 ? geo.ny : geo.nz;
      int lo = first, hi = last;
      if (OP == FR_F2C) {
        lo = 2 * first - 1;
        hi = 2 * last;
      } else if (OP == FR_C2F) {
        // The coarse index is non-decreasing in the fine position.
        auto ic = [n](int i) { return i < 2 ? 1 : i > 2 * n - 1 ? n - 1 : i / 2; };
        lo = ic(first);
        hi = ic(last) + 1;
      }
      if (!inside(lo, hi, geo, a)) die(fn, "source box outside the block array");
    }
  }
  const size_t buf_bytes = size_t(cnt) * sizeof(double);
  if (DSTBUF) need(fn, "buffer", dst, buf_bytes);
  else need_field(fn, "data_dst", dst, geo);
  if (SRCBUF) need(fn, "buffer", src, buf_bytes);
  else need_field(fn, "data_src", src, geo);
  ThreadState& s = state(fn, 0);
  hipLaunchKernelGGL((k_fb_region<OP, SRCBUF, DSTBUF>), dim3(blocks_for(cnt)), dim3(kFWG), 0, s.stream, dst,
                     (const double*)src, box, geo);
  finish(fn, s);
}

template <int OP>
void run_face(const char* fn, int face, double* x, double* Ap, double* An, double* b, double xc, const int* sz,
              const int* g, const double* center = nullptr, const double* org = nullptr,
              const double* csize = nullptr) {
  const bool coef = OP == FF_AD || OP == FF_AN || OP == FF_POIS_AU || OP == FF_POIS_AP;
  const FGeo geo = geo_of(fn, sz, g, coef ? 0 : 2);
  const int axis = face >> 1;
  const int n[3] = {geo.nx, geo.ny, geo.nz};
  if (OP == FF_PERIODIC && n[axis] < 2) die(fn, "the periodic copy needs at least 2 cells along its axis");
  if (coef) {
    need_field(fn, "Ap", Ap, geo);
    need_field(fn, "Anb", An, geo);
    need_field(fn, "b", b, geo);
  } else {
    need_field(fn, "x", x, geo);
  }
  FFaceArgs f{};
  f.face = face;
  f.xc = xc;
  if (center) {
    if (!org || !csize) die(fn, "NULL geometry argument");
    for (int a = 0; a < 3; ++a) {
      f.center[a] = center[a];
      f.org[a] = org[a];
      f.csize[a] = csize[a];
    }
  }
  const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
  const long long wide = OP == FF_PERIODIC ? 2 * geo.g : 0;
  const long long cnt = (n[ta] + wide) * (n[tb] + wide);
  ThreadState& s = state(fn, 0);
  hipLaunchKernelGGL(k_fb_face<OP>, dim3(blocks_for(cnt)), dim3(kFWG), 0, s.stream, x, Ap, An, b, f, geo);
  finish(fn, s);
}

}  // namespace
}  // namespace bcm

using namespace bcm;

// blas.h ---------------------------------------------------------------
extern "C" void fill_(double* x, double* a, int* sz, int* g) {
  const FGeo geo = geo_of("fill_", sz, g, 0);
  if (!a) die("fill_", "NULL value argument");
  need_field("fill_", "x", x, geo);
  ThreadState& s = state("fill_", 0);
  hipLaunchKernelGGL(k_fb_fill<double>, dim3(blocks_for(geo.vol)), dim3(kFWG), 0, s.stream, x, *a, geo.vol);
  finish("fill_", s);
}

extern "C" void fill_vf3d_(double* x, double* a, int* sz, int* g, int* ne) {
  const FGeo geo = geo_of("fill_vf3d_", sz, g, 0);
  if (!a || !ne || *ne < 0) die("fill_vf3d_", "bad value or component count");
  const long long count = geo.vol * *ne;
  if (count == 0) return;
  need("fill_vf3d_", "x", x, size_t(count) * sizeof(double));
  ThreadState& s = state("fill_vf3d_", 0);
  hipLaunchKernelGGL(k_fb_fill<double>, dim3(blocks_for(count)), dim3(kFWG), 0, s.stream, x, *a, count);
  finish("fill_vf3d_", s);
}

extern "C" void copy_(double* y, double* x, int* sz, int* g) {
  const FGeo geo = geo_of("copy_", sz, g, 0);
  need_field("copy_", "y", y, geo);
  need_field("copy_", "x", x, geo);
  ThreadState& s = state("copy_", 0);
  hipLaunchKernelGGL(k_fb_copy<double>, dim3(blocks_for(geo.vol)), dim3(kFWG), 0, s.stream, y, (const double*)x,
                     geo.vol);
  finish("copy_", s);
}

extern "C" void copy_integer_(int* y, int* x, int* sz, int* g) {
  const FGeo geo = geo_of("copy_integer_", sz, g, 0);
  need("copy_integer_", "y", y, size_t(geo.vol) * sizeof(int));
  need("copy_integer_", "x", x, size_t(geo.vol) * sizeof(int));
  ThreadState& s = state("copy_integer_", 0);
  hipLaunchKernelGGL(k_fb_copy<int>, dim3(blocks_for(geo.vol)), dim3(kFWG), 0, s.stream, y, (const int*)x,
                     geo.vol);
  finish("copy_integer_", s);
}

extern "C" void add_(double* C, double* A, double* B, int* sz, int* g) {
  run_blas<FB_ADD>("add_", C, A, B, 0.0, 0.0, sz, g);
}
extern "C" void triad_(double* C, double* A, double* B, double* d, int* sz, int* g) {
  if (!d) die("triad_", "NULL scalar argument");
  run_blas<FB_TRIAD>("triad_", C, A, B, *d, 0.0, sz, g);
}
extern "C" void scal_(double* y, double* a, int* sz, int* g) {
  if (!a) die("scal_", "NULL scalar argument");
  run_blas<FB_SCAL>("scal_", y, nullptr, nullptr, *a, 0.0, sz, g);
}
extern "C" void axpy_(double* y, double* x, double* a, int* sz, int* g) {
  if (!a) die("axpy_", "NULL scalar argument");
  run_blas<FB_AXPY>("axpy_", y, x, nullptr, *a, 0.0, sz, g);
}
extern "C" void xpay_(double* y, double* x, double* a, int* sz, int* g) {
  if (!a) die("xpay_", "NULL scalar argument");
  run_blas<FB_XPAY>("xpay_", y, x, nullptr, *a, 0.0, sz, g);
}
extern "C" void axpyz_(double* z, double* x, double* y, double* a, int* sz, int* g) {
  if (!a) die("axpyz_", "NULL scalar argument");
  run_blas<FB_AXPYZ>("axpyz_", z, x, y, *a, 0.0, sz, g);
}
extern "C" void axpbypz_(double* z, double* x, double* y, double* a, double* b, int* sz, int* g) {
  if (!a || !b) die("axpbypz_", "NULL scalar argument");
  run_blas<FB_AXPBYPZ>("axpbypz_", z, x, y, *a, *b, sz, g);
}
extern "C" void dot_(double* xy, double* y, double* x, int* sz, int* g) { run_dot("dot_", xy, x, y, nullptr, sz, g); }
extern "C" void dotx2_(double* xy, double* xz, double* x, double* y, double* z, int* sz, int* g) {
  if (!xy || !xz) die("dotx2_", "NULL result argument");
  double v[2];
  run_dot("dotx2_", v, x, y, z, sz, g);
  *xy = v[0];
  *xz = v[1];
}

extern "C" void jacobi_smoother_(double* x1, double* x0, double* Ap, double* Aw, double* Ae, double* As, double* An,
                                 double* Ab, double* At, double* b, double* omega, int* sz, int* g) {
  if (!omega) die("jacobi_smoother_", "NULL omega");
  if (x1 == x0) die("jacobi_smoother_", "x1 and x0 must be different arrays");
  run_stencil<FS_JACOBI>("jacobi_smoother_", x1, x0, coef7(Ap, Aw, Ae, As, An, Ab, At), b, *omega, sz, g, nullptr);
}
extern "C" void jacobi_smoother2_(double* x1, double* x0, double* A, double* b, double* omega, int* sz, int* g) {
  if (!omega) die("jacobi_smoother2_", "NULL omega");
  if (x1 == x0) die("jacobi_smoother2_", "x1 and x0 must be different arrays");
  FCoef c{};
  c.packed = A;
  if (!A) die("jacobi_smoother2_", "argument A is NULL");
  run_stencil<FS_JACOBI>("jacobi_smoother2_", x1, x0, c, b, *omega, sz, g, nullptr);
}
extern "C" void calc_ax_(double* Ax, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab,
                         double* At, double* x, int* sz, int* g) {
  if (Ax == x) die("calc_ax_", "Ax and x must be different arrays");
  run_stencil<FS_AX>("calc_ax_", Ax, x, coef7(Ap, Aw, Ae, As, An, Ab, At), nullptr, 0.0, sz, g, nullptr);
}
extern "C" void calc_r_(double* r, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab,
                        double* At, double* x, double* b, int* sz, int* g) {
  if (r == x) die("calc_r_", "r and x must be different arrays");
  run_stencil<FS_R>("calc_r_", r, x, coef7(Ap, Aw, Ae, As, An, Ab, At), b, 0.0, sz, g, nullptr);
}
extern "C" void calc_r2_(double* rr, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab,
                         double* At, double* x, double* b, int* sz, int* g) {
  run_stencil<FS_R2>("calc_r2_", nullptr, x, coef7(Ap, Aw, Ae, As, An, Ab, At), b, 0.0, sz, g, rr);
}

extern "C" void setup_mask_(double* m, int* mask, int* sz, int* g) {
  const FGeo geo = geo_of("setup_mask_", sz, g, 0);
  need_field("setup_mask_", "m", m, geo);
  need("setup_mask_", "mask", mask, size_t(geo.vol) * sizeof(int));
  ThreadState& s = state("setup_mask_", 0);
  hipLaunchKernelGGL(k_fb_setup_mask, dim3(blocks_for(geo.vol)), dim3(kFWG), 0, s.stream, m, mask, geo);
  finish("setup_mask_", s);
}
extern "C" void copy_mask_(double* y, double* x, int* mask, int* sz, int* g) {
  run_masked<FM_COPY>("copy_mask_", y, x, nullptr, FCoef{}, nullptr, mask, 0.0, sz, g, nullptr);
}
extern "C" void dot_mask_(double* xy, double* y, double* x, int* mask, int* sz, int* g) {
  run_masked<FM_DOT>("dot_mask_", nullptr, x, y, FCoef{}, nullptr, mask, 0.0, sz, g, xy);
}
extern "C" void calc_ax_mask_(double* Ax, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab,
                              double* At, double* x, int* mask, int* sz, int* g) {
  if (Ax == x) die("calc_ax_mask_", "Ax and x must be different arrays");
  run_masked<FM_AX>("calc_ax_mask_", Ax, x, nullptr, coef7(Ap, Aw, Ae, As, An, Ab, At), nullptr, mask, 0.0, sz, g,
                    nullptr);
}
extern "C" void jacobi_smoother_mask_(double* x1, double* x0, double* Ap, double* Aw, double* Ae, double* As,
                                      double* An, double* Ab, double* At, double* b, int* mask, double* omega,
                                      int* sz, int* g) {
  if (!omega) die("jacobi_smoother_mask_", "NULL omega");
  if (x1 == x0) die("jacobi_smoother_mask_", "x1 and x0 must be different arrays");
  run_masked<FM_JACOBI>("jacobi_smoother_mask_", x1, x0, nullptr, coef7(Ap, Aw, Ae, As, An, Ab, At), b, mask,
                        *omega, sz, g, nullptr);
}

// comm.h (one rank) ----------------------------------------------------
extern "C" void allreduce_(double* total, double* a) {
  if (!total || !a) die("allreduce_", "NULL argument");
  *total = *a;
}
extern "C" void allreduce_max_(double* total, double* a) {
  if (!total || !a) die("allreduce_max_", "NULL argument");
  *total = *a;
}
extern "C" void allreduce_min_(double* total, double* a) {
  if (!total || !a) die("allreduce_min_", "NULL argument");
  *total = *a;
}
extern "C" void dot_all_(double* xy, double* y, double* x, int* sz, int* g) {
  run_dot("dot_all_", xy, x, y, nullptr, sz, g);
}

// sup.h ----------------------------------------------------------------
extern "C" void sup_copy_from_neighbor_(double* data_dst, int* i1_dst, double* data_src, int* i1_src, int* sz_c,
                                        int* sz, int* g) {
  run_region<FR_COPY, false, false>("sup_copy_from_neighbor_", data_dst, i1_dst, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_from_neighbor_c2f_(double* data_dst, int* i1_dst, double* data_src, int* i1_src,
                                            int* sz_c, int* sz, int* g) {
  run_region<FR_C2F, false, false>("sup_copy_from_neighbor_c2f_", data_dst, i1_dst, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_from_neighbor_f2c_(double* data_dst, int* i1_dst, double* data_src, int* i1_src,
                                            int* sz_c, int* sz, int* g) {
  run_region<FR_F2C, false, false>("sup_copy_from_neighbor_f2c_", data_dst, i1_dst, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_to_buffer_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz, int* g) {
  run_region<FR_COPY, false, true>("sup_copy_to_buffer_", buffer, nullptr, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_to_buffer_c2f_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz,
                                        int* g) {
  run_region<FR_C2F, false, true>("sup_copy_to_buffer_c2f_", buffer, nullptr, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_to_buffer_f2c_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz,
                                        int* g) {
  run_region<FR_F2C, false, true>("sup_copy_to_buffer_f2c_", buffer, nullptr, data_src, i1_src, sz_c, sz, g);
}
extern "C" void sup_copy_from_buffer_(double* data_dst, int* i1_dst, double* buffer, int* sz_c, int* sz, int* g) {
  run_region<FR_COPY, true, false>("sup_copy_from_buffer_", data_dst, i1_dst, buffer, nullptr, sz_c, sz, g);
}

// bcx.h: ghost values on one face ---------------------------------------
#define BCM_BCX(name, op, face)                                                  \
  extern "C" void name(double* x, double* xc, int* sz, int* g) {                 \
    if (!xc) die(#name, "NULL value argument");                                  \
    run_face<op>(#name, face, x, nullptr, nullptr, nullptr, *xc, sz, g);         \
  }
BCM_BCX(bc_x1_d_, FF_D, 1)
BCM_BCX(bc_x3_d_, FF_D, 0)
BCM_BCX(bc_x2_d_, FF_D, 3)
BCM_BCX(bc_x4_d_, FF_D, 2)
BCM_BCX(bc_x5_d_, FF_D, 5)
BCM_BCX(bc_x6_d_, FF_D, 4)
BCM_BCX(bc_x1_d_f_, FF_DF, 1)
BCM_BCX(bc_x3_d_f_, FF_DF, 0)
BCM_BCX(bc_x2_d_f_, FF_DF, 3)
BCM_BCX(bc_x4_d_f_, FF_DF, 2)
BCM_BCX(bc_x5_d_f_, FF_DF, 5)
BCM_BCX(bc_x6_d_f_, FF_DF, 4)
BCM_BCX(bc_x1_n_, FF_N, 1)
BCM_BCX(bc_x3_n_, FF_N, 0)
BCM_BCX(bc_x2_n_, FF_N, 3)
BCM_BCX(bc_x4_n_, FF_N, 2)
BCM_BCX(bc_x5_n_, FF_N, 5)
BCM_BCX(bc_x6_n_, FF_N, 4)
// The periodic copies ignore xc (bcx.f90:561-652).
extern "C" void bc_x1_p_(double* x, double* xc, int* sz, int* g) {
  (void)xc;
  run_face<FF_PERIODIC>("bc_x1_p_", 1, x, nullptr, nullptr, nullptr, 0.0, sz, g);
}
extern "C" void bc_x2_p_(double* x, double* xc, int* sz, int* g) {
  (void)xc;
  run_face<FF_PERIODIC>("bc_x2_p_", 3, x, nullptr, nullptr, nullptr, 0.0, sz, g);
}
extern "C" void bc_x5_p_(double* x, double* xc, int* sz, int* g) {
  (void)xc;
  run_face<FF_PERIODIC>("bc_x5_p_", 5, x, nullptr, nullptr, nullptr, 0.0, sz, g);
}
#undef BCM_BCX

// bca.h: a boundary condition folded into one face's coefficients -------
#define BCM_BCA(name, op, face)                                                  \
  extern "C" void name(double* Ap, double* Anb, double* b, double* xc, int* sz, int* g) { \
    if (!xc) die(#name, "NULL value argument");                                  \
    run_face<op>(#name, face, nullptr, Ap, Anb, b, *xc, sz, g);                  \
  }
BCM_BCA(bc_aw_d_, FF_AD, 0)
BCM_BCA(bc_ae_d_, FF_AD, 1)
BCM_BCA(bc_as_d_, FF_AD, 2)
BCM_BCA(bc_an_d_, FF_AD, 3)
BCM_BCA(bc_ab_d_, FF_AD, 4)
BCM_BCA(bc_at_d_, FF_AD, 5)
BCM_BCA(bc_aw_n_, FF_AN, 0)
BCM_BCA(bc_ae_n_, FF_AN, 1)
BCM_BCA(bc_as_n_, FF_AN, 2)
BCM_BCA(bc_an_n_, FF_AN, 3)
BCM_BCA(bc_ab_n_, FF_AN, 4)
BCM_BCA(bc_at_n_, FF_AN, 5)
#undef BCM_BCA

// bcax.h: the Poiseuille inflow on the X_M face -------------------------
extern "C" void bc_x3_poiseuille_u_(double* x, double* xc, int* sz, int* g, double* center, double* org,
                                    double* bsize, double* csize) {
  (void)bsize;
  if (!xc || !center) die("bc_x3_poiseuille_u_", "NULL argument");
  run_face<FF_POIS_U>("bc_x3_poiseuille_u_", 0, x, nullptr, nullptr, nullptr, *xc, sz, g, center, org, csize);
}
extern "C" void bc_aw_poiseuille_u_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g, double* center,
                                    double* org, double* bsize, double* csize) {
  (void)bsize;
  if (!xc || !center) die("bc_aw_poiseuille_u_", "NULL argument");
  run_face<FF_POIS_AU>("bc_aw_poiseuille_u_", 0, nullptr, Ap, Aw, b, *xc, sz, g, center, org, csize);
}
extern "C" void bc_x3_poiseuille_p_(double* x, double* xc, int* sz, int* g, double* center, double* org,
                                    double* bsize, double* csize) {
  (void)bsize;
  if (!xc || !center) die("bc_x3_poiseuille_p_", "NULL argument");
  run_face<FF_POIS_P>("bc_x3_poiseuille_p_", 0, x, nullptr, nullptr, nullptr, *xc, sz, g, center, org, csize);
}
extern "C" void bc_aw_poiseuille_p_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g, double* center,
                                    double* org, double* bsize, double* csize) {
  (void)bsize;
  if (!xc || !center) die("bc_aw_poiseuille_p_", "NULL argument");
  run_face<FF_POIS_AP>("bc_aw_poiseuille_p_", 0, nullptr, Ap, Aw, b, *xc, sz, g, center, org, csize);
}
