// Internal definitions shared by the HIP translation units of libbcm_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "bcm_hip.h"

#define BCM_HIP_TRY(expr)                                                            \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "bcm: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                              \
      return BCM_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define BCM_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != BCM_OK) return rc_; \
  } while (0)

namespace bcm {

// Per-block constants used on the device.
struct BlockMeta {
  double cx, cy, cz;      // 1/h^2 per axis, h = block_size / n
  double hx;              // cell size along x, the scale of a Neumann value
  int32_t dir_mask;       // bit f: face f carries a Dirichlet condition
  int32_t neu_mask;       // bit f: face f carries a Neumann condition
  int32_t bc_id[6];       // boundary-value index per face (-1 if none)
};

// One ghost-fill task: dst block, src block, code = face | subface<<3 | kind<<5.
enum TaskKind { KIND_COPY = 0, KIND_F2C = 1, KIND_C2F = 2 };

struct EventPair {
  hipEvent_t start, stop;
  int kernel;
};

struct Krylov;  // fused solver state (bcm_krylov.hip)

}  // namespace bcm

struct bcm_mesh {
  int n = 0, vc = 0, nblocks = 0;
  int nx0 = 0;                    // n + 2 vc
  size_t bvol = 0;                // (n+2vc)^3 values per block
  size_t ncell = 0;               // nblocks * n^3 interior cells
  hipStream_t stream = nullptr;
  int device = 0;
  std::vector<bcm_block_info> blocks;
  std::vector<bcm::BlockMeta> meta;

  // Ghost-fill tasks ordered by [axis][kind]; off[a][k] .. off[a][k+1].
  int4* d_tasks = nullptr;
  std::vector<int4> h_tasks;      // the same table on the host
  int ntask_total = 0;
  int off[3][4] = {};
  int kind_count[3] = {};
  // Boundary face tasks {block, face, bc_id, type}.
  int4* d_bc_tasks = nullptr;
  int n_bc_tasks = 0;
  int n_dirichlet = 0;
  int n_neumann = 0;
  // Blocks with a face on the domain boundary {block, mask of those faces},
  // built on first use by the per-field face conditions.
  int2* d_outer = nullptr;
  int n_outer = -1;
  bcm::BlockMeta* d_meta = nullptr;

  // Reduction scratch: one double per reduction work-group, plus results.
  double* d_partials = nullptr;
  size_t n_partials = 0;
  double* d_result = nullptr;     // 16 doubles
  double* h_result = nullptr;     // pinned mirror of d_result

  // Boundary-value table on the device.
  double* d_bvalues = nullptr;
  int n_bvalues_cap = 0;
  std::vector<double> h_bvalues;  // last table uploaded (skip identical re-uploads)
  bool bvalues_valid = false;
  int need_bvalues = 0;           // 1 + largest Dirichlet bc_id

  // SOR coefficient cache (per block: c0, cx, cy, cz) for one omega/form.
  double* d_sor_coef = nullptr;
  double sor_omega = 0.0;
  int sor_form = -1;

  // Timing.
  bool timing = false;
  double kernel_ms[BCM_K_COUNT] = {};
  long long kernel_launches[BCM_K_COUNT] = {};
  std::vector<bcm::EventPair> pending;
  std::vector<hipEvent_t> event_pool;

  // Solver workspace (lazily allocated, kept for the mesh's lifetime).
  std::vector<void*> ws;

  // Halo-gathering sweep: per block, 6 faces x 8 ints
  // {kind, nb0, nb1, nb2, nb3, subface, bc_id, 0}; usable when every face
  // has a neighbour on each subface or a Dirichlet condition.
  int* d_desc = nullptr;
  bool halo_ok = false;
  // X-face column exports (columns 0, 1, n-2, n-1 of every block), two
  // buffers for the ping-pong of iterates; bytes sized for FP64.
  void* d_xexp[2] = {nullptr, nullptr};
  // The Jacobi form of the halo schedule: every face has a neighbour on
  // every subface (copy, F2C with all four, C2F) or none at all.
  bool jhalo_ok = false;

  // Segments: independent meshes sharing the arena, segment s owning blocks
  // seg_first[s] .. seg_first[s+1]-1 (one segment for a plain mesh).
  int nseg = 1;
  std::vector<int> seg_first;
  int* d_seg_first = nullptr;
  int* d_blk_seg = nullptr;

  // Fused Krylov engine (vectors, per-segment states, captured graph).
  bcm::Krylov* kry = nullptr;
  // Progress lines of long solves on stderr (seconds between lines, 0: off).
  double progress_s = 0.0;
  // Scheduling of bcm_sor_iterate (BCM_SOR_*), set by hand or by the autotuner.
  int sor_variant = 0;
  // Live fields of the mesh: destroying the mesh frees their device memory
  // and detaches them, so a later bcm_field_destroy only frees the handle.
  std::vector<struct bcm_field*> fields;
};

namespace bcm {
enum FaceKind { FD_NONE = 0, FD_COPY = 1, FD_F2C = 2, FD_C2F = 3, FD_DIR = 4 };
}

struct bcm_field {
  bcm_mesh* mesh = nullptr;
  int dtype = BCM_F64;
  void* data = nullptr;
  size_t count = 0;  // values
  size_t bytes = 0;
};

// A field of three interleaved components per cell (Vector3D's layout).
struct bcm_vfield {
  bcm_field base;  // count = 3 * nblocks * (n+2vc)^3
};

namespace bcm {

// Timing helpers (bcm_mesh.hip).
int timer_begin(bcm_mesh* m, int kernel, EventPair* ev);
int timer_end(bcm_mesh* m, EventPair* ev);

// Upload the boundary-value table (grows as needed).
int set_bvalues(bcm_mesh* m, const double* values, int n);

// Reduction of the per-work-group partials into d_result[slot] and,
// when `to_host` is set, copy back (synchronises the stream).
int reduce_partials(bcm_mesh* m, int nparts, int slot);
int reduce_partials_at(bcm_mesh* m, const double* parts, int nparts, int slot);
int fetch_result(bcm_mesh* m, int nslots, double* out);

// Deterministic sum of parts[0..n) by one 256-thread work-group: each
// thread sums its partials in order (the loads of eight issued together,
// then added one after the other), then a wave64 butterfly and a
// fixed-order sum of the four wave totals.  Thread 0 returns the total.
// red: 4 doubles of LDS.
__device__ __forceinline__ double sum_partials_wg256(const double* __restrict__ parts, int n, double* red) {
  constexpr int UNR = 8;
  double s = 0.0;
  int i = threadIdx.x;
  for (; i + (UNR - 1) * 256 < n; i += UNR * 256) {
    double v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) v[u] = parts[i + u * 256];
#pragma unroll
    for (int u = 0; u < UNR; ++u) s += v[u];
  }
  for (; i < n; i += 256) s += parts[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

// Device allocation helpers.
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);

// Raw-pointer kernels shared between the C ABI and the solvers.
template <typename T> int ghost_fill_raw(bcm_mesh* m, T* data, int mask);
template <typename T> int jacobi_poisson_raw(bcm_mesh* m, T* x1, const T* x0, const T* b, double omega);
template <typename T> int residual_poisson_raw(bcm_mesh* m, T* r, const T* x, const T* b, bool want_norm, int slot);
template <typename T> int spmv_poisson_raw(bcm_mesh* m, T* ax, const T* x);
template <typename T> int dot_raw(bcm_mesh* m, const T* x, const T* y, int slot, const T* mask = nullptr);
template <typename T> int blas_raw(bcm_mesh* m, int op, T* z, const T* x, const T* y, double a, double b,
                                   const T* mask = nullptr);
template <typename T> int fill_raw(bcm_mesh* m, T* x, double a);
int assemble_poisson_b_raw(bcm_mesh* m, double* b, const double* rhs);
int residual_d2f_raw(bcm_mesh* m, float* r32, const double* x, const double* b, int slot);
int convert_raw(bcm_mesh* m, int op, void* dst, const void* src);
// Dirichlet mirror with the boundary-value table already on the device.
int dirichlet_device(bcm_mesh* m, bcm_field* f);
// Halo-gathering red-black sweep A -> B (bcm_sweep.hip) and the X-face
// column export it reads.
template <typename T>
int halo_sweep_raw(bcm_mesh* m, const T* A, T* B, const T* s, double omega, const T* xin, T* xout, T* gout);
// Halo-gathering Jacobi of the matrix-free Poisson operator (bcm_sweep.hip):
// step 1 from the field x with its stored ghosts into compact buffers,
// then steps on compact iterates; the last one (gout) writes x's rows.
template <typename T>
int jacobi_first_raw(bcm_mesh* m, const T* x, const T* rhs, double omega, T* out, T* xout, T* rout);
template <typename T>
int jacobi_halo_raw(bcm_mesh* m, const T* A, T* B, const T* rhs, double omega, const T* xin, T* xout, T* x,
                    bool gout);
// Block-wise ghost fill (all faces, all kinds; n in {8, 16, 32}, vc = 1).
template <typename T> int ghost_gather_raw(bcm_mesh* m, T* data);
// Frees the fused solver's state (bcm_krylov.hip).
void krylov_release(bcm_mesh* m);
// Phases of a fused solve on the mesh stream (bcm_krylov.hip): begin
// (asynchronous set-up), advance (launch the next block of iterations once
// the previous one is done; wait = block until it is; *active = false when
// every segment has finished), finish (solution ghosts, outcome; syncs).
int krylov_check(const bcm_mesh* m, const bcm_solve_params* p, const double* seg_tol);
int krylov_begin(bcm_mesh* m, double* x, const double* rhs, const bcm_solve_params& p, const double* seg_tol);
int krylov_advance(bcm_mesh* m, bool wait, bool* active);
int krylov_finish(bcm_mesh* m, bcm_solve_stats* stats);
// Without waiting: *state = 0 while the last block of iterations runs, 1
// when it is done and another block is due, 2 when the solve is over.
int krylov_state(bcm_mesh* m, int* state);
// The current iterate of segment seg of a running fused solve (x plus the
// pending FP64 correction on the interior, the segment's connected faces'
// ghosts gathered) into y, the segment's ghosted blocks (device); waits
// for the iterations already queued.  *iters: iterations done.
int krylov_snapshot(bcm_mesh* m, int seg, double* y, int* iters);
// Restart files (bcm_io.hip).  dump2_raw: the Dump2 file of nblk ghosted
// blocks of the mesh's edge and ghost width held at dev (device memory,
// esize bytes per value).  load2_header: checks a Dump2 file against edge
// n (or n/2), ghost width vc and nblk blocks; *nfile = its edge.
// load2_into: blocks first.. of f from such a file.
int dump2_raw(bcm_mesh* m, const void* dev, int nblk, int esize, const char* dir, const char* label, int rank,
              int step);
int load2_header(const char* dir, const char* label, int rank, int step, int n, int vc, int nblk, int* nfile);
int load2_into(bcm_field* f, int first, int nblk, const char* dir, const char* label, int rank, int step);
// Device bytes the fused solver allocates for a mesh of nblocks n^3 blocks.
size_t krylov_bytes(int nblocks, int n, int precision, int count_pre);
// Field-sized workspace buffer kept on the mesh (bcm_solver.hip).
void* workspace(bcm_mesh* m, int slot, size_t bytes, int* rc);
// Compact iterates A, B and source of the halo-gathering SOR schedule.
enum WorkspaceSlot {
  WS_SOR_CA_F64 = 30, WS_SOR_CB_F64, WS_SOR_CS_F64, WS_SOR_CA_F32, WS_SOR_CB_F32, WS_SOR_CS_F32,
  WS_RES_PUB,   // bcm_jacobi_poisson_iterate: published interiors, two per block
  WS_RES_PING   // bcm_jacobi_poisson_iterate: the second ghosted buffer of the two-launch form
};

enum BlasOp { OP_AXPY = 0, OP_XPAY = 1, OP_AXPYZ = 2, OP_AXPBYPZ = 3, OP_COPY_INTERIOR = 4 };
enum ConvOp { CONV_D2F = 0, CONV_ADD_F2D = 1 };

}  // namespace bcm
