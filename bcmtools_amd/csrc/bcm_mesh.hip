// Device mesh, fields, ghost fill and boundary mirror.
//
// Ghost fill is one launch over a precomputed face-task table: each
// work-group takes one task (a face, or one subface for F2C) and writes
// that face's ghost layer from the neighbour's interior cells only
// (include/Scalar3DUpdater.h:190-266, 321-347, 519-708).  Because every
// task reads interiors and writes ghosts, the tasks are independent and
// the result does not depend on launch order.
#include <algorithm>
#include <cstring>
#include <new>

#include "bcm_host_par.h"
#include "bcm_internal.h"
#include "bcm_interp.h"

using bcm::BlockMeta;

namespace bcm {

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return BCM_ERR_NOMEM;
  }
  return BCM_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

static hipEvent_t get_event(bcm_mesh* m) {
  if (!m->event_pool.empty()) {
    hipEvent_t e = m->event_pool.back();
    m->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int timer_begin(bcm_mesh* m, int kernel, EventPair* ev) {
  ev->start = ev->stop = nullptr;
  ev->kernel = kernel;
  if (!m->timing) return BCM_OK;
  ev->start = get_event(m);
  ev->stop = get_event(m);
  if (!ev->start || !ev->stop) return BCM_ERR_HIP;
  BCM_HIP_TRY(hipEventRecord(ev->start, m->stream));
  return BCM_OK;
}

int timer_end(bcm_mesh* m, EventPair* ev) {
  if (!ev->start) return BCM_OK;
  BCM_HIP_TRY(hipEventRecord(ev->stop, m->stream));
  m->pending.push_back(*ev);
  // Bound the number of outstanding pairs.
  if (m->pending.size() > 4096) {
    BCM_HIP_TRY(hipStreamSynchronize(m->stream));
    for (auto& p : m->pending) {
      float ms = 0.f;
      BCM_HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
      m->kernel_ms[p.kernel] += ms;
      m->kernel_launches[p.kernel] += 1;
      m->event_pool.push_back(p.start);
      m->event_pool.push_back(p.stop);
    }
    m->pending.clear();
  }
  return BCM_OK;
}

int set_bvalues(bcm_mesh* m, const double* values, int n) {
  if (n < 0 || (n > 0 && !values)) return BCM_ERR_ARG;
  // Every Dirichlet or Neumann face must index a supplied value.
  if (n < m->need_bvalues) return BCM_ERR_ARG;
  if (m->bvalues_valid && int(m->h_bvalues.size()) == n &&
      std::equal(m->h_bvalues.begin(), m->h_bvalues.end(), values))
    return BCM_OK;
  if (n > m->n_bvalues_cap) {
    BCM_HIP_TRY(hipStreamSynchronize(m->stream));
    dev_free(m->d_bvalues);
    m->d_bvalues = nullptr;
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&m->d_bvalues), sizeof(double) * size_t(n)));
    m->n_bvalues_cap = n;
  }
  m->bvalues_valid = false;
  if (n > 0) {
    // Kernels already queued may still read the old table: wait, then copy.
    BCM_HIP_TRY(hipStreamSynchronize(m->stream));
    BCM_HIP_TRY(hipMemcpy(m->d_bvalues, values, sizeof(double) * size_t(n), hipMemcpyHostToDevice));
  }
  m->h_bvalues.assign(values, values + n);
  m->bvalues_valid = true;
  return BCM_OK;
}

// ---------------------------------------------------------------------
// Ghost fill
// ---------------------------------------------------------------------

struct Geo {
  int n, vc, nx0;
  long long bvol;
  __device__ __forceinline__ int idx(int i, int j, int k) const {
    return (i + vc) + nx0 * ((j + vc) + nx0 * (k + vc));
  }
};

// Component q of a cell of a field with NC interleaved components per
// cell (NC = 1: a scalar field; NC = 3: Vector3D's layout).
template <typename T, int NC = 1>
__device__ __forceinline__ T f2c_value(const T* __restrict__ f, const Geo& g, int I, int J, int K, int q = 0) {
  return f2c_sum<T>([&](int i, int j, int k) { return f[g.idx(i, j, k) * NC + q]; }, I, J, K);
}

template <typename T, int NC = 1>
__device__ __forceinline__ T c2f_value(const T* __restrict__ c, const Geo& g, int i, int j, int k, int q = 0) {
  return c2f_interp<T>([&](int a, int b, int d) { return c[g.idx(a, b, d) * NC + q]; }, g.n, i, j, k);
}

// One work-group per face task; every component of a ghost cell gets the
// scalar rule applied to that component.
template <typename T, int NC = 1>
__global__ __launch_bounds__(256) void k_ghost_fill(T* __restrict__ d, const int4* __restrict__ tasks, Geo g) {
  const int4 tk = tasks[blockIdx.x];
  const int face = tk.z & 7, sub = (tk.z >> 3) & 3, kind = (tk.z >> 5) & 3;
  const int axis = face >> 1;
  const bool plus = face & 1;
  T* __restrict__ dst = d + (long long)tk.x * g.bvol * NC;
  const T* __restrict__ src = d + (long long)tk.y * g.bvol * NC;
  const int n = g.n, vc = g.vc, h = n / 2;
  const int tn = (kind == KIND_F2C) ? h : n;  // tangential extent
  // Loop space (x fastest): depth vc along the face normal.
  const int ni = axis == 0 ? vc : tn;
  const int nj = axis == 1 ? vc : tn;
  const int nk = axis == 2 ? vc : tn;
  const int count = ni * nj * nk;
  const int so0 = sub & 1, so1 = sub >> 1;
  for (int e = threadIdx.x; e < count; e += blockDim.x) {
    const int i = e % ni;
    const int j = (e / ni) % nj;
    const int k = e / (ni * nj);
    int dc[3] = {i, j, k};
    int sc[3] = {i, j, k};
    const int dpt = dc[axis];
    if (kind == KIND_COPY) {
      dc[axis] = plus ? n + dpt : dpt - vc;
      sc[axis] = plus ? dpt : n - vc + dpt;
#pragma unroll
      for (int q = 0; q < NC; ++q)
        dst[g.idx(dc[0], dc[1], dc[2]) * NC + q] = src[g.idx(sc[0], sc[1], sc[2]) * NC + q];
    } else if (kind == KIND_F2C) {
      // Coarse ghost at the subface's quarter (copyFromNeighborF2C_0).
      const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;  // X:(y,z) Y:(z,x) Z:(x,y)
      dc[axis] = plus ? n + dpt : dpt - vc;
      dc[t0] += h * so0;
      dc[t1] += h * so1;
      sc[axis] = plus ? dpt : dpt + h - vc;
#pragma unroll
      for (int q = 0; q < NC; ++q)
        dst[g.idx(dc[0], dc[1], dc[2]) * NC + q] = f2c_value<T, NC>(src, g, sc[0], sc[1], sc[2], q);
    } else {
      // Fine ghost interpolated from the coarse neighbour (copyFromNeighborC2F_0).
      const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;
      dc[axis] = plus ? n + dpt : dpt - vc;
      sc[axis] = plus ? dpt : dpt + 2 * n - vc;
      sc[t0] += n * so0;
      sc[t1] += n * so1;
#pragma unroll
      for (int q = 0; q < NC; ++q)
        dst[g.idx(dc[0], dc[1], dc[2]) * NC + q] = c2f_value<T, NC>(src, g, sc[0], sc[1], sc[2], q);
    }
  }
}

template <typename T, int NC>
static int ghost_fill_nc(bcm_mesh* m, T* data, int mask) {
  const int kinds = mask & BCM_GF_ALL;
  int axes = (mask >> 3) & 7;
  if (axes == 0) axes = 7;
  if (kinds == 0) return BCM_OK;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_GHOST, &ev));
  auto launch = [&](int lo, int hi) -> int {
    if (hi <= lo) return BCM_OK;
    hipLaunchKernelGGL((k_ghost_fill<T, NC>), dim3(hi - lo), dim3(256), 0, m->stream, data, m->d_tasks + lo, g);
    BCM_HIP_TRY(hipGetLastError());
    return BCM_OK;
  };
  if (kinds == BCM_GF_ALL && axes == 7) {
    BCM_TRY(launch(0, m->ntask_total));
  } else {
    for (int a = 0; a < 3; ++a) {
      if (!(axes & (1 << a))) continue;
      if (kinds == BCM_GF_ALL) {
        BCM_TRY(launch(m->off[a][0], m->off[a][3]));
        continue;
      }
      for (int k = 0; k < 3; ++k)
        if (kinds & (1 << k)) BCM_TRY(launch(m->off[a][k], m->off[a][k + 1]));
    }
  }
  return timer_end(m, &ev);
}

// The complete fill (every kind, every axis) of a large mesh of blocks of
// 16^3 or fewer cells runs block-wise: one work-group gathers the six
// faces of a block, the same values as the face tasks.  Measured on
// 26,888 blocks of 16^3: 0.60 ms against 0.95 ms; with 32^3 blocks or a
// few hundred blocks the face tasks are the faster ones.  The gather
// knows a face as connected on every subface or not at all, so a table
// with a partly connected face keeps the face tasks.
static bool blockwise_pays(const bcm_mesh* m, int mask) {
  return (mask & 63) == BCM_GF_ALL && m->vc == 1 && m->n >= 4 && m->n <= 16 && m->nblocks >= 1024 && m->jhalo_ok;
}

template <typename T>
static int ghost_fill_blockwise(bcm_mesh* m, T* data) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_GHOST, &ev));
  BCM_TRY(ghost_gather_raw<T>(m, data));
  return timer_end(m, &ev);
}

template <typename T>
int ghost_fill_raw(bcm_mesh* m, T* data, int mask) {
  if (blockwise_pays(m, mask)) return ghost_fill_blockwise<T>(m, data);
  return ghost_fill_nc<T, 1>(m, data, mask);
}

template int ghost_fill_raw<double>(bcm_mesh*, double*, int);
template int ghost_fill_raw<float>(bcm_mesh*, float*, int);

// The face tasks whose destination is (block, face) -- and, for a finer
// neighbour, subface -- one work-group each (the per-face exchange of
// VCUpdater::copyFromNeighbor).
template <typename T, int NC>
static int ghost_fill_face_nc(bcm_mesh* m, T* data, int block, int face, int subface) {
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  int found = 0;
  for (int t = 0; t < int(m->h_tasks.size()); ++t) {
    const int4 tk = m->h_tasks[t];
    const int tf = tk.z & 7, ts = (tk.z >> 3) & 3, kind = (tk.z >> 5) & 3;
    if (tk.x != block || tf != face || (kind == KIND_F2C && ts != subface)) continue;
    hipLaunchKernelGGL((k_ghost_fill<T, NC>), dim3(1), dim3(256), 0, m->stream, data, m->d_tasks + t, g);
    BCM_HIP_TRY(hipGetLastError());
    ++found;
  }
  return found ? BCM_OK : BCM_ERR_ARG;
}

// ---------------------------------------------------------------------
// Dirichlet mirror: g = 2v - u0 on the first ghost layer
// (examples/sor1/Solver.cpp:232-294).
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_dirichlet(T* __restrict__ d, const int4* __restrict__ tasks,
                                                   const double* __restrict__ vals, Geo g) {
  const int4 tk = tasks[blockIdx.x];
  if (tk.w != BCM_BC_DIRICHLET) return;
  const int face = tk.y, axis = face >> 1;
  const bool plus = face & 1;
  T* __restrict__ f = d + (long long)tk.x * g.bvol;
  const double v = vals[tk.z];
  const int n = g.n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int u = e % n, w = e / n;
    int c[3], s[3];
    // X faces loop (j,k), Y faces (i,k), Z faces (i,j).
    if (axis == 0) { c[1] = u; c[2] = w; }
    else if (axis == 1) { c[0] = u; c[2] = w; }
    else { c[0] = u; c[1] = w; }
    s[0] = c[0]; s[1] = c[1]; s[2] = c[2];
    c[axis] = plus ? n : -1;
    s[axis] = plus ? n - 1 : 0;
    f[g.idx(c[0], c[1], c[2])] = T(2.0 * v - double(f[g.idx(s[0], s[1], s[2])]));
  }
}

template <typename T>
static int dirichlet_raw(bcm_mesh* m, T* data) {
  if (m->n_bc_tasks == 0) return BCM_OK;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_DIRICHLET, &ev));
  hipLaunchKernelGGL(k_dirichlet<T>, dim3(m->n_bc_tasks), dim3(256), 0, m->stream, data, m->d_bc_tasks,
                     m->d_bvalues, g);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

// ---------------------------------------------------------------------
// Outer-face conditions of BlockScalar3D::ImposeBoundaryCondition
// (examples/SolverTest/BlockScalar3D.h:199-206 with bcx.f90:18-376), per
// outer face and adjacent interior cell u:
//   Dirichlet  g1 = 2 v - u,   g2 = v             (bc_x*_d)
//   Neumann    g1 = u + xc,    g2 = u + 2 xc      on a plus face
//              g1 = u - xc,    g2 = u - 2 xc      on a minus face (bc_x*_n)
// with xc = v * cellSize.x (BlockScalar3D.h:98-100).  g2 is written only
// when the field has a second ghost layer (the Fortran writes it regardless).
// The arithmetic is FP64, rounded once to the field type.
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_impose_bc(T* __restrict__ d, const int4* __restrict__ tasks,
                                                   const double* __restrict__ vals,
                                                   const BlockMeta* __restrict__ meta, Geo g) {
  const int4 tk = tasks[blockIdx.x];
  const int face = tk.y, axis = face >> 1;
  const bool plus = face & 1;
  const bool neu = tk.w == BCM_BC_NEUMANN;
  T* __restrict__ f = d + (long long)tk.x * g.bvol;
  const double v = vals[tk.z];
  const double xc = neu ? v * meta[tk.x].hx : v;
  const int n = g.n;
  const bool two = g.vc >= 2;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int p = e % n, q = e / n;
    int c[3];
    // X faces loop (j,k), Y faces (i,k), Z faces (i,j).
    if (axis == 0) { c[1] = p; c[2] = q; }
    else if (axis == 1) { c[0] = p; c[2] = q; }
    else { c[0] = p; c[1] = q; }
    c[axis] = plus ? n - 1 : 0;
    const double u = double(f[g.idx(c[0], c[1], c[2])]);
    double g1, g2;
    if (neu) {
      g1 = plus ? u + xc : u - xc;
      g2 = plus ? u + xc * 2.0 : u - xc * 2.0;
    } else {
      g1 = 2.0 * xc - u;
      g2 = xc;
    }
    c[axis] = plus ? n : -1;
    f[g.idx(c[0], c[1], c[2])] = T(g1);
    if (two) {
      c[axis] = plus ? n + 1 : -2;
      f[g.idx(c[0], c[1], c[2])] = T(g2);
    }
  }
}

template <typename T>
static int impose_bc_raw(bcm_mesh* m, T* data) {
  if (m->n_bc_tasks == 0) return BCM_OK;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_DIRICHLET, &ev));
  hipLaunchKernelGGL(k_impose_bc<T>, dim3(m->n_bc_tasks), dim3(256), 0, m->stream, data, m->d_bc_tasks,
                     m->d_bvalues, m->d_meta, g);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

int dirichlet_device(bcm_mesh* m, bcm_field* f) {
  if (f->dtype == BCM_F64) return dirichlet_raw<double>(m, static_cast<double*>(f->data));
  return dirichlet_raw<float>(m, static_cast<float*>(f->data));
}

// ---------------------------------------------------------------------
// Reductions: deterministic two-stage sums.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ parts, int n,
                                                         double* __restrict__ out) {
  __shared__ double red[4];
  const double s = sum_partials_wg256(parts, n, red);
  if (threadIdx.x == 0) out[0] = s;
}

int reduce_partials(bcm_mesh* m, int nparts, int slot) { return reduce_partials_at(m, m->d_partials, nparts, slot); }

int reduce_partials_at(bcm_mesh* m, const double* parts, int nparts, int slot) {
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_REDUCE, &ev));
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, m->stream, parts, nparts, m->d_result + slot);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

int fetch_result(bcm_mesh* m, int nslots, double* out) {
  BCM_HIP_TRY(hipMemcpyAsync(m->h_result, m->d_result, sizeof(double) * size_t(nslots), hipMemcpyDeviceToHost,
                             m->stream));
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  for (int i = 0; i < nslots; ++i) out[i] = m->h_result[i];
  return BCM_OK;
}

}  // namespace bcm

using namespace bcm;

// ---------------------------------------------------------------------
// C ABI: mesh
// ---------------------------------------------------------------------

static void mesh_release(bcm_mesh* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  krylov_release(m);
  dev_free(m->d_seg_first);
  dev_free(m->d_blk_seg);
  for (auto& p : m->pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : m->event_pool) (void)hipEventDestroy(e);
  dev_free(m->d_tasks);
  dev_free(m->d_bc_tasks);
  dev_free(m->d_outer);
  dev_free(m->d_meta);
  dev_free(m->d_partials);
  dev_free(m->d_result);
  dev_free(m->d_bvalues);
  dev_free(m->d_sor_coef);
  dev_free(m->d_desc);
  dev_free(m->d_xexp[0]);
  dev_free(m->d_xexp[1]);
  for (void* p : m->ws) dev_free(p);
  for (bcm_field* f : m->fields) {
    dev_free(f->data);
    f->data = nullptr;
    f->mesh = nullptr;
  }
  if (m->h_result) (void)hipHostFree(m->h_result);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

static int mesh_create(const bcm_block_info* blocks, int nblocks, int n, int vc, const int* seg_counts, int nseg,
                       bcm_mesh_t* out) {
  if (!out) return BCM_ERR_ARG;
  *out = nullptr;
  if (!blocks || nblocks <= 0 || n < 2 || (n % 2) != 0 || vc < 1 || n / 2 < vc || n > 256) return BCM_ERR_ARG;
  // Validate the neighbour table (chunks of blocks on the host threads).
  const int nchunk = chunks_for(size_t(nblocks), 4096);
  std::atomic<bool> bad{false};
  parallel_chunks(size_t(nblocks), nchunk, [&](size_t b0, size_t b1, int) {
    for (size_t b = b0; b < b1 && !bad.load(std::memory_order_relaxed); ++b) {
      const bcm_block_info& bi = blocks[b];
      for (int f = 0; f < 6; ++f) {
        const int ld = bi.nb_level_diff[f];
        bool ok = ld >= -1 && ld <= 1;
        for (int s = 0; s < 4; ++s) ok = ok && bi.nb_id[f][s] >= -1 && bi.nb_id[f][s] < nblocks;
        ok = ok && !(ld == -1 && (bi.nb_subface[f] < 0 || bi.nb_subface[f] > 3));
        ok = ok && bi.bc_type[f] >= BCM_BC_INNER && bi.bc_type[f] <= BCM_BC_PERIODIC;
        ok = ok && !((bi.bc_type[f] == BCM_BC_DIRICHLET || bi.bc_type[f] == BCM_BC_NEUMANN) && bi.bc_id[f] < 0);
        if (!ok) bad.store(true);
      }
    }
  });
  if (bad.load()) return BCM_ERR_ARG;
  bcm_mesh* m = new (std::nothrow) bcm_mesh();
  if (!m) return BCM_ERR_NOMEM;
  m->n = n;
  m->vc = vc;
  m->nblocks = nblocks;
  m->nx0 = n + 2 * vc;
  m->bvol = size_t(m->nx0) * m->nx0 * m->nx0;
  m->ncell = size_t(nblocks) * n * n * n;
  int rc = BCM_OK;
  try {
    m->blocks.assign(blocks, blocks + nblocks);
    do {
      if (hipGetDevice(&m->device) != hipSuccess) { rc = BCM_ERR_HIP; break; }
      if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) { rc = BCM_ERR_HIP; break; }
      // Face tasks, as BlockManager::setFaceLists (src/BlockManager.cpp:471-530)
      // walks them: per block, per face, per subface.  Chunks of blocks build
      // their own lists on the host threads, merged in chunk order, so every
      // list is in block order whatever the chunk count.
      struct Part {
        std::vector<int4> lists[3][3];
        std::vector<int4> bc;
        int n_dirichlet = 0, n_neumann = 0, need_bvalues = 0;
        bool halo_ok = true, jhalo_ok = true;
      };
      std::vector<Part> parts(static_cast<size_t>(nchunk));
      m->meta.resize(nblocks);
      // Face descriptors of the halo-gathering sweep.
      std::vector<int> desc(size_t(nblocks) * 48, 0);
      parallel_chunks(size_t(nblocks), nchunk, [&](size_t b0, size_t b1, int c) {
        Part& P = parts[size_t(c)];
        for (size_t bb = b0; bb < b1; ++bb) {
          const int b = int(bb);
          const bcm_block_info& bi = blocks[b];
          BlockMeta& me = m->meta[b];
          const double hx = bi.block_size[0] / n, hy = bi.block_size[1] / n, hz = bi.block_size[2] / n;
          me.cx = 1.0 / (hx * hx);
          me.cy = 1.0 / (hy * hy);
          me.cz = 1.0 / (hz * hz);
          me.hx = hx;
          me.dir_mask = 0;
          me.neu_mask = 0;
          for (int f = 0; f < 6; ++f) {
            me.bc_id[f] = bi.bc_id[f];
            const int axis = f / 2;
            const int ld = bi.nb_level_diff[f];
            if (ld == 0 || ld == -1) {
              const int nb = bi.nb_id[f][0];
              if (nb >= 0) {
                const int kind = ld == 0 ? KIND_COPY : KIND_C2F;
                const int sub = ld == 0 ? 0 : bi.nb_subface[f];
                P.lists[axis][kind].push_back(make_int4(b, nb, f | (sub << 3) | (kind << 5), 0));
              }
            } else {
              for (int q = 0; q < 4; ++q) {
                const int nb = bi.nb_id[f][q];
                if (nb >= 0) P.lists[axis][KIND_F2C].push_back(make_int4(b, nb, f | (q << 3) | (KIND_F2C << 5), 0));
              }
            }
            if (bi.bc_type[f] == BCM_BC_DIRICHLET || bi.bc_type[f] == BCM_BC_NEUMANN) {
              P.bc.push_back(make_int4(b, f, bi.bc_id[f], bi.bc_type[f]));
              if (bi.bc_type[f] == BCM_BC_DIRICHLET) {
                me.dir_mask |= 1 << f;
                P.n_dirichlet++;
              } else {
                me.neu_mask |= 1 << f;
                P.n_neumann++;
              }
              P.need_bvalues = std::max(P.need_bvalues, bi.bc_id[f] + 1);
            }
            int* d = &desc[size_t(b) * 48 + f * 8];
            for (int q = 0; q < 4; ++q) d[1 + q] = bi.nb_id[f][q];
            d[5] = bi.nb_subface[f];
            d[6] = bi.bc_id[f];
            if (ld == 0 && bi.nb_id[f][0] >= 0) d[0] = FD_COPY;
            else if (ld == -1 && bi.nb_id[f][0] >= 0) d[0] = FD_C2F;
            else if (ld == 1 && bi.nb_id[f][0] >= 0 && bi.nb_id[f][1] >= 0 && bi.nb_id[f][2] >= 0 &&
                     bi.nb_id[f][3] >= 0) d[0] = FD_F2C;
            else if (bi.bc_type[f] == BCM_BC_DIRICHLET) d[0] = FD_DIR;
            else {
              d[0] = FD_NONE;
              P.halo_ok = false;
            }
            if (d[0] == FD_DIR || d[0] == FD_NONE)
              for (int q = 0; q < 4; ++q)
                if (bi.nb_id[f][q] >= 0) P.jhalo_ok = false;
          }
        }
      });
      m->halo_ok = true;
      m->jhalo_ok = true;
      std::vector<int4> bc;
      for (const Part& P : parts) {
        bc.insert(bc.end(), P.bc.begin(), P.bc.end());
        m->n_dirichlet += P.n_dirichlet;
        m->n_neumann += P.n_neumann;
        m->need_bvalues = std::max(m->need_bvalues, P.need_bvalues);
        m->halo_ok = m->halo_ok && P.halo_ok;
        m->jhalo_ok = m->jhalo_ok && P.jhalo_ok;
      }
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_desc), sizeof(int) * desc.size())) != BCM_OK) break;
      if (hipMemcpy(m->d_desc, desc.data(), sizeof(int) * desc.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      std::vector<int4>& all = m->h_tasks;
      all.clear();
      for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < 3; ++k) {
          m->off[a][k] = int(all.size());
          for (const Part& P : parts) {
            all.insert(all.end(), P.lists[a][k].begin(), P.lists[a][k].end());
            m->kind_count[k] += int(P.lists[a][k].size());
          }
        }
        m->off[a][3] = int(all.size());
      }
      m->ntask_total = int(all.size());
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_tasks), sizeof(int4) * all.size())) != BCM_OK) break;
      if (!all.empty() &&
          hipMemcpy(m->d_tasks, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      m->n_bc_tasks = int(bc.size());
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_bc_tasks), sizeof(int4) * bc.size())) != BCM_OK) break;
      if (!bc.empty() &&
          hipMemcpy(m->d_bc_tasks, bc.data(), sizeof(int4) * bc.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_meta), sizeof(BlockMeta) * size_t(nblocks))) != BCM_OK)
        break;
      if (hipMemcpy(m->d_meta, m->meta.data(), sizeof(BlockMeta) * size_t(nblocks), hipMemcpyHostToDevice) !=
          hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
      // One partial per 256-cell work-group of a cell sweep.
      m->n_partials = (m->ncell + 255) / 256 + 1024;
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_partials), sizeof(double) * m->n_partials)) != BCM_OK)
        break;
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_result), sizeof(double) * 16)) != BCM_OK) break;
      if (hipHostMalloc(reinterpret_cast<void**>(&m->h_result), sizeof(double) * 16, hipHostMallocDefault) !=
          hipSuccess) {
        m->h_result = nullptr;
        rc = BCM_ERR_NOMEM;
        break;
      }
      // Segment tables.
      m->nseg = nseg;
      m->seg_first.assign(size_t(nseg) + 1, 0);
      std::vector<int> blk_seg(size_t(nblocks), 0);
      for (int s = 0, b = 0; s < nseg; ++s) {
        m->seg_first[s] = b;
        for (int e = 0; e < seg_counts[s]; ++e) blk_seg[size_t(b++)] = s;
        m->seg_first[s + 1] = b;
      }
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_seg_first), sizeof(int) * (size_t(nseg) + 1))) != BCM_OK)
        break;
      if ((rc = dev_alloc(reinterpret_cast<void**>(&m->d_blk_seg), sizeof(int) * size_t(nblocks))) != BCM_OK) break;
      if (hipMemcpy(m->d_seg_first, m->seg_first.data(), sizeof(int) * (size_t(nseg) + 1), hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(m->d_blk_seg, blk_seg.data(), sizeof(int) * size_t(nblocks), hipMemcpyHostToDevice) != hipSuccess) {
        rc = BCM_ERR_HIP;
        break;
      }
    } while (false);
  } catch (const std::bad_alloc&) {  // host tables (also from the threads that build them)
    rc = BCM_ERR_NOMEM;
  }
  if (rc != BCM_OK) {
    mesh_release(m);
    return rc;
  }
  *out = m;
  return BCM_OK;
}

extern "C" int bcm_mesh_create(const bcm_block_info* blocks, int nblocks, int n, int vc, bcm_mesh_t* out) {
  const int one[1] = {nblocks};
  return mesh_create(blocks, nblocks, n, vc, one, 1, out);
}

extern "C" int bcm_mesh_create_batch(const bcm_block_info* blocks, const int* counts, int nseg, int n, int vc,
                                     bcm_mesh_t* out) {
  if (!out) return BCM_ERR_ARG;
  *out = nullptr;
  if (!blocks || !counts || nseg <= 0) return BCM_ERR_ARG;
  long long total = 0;
  for (int s = 0; s < nseg; ++s) {
    if (counts[s] <= 0) return BCM_ERR_ARG;
    total += counts[s];
  }
  if (total > (1LL << 30)) return BCM_ERR_ARG;
  // Shift each segment's neighbour ids (local to the segment) into the
  // concatenated table.
  std::vector<bcm_block_info> all;
  try {
    all.assign(blocks, blocks + total);
  } catch (const std::bad_alloc&) {
    return BCM_ERR_NOMEM;
  }
  for (int s = 0, b0 = 0; s < nseg; b0 += counts[s], ++s) {
    for (int e = 0; e < counts[s]; ++e) {
      bcm_block_info& bi = all[size_t(b0 + e)];
      for (int f = 0; f < 6; ++f)
        for (int q = 0; q < 4; ++q) {
          const int nb = bi.nb_id[f][q];
          if (nb < -1 || nb >= counts[s]) return BCM_ERR_ARG;
          if (nb >= 0) bi.nb_id[f][q] = nb + b0;
        }
    }
  }
  return mesh_create(all.data(), int(total), n, vc, counts, nseg, out);
}

extern "C" int bcm_mesh_num_segments(bcm_mesh_t m) { return m ? m->nseg : BCM_ERR_ARG; }

extern "C" int bcm_mesh_segment(bcm_mesh_t m, int seg, int* first_block, int* nblocks) {
  if (!m || seg < 0 || seg >= m->nseg || !first_block || !nblocks) return BCM_ERR_ARG;
  *first_block = m->seg_first[size_t(seg)];
  *nblocks = m->seg_first[size_t(seg) + 1] - m->seg_first[size_t(seg)];
  return BCM_OK;
}

extern "C" int bcm_mesh_destroy(bcm_mesh_t mesh) {
  mesh_release(mesh);
  return BCM_OK;
}

extern "C" int bcm_mesh_task_counts(bcm_mesh_t m, int counts[4]) {
  if (!m || !counts) return BCM_ERR_ARG;
  counts[0] = m->kind_count[KIND_COPY];
  counts[1] = m->kind_count[KIND_F2C];
  counts[2] = m->kind_count[KIND_C2F];
  counts[3] = m->n_bc_tasks;
  return BCM_OK;
}

extern "C" int bcm_mesh_sync(bcm_mesh_t m) {
  if (!m) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  return BCM_OK;
}

extern "C" int bcm_mesh_set_timing(bcm_mesh_t m, int enable) {
  if (!m) return BCM_ERR_ARG;
  m->timing = enable != 0;
  return BCM_OK;
}

extern "C" int bcm_mesh_kernel_time(bcm_mesh_t m, int kernel, double* total_ms, long long* launches) {
  if (!m || kernel < 0 || kernel >= BCM_K_COUNT) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  for (auto& p : m->pending) {
    float ms = 0.f;
    BCM_HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
    m->kernel_ms[p.kernel] += ms;
    m->kernel_launches[p.kernel] += 1;
    m->event_pool.push_back(p.start);
    m->event_pool.push_back(p.stop);
  }
  m->pending.clear();
  if (total_ms) *total_ms = m->kernel_ms[kernel];
  if (launches) *launches = m->kernel_launches[kernel];
  return BCM_OK;
}

extern "C" int bcm_mesh_reset_timing(bcm_mesh_t m) {
  if (!m) return BCM_ERR_ARG;
  double dummy;
  BCM_TRY(bcm_mesh_kernel_time(m, 0, &dummy, nullptr));
  for (int k = 0; k < BCM_K_COUNT; ++k) {
    m->kernel_ms[k] = 0.0;
    m->kernel_launches[k] = 0;
  }
  return BCM_OK;
}

// ---------------------------------------------------------------------
// C ABI: fields
// ---------------------------------------------------------------------

extern "C" int bcm_field_create(bcm_mesh_t m, int dtype, bcm_field_t* out) {
  if (!m || !out || (dtype != BCM_F64 && dtype != BCM_F32)) return BCM_ERR_ARG;
  *out = nullptr;
  bcm_field* f = new (std::nothrow) bcm_field();
  if (!f) return BCM_ERR_NOMEM;
  f->mesh = m;
  f->dtype = dtype;
  f->count = size_t(m->nblocks) * m->bvol;
  f->bytes = f->count * (dtype == BCM_F64 ? 8 : 4);
  int rc = dev_alloc(&f->data, f->bytes);
  if (rc != BCM_OK) {
    delete f;
    return rc;
  }
  if (hipMemsetAsync(f->data, 0, f->bytes, m->stream) != hipSuccess) {
    dev_free(f->data);
    delete f;
    return BCM_ERR_HIP;
  }
  m->fields.push_back(f);
  *out = f;
  return BCM_OK;
}

extern "C" int bcm_field_destroy(bcm_field_t f) {
  if (!f) return BCM_OK;
  if (bcm_mesh* m = f->mesh) {
    (void)hipStreamSynchronize(m->stream);
    m->fields.erase(std::remove(m->fields.begin(), m->fields.end(), f), m->fields.end());
  }
  dev_free(f->data);
  delete f;
  return BCM_OK;
}

extern "C" int bcm_field_upload(bcm_field_t f, const void* host, size_t bytes) {
  if (!f || !host || bytes != f->bytes) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipMemcpyAsync(f->data, host, bytes, hipMemcpyHostToDevice, f->mesh->stream));
  BCM_HIP_TRY(hipStreamSynchronize(f->mesh->stream));
  return BCM_OK;
}

extern "C" int bcm_host_alloc(size_t bytes, void** out) {
  if (!out) return BCM_ERR_ARG;
  *out = nullptr;
  if (bytes == 0) return BCM_ERR_ARG;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) return BCM_ERR_NOMEM;
  std::memset(p, 0, bytes);
  *out = p;
  return BCM_OK;
}

extern "C" int bcm_host_free(void* p) {
  if (!p) return BCM_OK;
  BCM_HIP_TRY(hipHostFree(p));
  return BCM_OK;
}

extern "C" int bcm_field_download(bcm_field_t f, void* host, size_t bytes) {
  if (!f || !host || bytes != f->bytes) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipMemcpyAsync(host, f->data, bytes, hipMemcpyDeviceToHost, f->mesh->stream));
  BCM_HIP_TRY(hipStreamSynchronize(f->mesh->stream));
  return BCM_OK;
}

// Blocks first_block .. first_block+nblocks-1 (e.g. one segment of a batch).
// Byte range of blocks first_block .. first_block+nblocks-1 (a block holds
// count / nblocks values: (n+2vc)^3 per component).
static int block_range(bcm_field_t f, int first_block, int nblocks, size_t bytes, size_t* off, size_t* len) {
  if (!f || !f->mesh || first_block < 0 || nblocks < 0 || first_block + nblocks > f->mesh->nblocks)
    return BCM_ERR_ARG;
  const size_t es = f->dtype == BCM_F64 ? 8 : 4;
  const size_t per = f->count / size_t(f->mesh->nblocks);
  *off = size_t(first_block) * per * es;
  *len = size_t(nblocks) * per * es;
  return bytes == *len ? BCM_OK : BCM_ERR_ARG;
}

extern "C" int bcm_field_upload_blocks(bcm_field_t f, int first_block, int nblocks, const void* host, size_t bytes) {
  size_t off = 0, len = 0;
  BCM_TRY(block_range(f, first_block, nblocks, bytes, &off, &len));
  if (!host && len) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipMemcpyAsync(static_cast<char*>(f->data) + off, host, len, hipMemcpyHostToDevice, f->mesh->stream));
  BCM_HIP_TRY(hipStreamSynchronize(f->mesh->stream));
  return BCM_OK;
}

extern "C" int bcm_field_download_blocks(bcm_field_t f, int first_block, int nblocks, void* host, size_t bytes) {
  size_t off = 0, len = 0;
  BCM_TRY(block_range(f, first_block, nblocks, bytes, &off, &len));
  if (!host && len) return BCM_ERR_ARG;
  BCM_HIP_TRY(hipMemcpyAsync(host, static_cast<const char*>(f->data) + off, len, hipMemcpyDeviceToHost, f->mesh->stream));
  BCM_HIP_TRY(hipStreamSynchronize(f->mesh->stream));
  return BCM_OK;
}

extern "C" void* bcm_field_device_ptr(bcm_field_t f) { return f ? f->data : nullptr; }

extern "C" size_t bcm_field_bytes(bcm_field_t f) { return f ? f->bytes : 0; }

extern "C" int bcm_ghost_fill(bcm_field_t f, int mask) {
  if (!f || (mask & ~127)) return BCM_ERR_ARG;
  if (mask & BCM_GF_BLOCKWISE) {
    // All faces of every block in one pass; only the complete fill.
    if ((mask & 63) != BCM_GF_ALL) return BCM_ERR_ARG;
    if (f->dtype == BCM_F64) return ghost_fill_blockwise<double>(f->mesh, static_cast<double*>(f->data));
    return ghost_fill_blockwise<float>(f->mesh, static_cast<float*>(f->data));
  }
  if (f->dtype == BCM_F64) return ghost_fill_raw<double>(f->mesh, static_cast<double*>(f->data), mask);
  return ghost_fill_raw<float>(f->mesh, static_cast<float*>(f->data), mask);
}

extern "C" int bcm_ghost_fill_face(bcm_field_t f, int block, int face, int subface) {
  if (!f || !f->mesh || block < 0 || block >= f->mesh->nblocks || face < 0 || face >= 6 || subface < 0 ||
      subface >= 4)
    return BCM_ERR_ARG;
  if (f->dtype == BCM_F64)
    return ghost_fill_face_nc<double, 1>(f->mesh, static_cast<double*>(f->data), block, face, subface);
  return ghost_fill_face_nc<float, 1>(f->mesh, static_cast<float*>(f->data), block, face, subface);
}

// ---------------------------------------------------------------------
// Face buffers (Scalar3DUpdater::copyToCommBuffer* / copyFromCommBuffer*,
// include/Scalar3DUpdater.h:377-515, 711-901): one face per work-group,
// its box enumerated x fastest (element m = i + w (j + h k)).
// ---------------------------------------------------------------------
namespace bcm {
namespace {

template <typename T, bool PACK>
__global__ __launch_bounds__(256) void k_comm(T* __restrict__ d, T* __restrict__ buf,
                                              const bcm_comm_face* __restrict__ faces, Geo g) {
  const bcm_comm_face fc = faces[blockIdx.x];
  const int axis = fc.face >> 1;
  const bool plus = fc.face & 1;
  const int n = g.n, vc = g.vc, h = n / 2;
  const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;  // X:(y,z) Y:(z,x) Z:(x,y)
  const int so0 = fc.subface & 1, so1 = fc.subface >> 1;
  const int tn = fc.kind == BCM_COMM_F2C ? h : n;
  int ext[3];
  ext[axis] = vc;
  ext[t0] = tn;
  ext[t1] = tn;
  const int count = ext[0] * ext[1] * ext[2];
  T* __restrict__ blk = d + (long long)fc.block * g.bvol;
  T* __restrict__ fb = buf + fc.offset;
  for (int m = threadIdx.x; m < count; m += 256) {
    int c[3] = {m % ext[0], (m / ext[0]) % ext[1], m / (ext[0] * ext[1])};
    if (PACK) {
      T v;
      if (fc.kind == BCM_COMM_SAME) {
        c[axis] += plus ? n - vc : 0;
        v = blk[g.idx(c[0], c[1], c[2])];
      } else if (fc.kind == BCM_COMM_F2C) {
        // copyToCommBufferF2C_0: coarse-cell coordinates of the fine block.
        c[axis] += plus ? h - vc : 0;
        v = f2c_value<T>(blk, g, c[0], c[1], c[2]);
      } else {
        // copyToCommBufferC2F_0: fine-cell coordinates of the coarse block.
        c[axis] += plus ? 2 * n - vc : 0;
        c[t0] += n * so0;
        c[t1] += n * so1;
        v = c2f_value<T>(blk, g, c[0], c[1], c[2]);
      }
      fb[m] = v;
    } else {
      c[axis] += plus ? n : -vc;
      if (fc.kind == BCM_COMM_F2C) {
        c[t0] += h * so0;
        c[t1] += h * so1;
      }
      blk[g.idx(c[0], c[1], c[2])] = fb[m];
    }
  }
}

// Checks a face list against the mesh; *need: the elements the buffer must
// hold (the largest offset + face size).
int comm_check(const bcm_mesh* m, const bcm_comm_face* faces, int nfaces, long long* need) {
  const long long full = (long long)m->n * m->n * m->vc;
  long long hi = 0;
  for (int i = 0; i < nfaces; ++i) {
    const bcm_comm_face& c = faces[i];
    if (c.block < 0 || c.block >= m->nblocks || c.face < 0 || c.face >= BCM_NUM_FACE || c.subface < 0 ||
        c.subface > 3 || c.kind < BCM_COMM_SAME || c.kind > BCM_COMM_C2F || c.offset < 0 ||
        c.offset > (1ll << 60))
      return BCM_ERR_ARG;
    hi = std::max(hi, (long long)c.offset + (c.kind == BCM_COMM_F2C ? full / 4 : full));
  }
  *need = hi;
  return BCM_OK;
}

template <typename T, bool PACK>
hipError_t comm_launch(bcm_field* f, const bcm_comm_face* d_faces, int nfaces, void* buffer) {
  bcm_mesh* m = f->mesh;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  hipLaunchKernelGGL((k_comm<T, PACK>), dim3(unsigned(nfaces)), dim3(256), 0, m->stream, static_cast<T*>(f->data),
                     static_cast<T*>(buffer), d_faces, g);
  return hipGetLastError();
}

template <typename T, bool PACK>
int comm_run(bcm_field* f, const bcm_comm_face* faces, int nfaces, void* buffer, size_t bytes) {
  bcm_mesh* m = f->mesh;
  long long need = 0;
  BCM_TRY(comm_check(m, faces, nfaces, &need));
  if (need > (long long)(bytes / sizeof(T))) return BCM_ERR_ARG;
  if (nfaces == 0) return BCM_OK;
  void* df = nullptr;
  BCM_TRY(dev_alloc(&df, sizeof(bcm_comm_face) * size_t(nfaces)));
  hipError_t e = hipMemcpyAsync(df, faces, sizeof(bcm_comm_face) * size_t(nfaces), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = comm_launch<T, PACK>(f, static_cast<const bcm_comm_face*>(df), nfaces, buffer);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  dev_free(df);
  return e == hipSuccess ? BCM_OK : BCM_ERR_HIP;
}

}  // namespace
}  // namespace bcm

// A face list uploaded once (what prepareForVCUpdate keeps between
// exchanges): packs and unpacks over it allocate nothing, copy nothing
// from the host and do not wait for the stream.
struct bcm_comm_list {
  bcm_mesh* mesh = nullptr;
  bcm_comm_face* d_faces = nullptr;
  int nfaces = 0;
  long long need = 0;  // elements a buffer must hold
};

extern "C" int bcm_comm_list_create(bcm_mesh_t m, const bcm_comm_face* faces, int nfaces, bcm_comm_list_t* out) {
  if (!out) return BCM_ERR_ARG;
  *out = nullptr;
  if (!m || nfaces < 0 || (nfaces > 0 && !faces)) return BCM_ERR_ARG;
  long long need = 0;
  BCM_TRY(bcm::comm_check(m, faces, nfaces, &need));
  bcm_comm_list* l = new (std::nothrow) bcm_comm_list;
  if (!l) return BCM_ERR_NOMEM;
  l->mesh = m;
  l->nfaces = nfaces;
  l->need = need;
  if (nfaces > 0) {
    const size_t bytes = sizeof(bcm_comm_face) * size_t(nfaces);
    int rc = bcm::dev_alloc(reinterpret_cast<void**>(&l->d_faces), bytes);
    if (rc == BCM_OK && (hipMemcpyAsync(l->d_faces, faces, bytes, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
                         hipStreamSynchronize(m->stream) != hipSuccess))
      rc = BCM_ERR_HIP;
    if (rc != BCM_OK) {
      bcm::dev_free(l->d_faces);
      delete l;
      return rc;
    }
  }
  *out = l;
  return BCM_OK;
}

extern "C" int bcm_comm_list_destroy(bcm_comm_list_t l) {
  if (!l) return BCM_OK;
  if (l->d_faces) (void)hipStreamSynchronize(l->mesh->stream);  // no launch may still read the list
  bcm::dev_free(l->d_faces);
  delete l;
  return BCM_OK;
}

template <bool PACK>
static int comm_list_run(bcm_field_t f, bcm_comm_list_t l, void* buffer, size_t bytes) {
  if (!f || !f->mesh || !l || l->mesh != f->mesh || (l->nfaces > 0 && !buffer)) return BCM_ERR_ARG;
  const size_t es = f->dtype == BCM_F64 ? sizeof(double) : sizeof(float);
  if (l->need > (long long)(bytes / es)) return BCM_ERR_ARG;
  if (l->nfaces == 0) return BCM_OK;
  const hipError_t e = f->dtype == BCM_F64 ? bcm::comm_launch<double, PACK>(f, l->d_faces, l->nfaces, buffer)
                                           : bcm::comm_launch<float, PACK>(f, l->d_faces, l->nfaces, buffer);
  return e == hipSuccess ? BCM_OK : BCM_ERR_HIP;
}

extern "C" int bcm_comm_pack_list(bcm_field_t f, bcm_comm_list_t l, void* buffer, size_t bytes) {
  return comm_list_run<true>(f, l, buffer, bytes);
}

extern "C" int bcm_comm_unpack_list(bcm_field_t f, bcm_comm_list_t l, const void* buffer, size_t bytes) {
  return comm_list_run<false>(f, l, const_cast<void*>(buffer), bytes);
}

extern "C" int64_t bcm_comm_face_size(bcm_field_t f, int kind) {
  if (!f || !f->mesh || kind < BCM_COMM_SAME || kind > BCM_COMM_C2F) return -BCM_ERR_ARG;
  const int64_t full = int64_t(f->mesh->n) * f->mesh->n * f->mesh->vc;
  return kind == BCM_COMM_F2C ? full / 4 : full;
}

extern "C" int bcm_comm_pack(bcm_field_t f, const bcm_comm_face* faces, int nfaces, void* buffer, size_t bytes) {
  if (!f || !f->mesh || nfaces < 0 || (nfaces > 0 && (!faces || !buffer))) return BCM_ERR_ARG;
  if (f->dtype == BCM_F64) return comm_run<double, true>(f, faces, nfaces, buffer, bytes);
  return comm_run<float, true>(f, faces, nfaces, buffer, bytes);
}

extern "C" int bcm_comm_unpack(bcm_field_t f, const bcm_comm_face* faces, int nfaces, const void* buffer,
                               size_t bytes) {
  if (!f || !f->mesh || nfaces < 0 || (nfaces > 0 && (!faces || !buffer))) return BCM_ERR_ARG;
  if (f->dtype == BCM_F64) return comm_run<double, false>(f, faces, nfaces, const_cast<void*>(buffer), bytes);
  return comm_run<float, false>(f, faces, nfaces, const_cast<void*>(buffer), bytes);
}

static int comm_host(bcm_field_t f, const bcm_comm_face* faces, int nfaces, void* host, size_t bytes, bool pack) {
  if (!f || !f->mesh || nfaces < 0 || (nfaces > 0 && (!faces || !host))) return BCM_ERR_ARG;
  if (nfaces == 0) return BCM_OK;
  if (bytes == 0) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  void* d = nullptr;
  BCM_TRY(dev_alloc(&d, bytes));
  int rc = BCM_OK;
  if (!pack && hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, m->stream) != hipSuccess) rc = BCM_ERR_HIP;
  if (rc == BCM_OK) rc = pack ? bcm_comm_pack(f, faces, nfaces, d, bytes) : bcm_comm_unpack(f, faces, nfaces, d, bytes);
  if (rc == BCM_OK && pack) {
    if (hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      rc = BCM_ERR_HIP;
  }
  (void)hipStreamSynchronize(m->stream);
  dev_free(d);
  return rc;
}

extern "C" int bcm_comm_pack_host(bcm_field_t f, const bcm_comm_face* faces, int nfaces, void* buffer, size_t bytes) {
  return comm_host(f, faces, nfaces, buffer, bytes, true);
}

extern "C" int bcm_comm_unpack_host(bcm_field_t f, const bcm_comm_face* faces, int nfaces, const void* buffer,
                                    size_t bytes) {
  return comm_host(f, faces, nfaces, const_cast<void*>(buffer), bytes, false);
}

// ---------------------------------------------------------------------
// Vector fields: three interleaved components per cell.
// ---------------------------------------------------------------------
extern "C" int bcm_vfield_create(bcm_mesh_t m, int dtype, bcm_vfield_t* out) {
  if (!m || !out || (dtype != BCM_F64 && dtype != BCM_F32)) return BCM_ERR_ARG;
  *out = nullptr;
  bcm_vfield* v = new (std::nothrow) bcm_vfield();
  if (!v) return BCM_ERR_NOMEM;
  bcm_field& f = v->base;
  f.mesh = m;
  f.dtype = dtype;
  f.count = 3 * size_t(m->nblocks) * m->bvol;
  f.bytes = f.count * (dtype == BCM_F64 ? 8 : 4);
  int rc = dev_alloc(&f.data, f.bytes);
  if (rc == BCM_OK && hipMemsetAsync(f.data, 0, f.bytes, m->stream) != hipSuccess) rc = BCM_ERR_HIP;
  if (rc != BCM_OK) {
    dev_free(f.data);
    delete v;
    return rc;
  }
  m->fields.push_back(&f);
  *out = v;
  return BCM_OK;
}

extern "C" int bcm_vfield_destroy(bcm_vfield_t v) {
  if (!v) return BCM_OK;
  bcm_field* f = &v->base;
  if (bcm_mesh* m = f->mesh) {
    (void)hipStreamSynchronize(m->stream);
    m->fields.erase(std::remove(m->fields.begin(), m->fields.end(), f), m->fields.end());
  }
  dev_free(f->data);
  delete v;
  return BCM_OK;
}

extern "C" int bcm_vfield_upload(bcm_vfield_t v, const void* host, size_t bytes) {
  if (!v || !v->base.mesh) return BCM_ERR_ARG;
  return bcm_field_upload(&v->base, host, bytes);
}

extern "C" int bcm_vfield_download(bcm_vfield_t v, void* host, size_t bytes) {
  if (!v || !v->base.mesh) return BCM_ERR_ARG;
  return bcm_field_download(&v->base, host, bytes);
}

extern "C" int bcm_vfield_upload_blocks(bcm_vfield_t v, int first_block, int nblocks, const void* host,
                                        size_t bytes) {
  if (!v) return BCM_ERR_ARG;
  return bcm_field_upload_blocks(&v->base, first_block, nblocks, host, bytes);
}

extern "C" int bcm_vfield_download_blocks(bcm_vfield_t v, int first_block, int nblocks, void* host, size_t bytes) {
  if (!v) return BCM_ERR_ARG;
  return bcm_field_download_blocks(&v->base, first_block, nblocks, host, bytes);
}

extern "C" int bcm_vfield_ghost_fill(bcm_vfield_t v, int mask) {
  if (!v || !v->base.mesh || (mask & ~63)) return BCM_ERR_ARG;
  bcm_field& f = v->base;
  if (f.dtype == BCM_F64) return ghost_fill_nc<double, 3>(f.mesh, static_cast<double*>(f.data), mask);
  return ghost_fill_nc<float, 3>(f.mesh, static_cast<float*>(f.data), mask);
}

extern "C" int bcm_vfield_ghost_fill_face(bcm_vfield_t v, int block, int face, int subface) {
  if (!v || !v->base.mesh) return BCM_ERR_ARG;
  bcm_field& f = v->base;
  if (block < 0 || block >= f.mesh->nblocks || face < 0 || face >= 6 || subface < 0 || subface >= 4)
    return BCM_ERR_ARG;
  if (f.dtype == BCM_F64)
    return ghost_fill_face_nc<double, 3>(f.mesh, static_cast<double*>(f.data), block, face, subface);
  return ghost_fill_face_nc<float, 3>(f.mesh, static_cast<float*>(f.data), block, face, subface);
}

extern "C" int bcm_impose_bc(bcm_field_t f, const double* values, int nvalues) {
  if (!f) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  BCM_TRY(set_bvalues(m, values, nvalues));
  if (f->dtype == BCM_F64) return impose_bc_raw<double>(m, static_cast<double*>(f->data));
  return impose_bc_raw<float>(m, static_cast<float*>(f->data));
}

extern "C" int bcm_update_virtual_cells(bcm_field_t f, int method) {
  if (!f) return BCM_ERR_ARG;
  switch (method) {
    case BCM_VC_AT_ONCE:
      return bcm_ghost_fill(f, BCM_GF_ALL);
    case BCM_VC_SEPARATE_XYZ:
      for (int ax : {BCM_GF_X, BCM_GF_Y, BCM_GF_Z}) BCM_TRY(bcm_ghost_fill(f, BCM_GF_ALL | ax));
      return BCM_OK;
    case BCM_VC_SEPARATE_XYZ_SEPARATE_LEVEL_DIFF:
      for (int kind : {BCM_GF_F2C, BCM_GF_FLAT, BCM_GF_C2F})
        for (int ax : {BCM_GF_X, BCM_GF_Y, BCM_GF_Z}) BCM_TRY(bcm_ghost_fill(f, kind | ax));
      return BCM_OK;
    default:
      return BCM_ERR_ARG;
  }
}

extern "C" int bcm_impose_boundary_condition(bcm_field_t f, int method, const double* values, int nvalues) {
  BCM_TRY(bcm_update_virtual_cells(f, method));
  return bcm_impose_bc(f, values, nvalues);
}

// ---------------------------------------------------------------------
// Per-field outer-face conditions: SolverTest's LocalScalar3D keeps one
// boundary type and value per domain face for each field
// (examples/SolverTest/LocalScalar3D.h:36-100, BlockScalar3D.h:76-198) and
// applies them to every block face without a neighbour, faces in X_M..Z_P
// order within a block (BlockScalar3D.h:199-214).  One work-group per block
// that has such a face, a barrier between its faces.
// ---------------------------------------------------------------------
struct FaceBC {
  int type[6];
  double value[6];
};

static int outer_blocks(bcm_mesh* m) {
  if (m->n_outer >= 0) return BCM_OK;
  std::vector<int2> list;
  for (int b = 0; b < m->nblocks; ++b) {
    int mask = 0;
    for (int f = 0; f < 6; ++f)
      if (m->blocks[b].nb_id[f][0] < 0) mask |= 1 << f;
    if (mask) list.push_back(make_int2(b, mask));
  }
  if (!list.empty()) {
    BCM_TRY(dev_alloc(reinterpret_cast<void**>(&m->d_outer), sizeof(int2) * list.size()));
    BCM_HIP_TRY(hipMemcpy(m->d_outer, list.data(), sizeof(int2) * list.size(), hipMemcpyHostToDevice));
  }
  m->n_outer = int(list.size());
  return BCM_OK;
}

// Ghost values (bcx.f90:18-376, 561-652 through BlockScalar3D.h:216-279):
// Dirichlet g1 = 2v - u, g2 = v; Neumann g1 = u -/+ xc, g2 = u -/+ 2 xc on a
// minus/plus face, xc = v cellSize.x; periodic (a plus face only) copies
// the block's own opposite layers over the whole face, ghosts included.
// The second layer is written when the field has one.  FP64 arithmetic,
// rounded once to the field type.
template <typename T>
__global__ __launch_bounds__(256) void k_face_bc(T* __restrict__ d, const int2* __restrict__ list, FaceBC bc,
                                                 const BlockMeta* __restrict__ meta, Geo g) {
  const int2 it = list[blockIdx.x];
  T* __restrict__ f = d + (long long)it.x * g.bvol;
  const int n = g.n;
  const bool two = g.vc >= 2;
  for (int face = 0; face < 6; ++face) {
    const int type = bc.type[face];
    const int axis = face >> 1;
    const bool plus = face & 1;
    if (!(it.y >> face & 1) || !(type == BCM_BC_DIRICHLET || type == BCM_BC_NEUMANN ||
                                 (type == BCM_BC_PERIODIC && plus)))
      continue;
    if (type == BCM_BC_PERIODIC) {
      const int w = n + 2 * g.vc;
      for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
        int c[3];
        const int p = e % w - g.vc, q = e / w - g.vc;
        if (axis == 0) { c[1] = p; c[2] = q; }
        else if (axis == 1) { c[0] = p; c[2] = q; }
        else { c[0] = p; c[1] = q; }
        auto at = [&](int a) { c[axis] = a; return g.idx(c[0], c[1], c[2]); };
        f[at(n)] = f[at(0)];
        if (two) f[at(n + 1)] = f[at(1)];
        f[at(-1)] = f[at(n - 1)];
        if (two) f[at(-2)] = f[at(n - 2)];
      }
    } else {
      const bool neu = type == BCM_BC_NEUMANN;
      const double v = bc.value[face];
      const double xc = neu ? v * meta[it.x].hx : v;
      for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        int c[3];
        const int p = e % n, q = e / n;
        if (axis == 0) { c[1] = p; c[2] = q; }
        else if (axis == 1) { c[0] = p; c[2] = q; }
        else { c[0] = p; c[1] = q; }
        c[axis] = plus ? n - 1 : 0;
        const double u = double(f[g.idx(c[0], c[1], c[2])]);
        double g1, g2;
        if (neu) {
          g1 = plus ? u + xc : u - xc;
          g2 = plus ? u + xc * 2.0 : u - xc * 2.0;
        } else {
          g1 = 2.0 * xc - u;
          g2 = xc;
        }
        c[axis] = plus ? n : -1;
        f[g.idx(c[0], c[1], c[2])] = T(g1);
        if (two) {
          c[axis] = plus ? n + 1 : -2;
          f[g.idx(c[0], c[1], c[2])] = T(g2);
        }
      }
    }
    __syncthreads();
  }
}

// The same conditions folded into the coefficients of the boundary cells
// (bca.f90:12-394 through BlockScalar3D.h:207-214, 283-345): Dirichlet
// Ap -= Anb, b -= Anb v 2; Neumann Ap += Anb, b += Anb xc on a minus face
// and b -= Anb xc on a plus face; Anb = 0.  Periodic faces are unchanged.
struct FoldPtrs {
  void* Ap;
  void* Anb[6];  // Aw, Ae, As, An, Ab, At
  void* b;
};

template <typename T>
__global__ __launch_bounds__(256) void k_fold_face_bc(FoldPtrs P, const int2* __restrict__ list, FaceBC bc,
                                                      const BlockMeta* __restrict__ meta, Geo g) {
  const int2 it = list[blockIdx.x];
  const long long base = (long long)it.x * g.bvol;
  T* __restrict__ Ap = static_cast<T*>(P.Ap) + base;
  T* __restrict__ b = static_cast<T*>(P.b) + base;
  const int n = g.n;
  for (int face = 0; face < 6; ++face) {
    const int type = bc.type[face];
    if (!(it.y >> face & 1) || !(type == BCM_BC_DIRICHLET || type == BCM_BC_NEUMANN)) continue;
    T* __restrict__ An = static_cast<T*>(P.Anb[face]) + base;
    const int axis = face >> 1;
    const bool plus = face & 1;
    const bool neu = type == BCM_BC_NEUMANN;
    const double v = bc.value[face];
    const double xc = neu ? v * meta[it.x].hx : v;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      int c[3];
      const int p = e % n, q = e / n;
      if (axis == 0) { c[1] = p; c[2] = q; }
      else if (axis == 1) { c[0] = p; c[2] = q; }
      else { c[0] = p; c[1] = q; }
      c[axis] = plus ? n - 1 : 0;
      const int o = g.idx(c[0], c[1], c[2]);
      const double a = double(An[o]);
      if (neu) {
        Ap[o] = T(double(Ap[o]) + a);
        b[o] = T(plus ? double(b[o]) - a * xc : double(b[o]) + a * xc);
      } else {
        Ap[o] = T(double(Ap[o]) - a);
        b[o] = T(double(b[o]) - a * xc * 2.0);
      }
      An[o] = T(0);
    }
    __syncthreads();
  }
}

static int face_bc_args(const int32_t* types, const double* values, FaceBC* bc) {
  if (!types || !values) return BCM_ERR_ARG;
  for (int f = 0; f < 6; ++f) {
    if (types[f] < BCM_BC_INNER || types[f] > BCM_BC_PERIODIC) return BCM_ERR_ARG;
    bc->type[f] = types[f];
    bc->value[f] = values[f];
  }
  return BCM_OK;
}

extern "C" int bcm_impose_face_bc(bcm_field_t f, const int32_t types[6], const double values[6]) {
  if (!f || !f->mesh) return BCM_ERR_ARG;
  FaceBC bc;
  BCM_TRY(face_bc_args(types, values, &bc));
  bcm_mesh* m = f->mesh;
  BCM_TRY(outer_blocks(m));
  if (m->n_outer == 0) return BCM_OK;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_DIRICHLET, &ev));
  if (f->dtype == BCM_F64)
    hipLaunchKernelGGL(k_face_bc<double>, dim3(m->n_outer), dim3(256), 0, m->stream, static_cast<double*>(f->data),
                       (const int2*)m->d_outer, bc, (const BlockMeta*)m->d_meta, g);
  else
    hipLaunchKernelGGL(k_face_bc<float>, dim3(m->n_outer), dim3(256), 0, m->stream, static_cast<float*>(f->data),
                       (const int2*)m->d_outer, bc, (const BlockMeta*)m->d_meta, g);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_fold_face_bc(const bcm_field_t A[7], bcm_field_t b, const int32_t types[6],
                                const double values[6]) {
  if (!A || !b || !b->mesh) return BCM_ERR_ARG;
  for (int i = 0; i < 7; ++i)
    if (!A[i] || A[i]->mesh != b->mesh || A[i]->dtype != b->dtype) return BCM_ERR_ARG;
  FaceBC bc;
  BCM_TRY(face_bc_args(types, values, &bc));
  bcm_mesh* m = b->mesh;
  BCM_TRY(outer_blocks(m));
  if (m->n_outer == 0) return BCM_OK;
  FoldPtrs P;
  P.Ap = A[0]->data;
  for (int i = 0; i < 6; ++i) P.Anb[i] = A[1 + i]->data;
  P.b = b->data;
  Geo g{m->n, m->vc, m->nx0, (long long)m->bvol};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_ASSEMBLE, &ev));
  if (b->dtype == BCM_F64)
    hipLaunchKernelGGL(k_fold_face_bc<double>, dim3(m->n_outer), dim3(256), 0, m->stream, P,
                       (const int2*)m->d_outer, bc, (const BlockMeta*)m->d_meta, g);
  else
    hipLaunchKernelGGL(k_fold_face_bc<float>, dim3(m->n_outer), dim3(256), 0, m->stream, P,
                       (const int2*)m->d_outer, bc, (const BlockMeta*)m->d_meta, g);
  BCM_HIP_TRY(hipGetLastError());
  return timer_end(m, &ev);
}

extern "C" int bcm_dirichlet_mirror(bcm_field_t f, const double* values, int nvalues) {
  if (!f) return BCM_ERR_ARG;
  bcm_mesh* m = f->mesh;
  for (const auto& bi : m->blocks)
    for (int k = 0; k < 6; ++k)
      if (bi.bc_type[k] == BCM_BC_NEUMANN) return BCM_ERR_ARG;  // not implemented, as in sor1/sor2
  BCM_TRY(set_bvalues(m, values, nvalues));
  if (f->dtype == BCM_F64) return dirichlet_raw<double>(m, static_cast<double*>(f->data));
  return dirichlet_raw<float>(m, static_cast<float*>(f->data));
}
