// Many sweeps of ILS3D's Jacobi iteration on the Poisson operator in one
// call: bcm_jacobi_poisson_iterate.
//
// Each step is jacobi_smoother (blas.f90:452-502, matrix-free form of
// bcm_jacobi7_poisson) followed by the ghost exchange (updateVC,
// include/Scalar3DUpdater.h:190-347, 519-708), the loop of
// ILS3D::Jacobi_Smoother (examples/SolverTest/ILS3D.h:65-126) without the
// host between steps.
//
// When every block fits in LDS at once (small meshes: one work-group per
// block, all co-resident), the whole iteration runs in one cooperative
// launch: each work-group keeps its block's ghosted array in LDS for every
// step, computes the new interior, publishes it to a global ping-pong
// buffer, waits at a grid barrier and fills its face ghosts from the
// neighbours' published interiors with the exchange's arithmetic.  A mesh
// too large for that runs the same steps as two launches per step on the
// mesh stream.  Both give the bits of the step-by-step calls.
#include <algorithm>
#include <cstdio>
#include <utility>
#include <vector>

#include "bcm_internal.h"
#include "bcm_interp.h"
#include "bcm_poisson.h"

namespace bcm {
namespace {

constexpr int kResWG = 512;

template <int N>
struct ResBlk {
  static constexpr int NX = N + 2;
  static constexpr int GV = NX * NX * NX;
  static constexpr int N3 = N * N * N;
  static constexpr int PER = (N3 + kResWG - 1) / kResWG;
  __device__ static __forceinline__ int at(int i, int j, int k) { return (i + 1) + NX * ((j + 1) + NX * (k + 1)); }
};

// A block's published interior (compact, x fastest).
template <int N>
struct Pub {
  const double* __restrict__ p;
  __device__ __forceinline__ double operator()(int blk, int i, int j, int k) const {
    return p[(long long)blk * (N * N * N) + i + N * (j + N * k)];
  }
};

// The ghost value of face f at in-face position (u, v) from the
// neighbours' published interiors (the task kernel's copy / F2C / C2F);
// returns false for a face the exchange leaves alone.
template <int N>
__device__ __forceinline__ bool exchange_value(const Pub<N>& src, const int* df, int f, int u, int v, int (&c)[3],
                                               double& val) {
  constexpr int h = N / 2;
  const int kind = df[0];
  if (kind != FD_COPY && kind != FD_F2C && kind != FD_C2F) return false;
  const int ax = f >> 1;
  const bool plus = f & 1;
  const int t0 = (ax + 1) % 3, t1 = (ax + 2) % 3;
  c[ax] = plus ? N : -1;
  c[t0] = u;
  c[t1] = v;
  if (kind == FD_COPY) {
    int sc[3] = {c[0], c[1], c[2]};
    sc[ax] = plus ? 0 : N - 1;
    val = src(df[1], sc[0], sc[1], sc[2]);
  } else if (kind == FD_F2C) {
    const int s0 = u >= h, s1 = v >= h;
    const int fb = df[1 + s0 + 2 * s1];
    int I[3];
    I[ax] = plus ? 0 : h - 1;
    I[t0] = u - h * s0;
    I[t1] = v - h * s1;
    val = f2c_sum<double>([&](int i, int j, int k) { return src(fb, i, j, k); }, I[0], I[1], I[2]);
  } else {
    const int sf = df[5], cb = df[1];
    int p[3];
    p[ax] = plus ? 0 : 2 * N - 1;
    p[t0] = u + N * (sf & 1);
    p[t1] = v + N * (sf >> 1);
    val = c2f_interp<double>([&](int i, int j, int k) { return src(cb, i, j, k); }, N, p[0], p[1], p[2]);
  }
  return true;
}

// Publication flags: ctl[0] is a failure flag, ctl[1 + b] the number of
// steps block b has published.  A work-group's stores of a step must have
// reached the device's L2 before its flag is raised: every thread waits for
// its own stores (s_waitcnt vmcnt(0)), the work-group barrier orders them
// before thread 0's release store.
__device__ __forceinline__ void publish(unsigned* ctl, int blk, unsigned steps) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(ctl + 1 + blk, steps, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until every neighbour in nbl[0..nnb) has published `steps` steps
// (one polling thread per neighbour, acquire at device scope), then a
// work-group barrier.  The wait is bounded: past the limit it raises
// ctl[0], which ends every later wait at once, so a wait that could never
// end returns an error instead of holding the GPU.
__device__ __forceinline__ void await_neighbours(unsigned* ctl, const int* nbl, int nnb, unsigned steps) {
  if (threadIdx.x < nnb) {
    const unsigned* flag = ctl + 1 + nbl[threadIdx.x];
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < steps) {
      if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > (1u << 24)) {
        __hip_atomic_store(ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
}

__global__ void k_res_reset(unsigned* ctl, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) ctl[i] = 0u;
}

template <int N>
__global__ __launch_bounds__(kResWG) void k_jacobi_resident(double* __restrict__ x, const double* __restrict__ b,
                                                            const BlockMeta* __restrict__ meta,
                                                            const int* __restrict__ desc, double omega, int nsweeps,
                                                            int nblocks, double* __restrict__ pub,
                                                            unsigned* __restrict__ ctl) {
  using B = ResBlk<N>;
  extern __shared__ double cur[];  // B::GV values: the block with its ghosts
  __shared__ int d[48];
  const int blk = blockIdx.x;
  const int tid = threadIdx.x;
  double* __restrict__ xb = x + (long long)blk * B::GV;
  for (int e = tid; e < B::GV; e += kResWG) cur[e] = xb[e];
  if (tid < 48) d[tid] = desc[blk * 48 + tid];
  // The blocks this block's exchange reads (24 at most: four per face).
  __shared__ int nbl[24];
  __shared__ int nnb;
  __syncthreads();
  if (tid == 0) {
    int k = 0;
    for (int f = 0; f < 6; ++f) {
      const int kind = d[f * 8];
      if (kind == FD_COPY || kind == FD_C2F) nbl[k++] = d[f * 8 + 1];
      else if (kind == FD_F2C)
        for (int q = 0; q < 4; ++q) nbl[k++] = d[f * 8 + 1 + q];
    }
    nnb = k;
  }
  const BlockMeta me = meta[blk];
  const double* __restrict__ bb = b + (long long)blk * B::GV;
  double bv[B::PER];
#pragma unroll
  for (int q = 0; q < B::PER; ++q) {
    const int c = tid + q * kResWG;
    if (c < B::N3) bv[q] = bb[B::at(c % N, (c / N) % N, c / (N * N))];
  }
  __syncthreads();
  for (int s = 0; s < nsweeps; ++s) {
    double* __restrict__ pw = pub + (long long)(s & 1) * nblocks * B::N3;
    double nv[B::PER];
#pragma unroll
    for (int q = 0; q < B::PER; ++q) {
      const int c = tid + q * kResWG;
      if (c >= B::N3) continue;
      const int i = c % N, j = (c / N) % N, k = c / (N * N);
      const Coef7<double> a = poisson_coef<double>(me, i, j, k, N);
      nv[q] = jacobi_cell(a, cur, B::at(i, j, k), B::NX, B::NX * B::NX, bv[q], omega);
      pw[(long long)blk * B::N3 + c] = nv[q];
    }
    publish(ctl, blk, unsigned(s + 1));  // also: every read of the old values is done
#pragma unroll
    for (int q = 0; q < B::PER; ++q) {
      const int c = tid + q * kResWG;
      if (c < B::N3) cur[B::at(c % N, (c / N) % N, c / (N * N))] = nv[q];
    }
    await_neighbours(ctl, nbl, nnb, unsigned(s + 1));
    const Pub<N> src{pw};
    for (int e = tid; e < 6 * N * N; e += kResWG) {
      const int f = e / (N * N), uv = e % (N * N);
      int c[3];
      double val;
      if (exchange_value<N>(src, d + f * 8, f, uv % N, uv / N, c, val)) cur[B::at(c[0], c[1], c[2])] = val;
    }
    __syncthreads();
  }
  for (int e = tid; e < B::GV; e += kResWG) xb[e] = cur[e];
}

// Every face with a neighbour is one the block-wise exchange can fill: a
// same-level or coarser neighbour, or a finer one on all four subfaces.
// The blocks a block's exchange reads must also read it (so a block that
// waits for its readers' next step never overwrites a value still to be
// read); every octree mesh has this symmetry, which is checked here.
bool exchange_complete(const bcm_mesh* m) {
  const int nb = m->nblocks;
  std::vector<std::vector<int>> reads(nb);
  for (int b = 0; b < nb; ++b) {
    const bcm_block_info& bi = m->blocks[b];
    for (int f = 0; f < 6; ++f) {
      if (bi.nb_id[f][0] < 0) continue;
      const int ns = bi.nb_level_diff[f] == 1 ? 4 : 1;
      for (int s = 0; s < ns; ++s) {
        const int c = bi.nb_id[f][s];
        if (c < 0 || c >= nb) return false;
        reads[b].push_back(c);
      }
    }
  }
  for (int b = 0; b < nb; ++b)
    for (int c : reads[b])
      if (std::find(reads[c].begin(), reads[c].end(), b) == reads[c].end()) return false;
  return true;
}

// The one-launch form applies when the device supports cooperative
// launches, every block's work-group can be resident at once and the
// exchange is complete.
template <int N>
int resident_ok(bcm_mesh* m, bool* ok) {
  *ok = false;
  int dev = 0, coop = 0, cus = 0;
  BCM_HIP_TRY(hipGetDevice(&dev));
  BCM_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  BCM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (!coop || !exchange_complete(m)) return BCM_OK;
  int per_cu = 0;
  BCM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jacobi_resident<N>, kResWG,
                                                           sizeof(double) * ResBlk<N>::GV));
  *ok = (long long)per_cu * cus >= m->nblocks;
  return BCM_OK;
}

template <int N>
int resident_launch(bcm_mesh* m, double* x, const double* b, double omega, int nsweeps) {
  int rc = BCM_OK;
  // Two published interiors per block, then the failure flag and one
  // publication count per block.
  const size_t pub_bytes = sizeof(double) * 2 * size_t(m->nblocks) * ResBlk<N>::N3;
  char* ws = static_cast<char*>(workspace(m, WS_RES_PUB, pub_bytes + sizeof(unsigned) * (size_t(m->nblocks) + 1),
                                          &rc));
  if (rc != BCM_OK) return rc;
  double* pub = reinterpret_cast<double*>(ws);
  unsigned* ctl = reinterpret_cast<unsigned*>(ws + pub_bytes);
  int nb = m->nblocks;
  const BlockMeta* meta = m->d_meta;
  const int* desc = m->d_desc;
  void* args[] = {&x, &b, &meta, &desc, &omega, &nsweeps, &nb, &pub, &ctl};
  hipLaunchKernelGGL(k_res_reset, dim3(1), dim3(256), 0, m->stream, ctl, nb + 1);
  BCM_HIP_TRY(hipGetLastError());
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_JACOBI, &ev));
  BCM_HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_jacobi_resident<N>), dim3(nb),
                                         dim3(kResWG), args, sizeof(double) * ResBlk<N>::GV, m->stream));
  BCM_HIP_TRY(hipGetLastError());
  BCM_TRY(timer_end(m, &ev));
  unsigned fail = 0;
  BCM_HIP_TRY(hipMemcpyAsync(&fail, ctl, sizeof(fail), hipMemcpyDeviceToHost, m->stream));
  BCM_HIP_TRY(hipStreamSynchronize(m->stream));
  if (fail) {
    std::fprintf(stderr, "bcm: bcm_jacobi_poisson_iterate: a wait for the neighbours did not complete\n");
    return BCM_ERR_HIP;
  }
  return BCM_OK;
}

// Meshes too large to be resident: the halo-gathering schedule.  Step 1
// reads x with its stored ghosts and writes the compact iterate, its X
// export and the compact right-hand side; steps 2..nsweeps build each
// block's ghost frame from the previous compact iterate (ping-pong in the
// SOR schedule's workspace buffers); the last writes x's rows; then one
// exchange completes x's ghosts.
int halo_iterate(bcm_mesh* m, double* x, const double* b, double omega, int nsweeps) {
  int rc = BCM_OK;
  const size_t xbytes = size_t(m->nblocks) * 4 * m->n * m->n * sizeof(double);
  for (int i = 0; i < 2; ++i)
    if (!m->d_xexp[i]) BCM_TRY(dev_alloc(&m->d_xexp[i], xbytes));
  const size_t cbytes = m->ncell * sizeof(double);
  double* A = static_cast<double*>(workspace(m, WS_SOR_CA_F64, cbytes, &rc));
  if (rc != BCM_OK) return rc;
  double* B = static_cast<double*>(workspace(m, WS_SOR_CB_F64, cbytes, &rc));
  if (rc != BCM_OK) return rc;
  double* R = static_cast<double*>(workspace(m, WS_SOR_CS_F64, cbytes, &rc));
  if (rc != BCM_OK) return rc;
  double* xe[2] = {static_cast<double*>(m->d_xexp[0]), static_cast<double*>(m->d_xexp[1])};
  EventPair ev;
  BCM_TRY(timer_begin(m, BCM_K_JACOBI, &ev));
  BCM_TRY(jacobi_first_raw<double>(m, x, b, omega, A, xe[0], R));
  BCM_TRY(timer_end(m, &ev));
  int cur = 0;
  for (int it = 1; it < nsweeps; ++it) {
    const bool last = it == nsweeps - 1;
    BCM_TRY(timer_begin(m, last ? BCM_K_JACOBI : BCM_K_JACOBI_HALO, &ev));
    BCM_TRY(jacobi_halo_raw<double>(m, A, B, R, omega, xe[cur], xe[cur ^ 1], x, last));
    BCM_TRY(timer_end(m, &ev));
    std::swap(A, B);
    cur ^= 1;
  }
  return ghost_fill_raw<double>(m, x, BCM_GF_ALL);
}

}  // namespace
}  // namespace bcm

using namespace bcm;

extern "C" int bcm_jacobi_poisson_iterate(bcm_field_t x, bcm_field_t b, double omega, int nsweeps) {
  if (!x || !b || x->mesh != b->mesh || x->dtype != BCM_F64 || b->dtype != BCM_F64 || nsweeps < 0)
    return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  if (nsweeps == 0) return BCM_OK;
  double* xd = static_cast<double*>(x->data);
  const double* bd = static_cast<const double*>(b->data);
  bool ok = false;
  if (m->vc == 1 && m->n == 16) {
    BCM_TRY(resident_ok<16>(m, &ok));
    if (ok) return resident_launch<16>(m, xd, bd, omega, nsweeps);
  } else if (m->vc == 1 && m->n == 8) {
    BCM_TRY(resident_ok<8>(m, &ok));
    if (ok) return resident_launch<8>(m, xd, bd, omega, nsweeps);
  }
  if (m->vc == 1 && (m->n == 16 || m->n == 32) && m->jhalo_ok && nsweeps >= 2)
    return halo_iterate(m, xd, bd, omega, nsweeps);
  // Two launches per step, ping-pong with a copy of x (so the cells the
  // exchange never writes keep x's values whichever buffer ends last).
  int rc = BCM_OK;
  double* w = static_cast<double*>(workspace(m, WS_RES_PING, sizeof(double) * m->nblocks * m->bvol, &rc));
  if (rc != BCM_OK) return rc;
  BCM_HIP_TRY(hipMemcpyAsync(w, xd, sizeof(double) * m->nblocks * m->bvol, hipMemcpyDeviceToDevice, m->stream));
  double* cur = xd;
  double* nxt = w;
  for (int s = 0; s < nsweeps; ++s) {
    BCM_TRY(jacobi_poisson_raw<double>(m, nxt, cur, bd, omega));
    std::swap(cur, nxt);
    BCM_TRY(ghost_fill_raw<double>(m, cur, BCM_GF_ALL));
  }
  if (cur != xd)
    BCM_HIP_TRY(hipMemcpyAsync(xd, cur, sizeof(double) * m->nblocks * m->bvol, hipMemcpyDeviceToDevice, m->stream));
  return BCM_OK;
}
