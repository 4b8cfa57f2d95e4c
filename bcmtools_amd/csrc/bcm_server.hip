// Request server: Poisson solves of independent meshes arriving over time,
// run with continuous batching on one GPU.
//
// Admitted requests run in groups.  A group owns a device mesh (its own
// HIP stream), its fields and its fused-solver state; the server loop keeps
// every group's stream fed with hipGraph-replayed blocks of iterations and
// polls completion events without blocking, so the work of all groups
// overlaps on the device.  With batch_blocks = 0 every request is a group
// of its own.  With batch_blocks > 0, requests of at most that many blocks
// are merged at admission, in arrival order and with the same precision,
// into one group of at most batch_blocks blocks: one batched mesh whose
// segments are the requests (bcm_mesh_create_batch), so one launch of each
// solver kernel serves all of them while their scalars and tolerances stay
// per segment.  Requests wait in FIFO order and are admitted (first fit)
// while the device bytes they need fit in the HBM budget and a group slot
// is free; a finished group's memory is released at once, which admits the
// next ones.  Each request is computed exactly as it would be alone (a
// segment's kernels read only its own blocks and its reductions are its
// own), so its result does not depend on what else is running or on its
// group.
//
// The reference has no counterpart (one BlockManager per process,
// include/BlockManager.h:153-156); a request's solve is bcm_solve's fused
// BiCGSTAB (examples/SolverTest/ILS3D.h:663-844).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <list>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "bcm_internal.h"

namespace bcm {
namespace {

// Synthetic right-hand side of a request (rhs == NULL at submission):
// f = sin(pi x) sin(pi y) sin(pi z) + sigma N(0, 1) at cell centres, the
// noise from a counter-based generator keyed by (seed, block pedigree,
// cell), so a request's data depends on nothing but the request.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct BlockGeo {
  double ox, oy, oz, hx, hy, hz;
  unsigned long long ped;
};

__global__ __launch_bounds__(256) void k_rhs_synth(double* __restrict__ rhs, const BlockGeo* __restrict__ geo, int n,
                                                   unsigned long long seed, double sigma) {
  const int b = blockIdx.x;
  const BlockGeo g = geo[b];
  const int nx = n + 2;
  const long long base = (long long)b * nx * nx * nx;
  const unsigned long long key = mix64(seed * 0xD1B54A32D192ED03ULL + mix64(g.ped));
  const double pi = 3.14159265358979323846;
  for (int c = threadIdx.x; c < n * n * n; c += 256) {
    const int i = c % n, j = (c / n) % n, k = c / (n * n);
    const double x = g.ox + (i + 0.5) * g.hx, y = g.oy + (j + 0.5) * g.hy, z = g.oz + (k + 0.5) * g.hz;
    const unsigned long long h1 = mix64(key + (unsigned long long)c);
    const unsigned long long h2 = mix64(h1 ^ 0x632BE59BD9B4E019ULL);
    const double u1 = 1.0 - double(h1 >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double nz = sqrt(-2.0 * log(u1)) * cos(2.0 * pi * u2);
    const double f = sin(pi * z) * sin(pi * y) * sin(pi * x) + sigma * nz;
    // Right-hand side of A u = rhs with A = lap_h: rhs = -f.
    rhs[base + (i + 1) + nx * ((j + 1) + nx * (k + 1))] = -f;
  }
}

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t tag = 0;
  std::vector<bcm_block_info> blocks;
  const double* rhs_host = nullptr;
  unsigned long long seed = 0;
  double tol = 1e-10;
  int precision = BCM_PREC_MIXED;
  double* x_out = nullptr;
  double t_submit = 0.0, t_admit = 0.0;
  size_t bytes = 0;
  // Initial guess from a restart file (bcm_server_submit_restart).
  bool restart = false;
  std::string r_dir, r_label;
  int r_rank = 0, r_step = 0;
};

// Requests sharing one device mesh (one segment each, in order).
struct Group {
  std::vector<Request*> reqs;
  std::vector<int> first;  // first block of each request's segment
  size_t bytes = 0;        // charged to the budget
  int nblocks = 0;
  bcm_mesh_t mesh = nullptr;
  bcm_field_t x = nullptr, rhs = nullptr;
  double ready_since = -1.0;  // when its next block of iterations became due and was held back
};

}  // namespace
}  // namespace bcm

struct bcm_server {
  bcm_server_params p{};
  std::list<bcm::Request*> queue;
  std::vector<bcm::Group*> active;
  std::deque<bcm_server_result> done;
  size_t in_use = 0;
  bcm::Clock::time_point t0;
  double now() const { return std::chrono::duration<double>(bcm::Clock::now() - t0).count(); }
};

namespace bcm {
namespace {

// Device bytes of one request: mesh tables, x and rhs, the solver state.
size_t request_bytes(const bcm_server* sv, int nblocks, int precision) {
  const int n = sv->p.n;
  const size_t ng = size_t(nblocks) * (n + 2) * (n + 2) * (n + 2);
  return size_t(nblocks) * 640 + 2 * ng * 8 + krylov_bytes(nblocks, n, precision, sv->p.count_pre);
}

void release(bcm_server* sv, Group* g) {
  if (g->x) bcm_field_destroy(g->x);
  if (g->rhs) bcm_field_destroy(g->rhs);
  if (g->mesh) bcm_mesh_destroy(g->mesh);
  sv->in_use -= g->bytes;
  for (Request* r : g->reqs) delete r;
  delete g;
}

void finish_result(bcm_server* sv, Request* r, int status, const bcm_solve_stats* st, double t_done = -1.0) {
  bcm_server_result o;
  std::memset(&o, 0, sizeof(o));
  o.tag = r->tag;
  o.status = status;
  o.nblocks = int(r->blocks.size());
  o.precision = r->precision;
  o.t_submit = r->t_submit;
  o.t_admit = r->t_admit;
  o.t_done = t_done >= 0.0 ? t_done : sv->now();
  if (st) {
    o.iters = st->iters;
    o.relres = st->relres;
  }
  sv->done.push_back(o);
}

int admit(bcm_server* sv, Group* g) {
  const int n = sv->p.n;
  const int nseg = int(g->reqs.size());
  std::vector<bcm_block_info> blocks;
  std::vector<int> counts;
  blocks.reserve(size_t(g->nblocks));
  for (Request* r : g->reqs) {
    g->first.push_back(int(blocks.size()));
    counts.push_back(int(r->blocks.size()));
    blocks.insert(blocks.end(), r->blocks.begin(), r->blocks.end());
  }
  if (nseg == 1) BCM_TRY(bcm_mesh_create(blocks.data(), int(blocks.size()), n, 1, &g->mesh));
  else BCM_TRY(bcm_mesh_create_batch(blocks.data(), counts.data(), nseg, n, 1, &g->mesh));
  BCM_TRY(bcm_field_create(g->mesh, BCM_F64, &g->x));
  BCM_TRY(bcm_field_create(g->mesh, BCM_F64, &g->rhs));
  bcm_mesh* m = g->mesh;
  const size_t bvol = size_t(n + 2) * (n + 2) * (n + 2);
  std::vector<BlockGeo> geo;
  for (int s = 0; s < nseg; ++s) {
    Request* r = g->reqs[size_t(s)];
    const int nb = int(r->blocks.size());
    if (r->rhs_host) {
      BCM_TRY(bcm_field_upload_blocks(g->rhs, g->first[size_t(s)], nb, r->rhs_host, sizeof(double) * bvol * nb));
      continue;
    }
    geo.resize(size_t(nb));
    for (int b = 0; b < nb; ++b) {
      const bcm_block_info& bi = r->blocks[size_t(b)];
      geo[size_t(b)] = BlockGeo{bi.origin[0], bi.origin[1], bi.origin[2], bi.block_size[0] / n,
                                bi.block_size[1] / n, bi.block_size[2] / n, (unsigned long long)bi.pedigree};
    }
    void* d = nullptr;
    BCM_TRY(dev_alloc(&d, sizeof(BlockGeo) * geo.size()));
    hipError_t e = hipMemcpyAsync(d, geo.data(), sizeof(BlockGeo) * geo.size(), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_rhs_synth, dim3(unsigned(nb)), dim3(256), 0, m->stream,
                         static_cast<double*>(g->rhs->data) + bvol * size_t(g->first[size_t(s)]),
                         static_cast<const BlockGeo*>(d), n, r->seed, 0.01);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    dev_free(d);
    if (e != hipSuccess) return BCM_ERR_HIP;
  }
  // Initial guesses of resumed requests; the others start from zero.
  for (int s = 0; s < nseg; ++s) {
    const Request* r = g->reqs[size_t(s)];
    if (r->restart)
      BCM_TRY(load2_into(g->x, g->first[size_t(s)], int(r->blocks.size()), r->r_dir.c_str(), r->r_label.c_str(),
                         r->r_rank, r->r_step));
  }
  const double zero = 0.0;
  BCM_TRY(set_bvalues(m, &zero, 1));
  bcm_solve_params P;
  std::memset(&P, 0, sizeof(P));
  P.precision = g->reqs[0]->precision;
  P.tol = g->reqs[0]->tol;
  P.max_iter = sv->p.max_iter;
  P.count_pre = sv->p.count_pre;
  P.omega = sv->p.omega;
  P.check_every = sv->p.check_every;
  P.use_graph = 1;
  std::vector<double> tol;
  for (Request* r : g->reqs) tol.push_back(r->tol);
  BCM_TRY(krylov_check(m, &P, tol.data()));
  return krylov_begin(m, static_cast<double*>(g->x->data), static_cast<const double*>(g->rhs->data), P,
                      tol.data());
}

// Finish a group whose solve is over: each request's outcome and solution.
int complete(bcm_server* sv, Group* g) {
  const int nseg = int(g->reqs.size());
  std::vector<bcm_solve_stats> st(static_cast<size_t>(nseg));
  BCM_TRY(krylov_finish(g->mesh, st.data()));
  const size_t bvol = size_t(sv->p.n + 2) * (sv->p.n + 2) * (sv->p.n + 2);
  const double t_done = sv->now();  // the group's solve is over for every member
  for (int s = 0; s < nseg; ++s) {
    Request* r = g->reqs[size_t(s)];
    const int nb = int(r->blocks.size());
    if (r->x_out)
      BCM_TRY(bcm_field_download_blocks(g->x, g->first[size_t(s)], nb, r->x_out, sizeof(double) * bvol * nb));
    finish_result(sv, r, st[size_t(s)].status, &st[size_t(s)], t_done);
  }
  return BCM_OK;
}

// Admit queued requests (first fit in arrival order) while group slots and
// the HBM budget allow; with batch_blocks > 0 a small request takes the
// later queued requests that fit with it (same precision) into its group.
// A request larger than the whole budget is refused.
void admit_ready(bcm_server* sv) {
  const int cap = sv->p.batch_blocks;
  for (auto it = sv->queue.begin(); it != sv->queue.end() && int(sv->active.size()) < sv->p.max_active;) {
    Request* r = *it;
    if (double(r->bytes) > sv->p.hbm_budget) {
      finish_result(sv, r, -BCM_ERR_BUSY, nullptr);
      it = sv->queue.erase(it);
      delete r;
      continue;
    }
    if (double(sv->in_use + r->bytes) > sv->p.hbm_budget) {
      ++it;
      continue;
    }
    it = sv->queue.erase(it);
    Group* g = new Group();
    const double now = sv->now();
    auto take = [&](Request* q) {
      q->t_admit = now;
      g->reqs.push_back(q);
      g->nblocks += int(q->blocks.size());
      g->bytes += q->bytes;
      sv->in_use += q->bytes;
    };
    take(r);
    if (cap > 0 && g->nblocks <= cap) {
      for (auto jt = it; jt != sv->queue.end();) {
        Request* q = *jt;
        if (q->precision == r->precision && g->nblocks + int(q->blocks.size()) <= cap &&
            double(sv->in_use + q->bytes) <= sv->p.hbm_budget) {
          if (jt == it) ++it;
          jt = sv->queue.erase(jt);
          take(q);
        } else {
          ++jt;
        }
      }
    }
    const int rc = admit(sv, g);
    if (rc != BCM_OK) {
      for (Request* q : g->reqs) finish_result(sv, q, -rc, nullptr);
      release(sv, g);
      continue;
    }
    sv->active.push_back(g);
  }
}

// Ends the solve of active group i (over, or failed with rc).
void finish_group(bcm_server* sv, size_t i, int rc) {
  Group* g = sv->active[i];
  if (rc == BCM_OK) rc = complete(sv, g);
  if (rc != BCM_OK)
    for (Request* q : g->reqs) finish_result(sv, q, -rc, nullptr);
  sv->active.erase(sv->active.begin() + long(i));
  release(sv, g);
}

// One pass, window_blocks = 0: every group whose previous block of
// iterations is done gets its next one.  Returns whether a group finished.
bool pass_all(bcm_server* sv) {
  bool moved = false;
  for (size_t i = 0; i < sv->active.size();) {
    bool act = true;
    const int rc = krylov_advance(sv->active[i]->mesh, false, &act);
    if (rc == BCM_OK && act) {
      ++i;
      continue;
    }
    moved = true;
    finish_group(sv, i, rc);
  }
  return moved;
}

// One pass, window_blocks > 0: the groups whose previous block is done are
// served smallest first, each while the blocks in flight stay within the
// window (or nothing is in flight, or it has waited max_wait seconds).
bool pass_windowed(bcm_server* sv) {
  bool moved = false;
  const double now = sv->now();
  long long inflight = 0;
  std::vector<Group*> ready;
  for (size_t i = 0; i < sv->active.size();) {
    Group* g = sv->active[i];
    int state = 0;
    const int rc = krylov_state(g->mesh, &state);
    if (rc == BCM_OK && state == 0) {
      inflight += g->nblocks;
      ++i;
    } else if (rc == BCM_OK && state == 1) {
      ready.push_back(g);
      ++i;
    } else {
      moved = true;
      finish_group(sv, i, rc);
    }
  }
  std::stable_sort(ready.begin(), ready.end(), [](const Group* a, const Group* b) { return a->nblocks < b->nblocks; });
  for (Group* g : ready) {
    if (g->ready_since < 0.0) g->ready_since = now;
    const bool aged = sv->p.max_wait > 0.0 && now - g->ready_since >= sv->p.max_wait;
    if (!(inflight == 0 || inflight + g->nblocks <= sv->p.window_blocks || aged)) continue;
    bool act = true;
    const int rc = krylov_advance(g->mesh, false, &act);
    if (rc == BCM_OK && act) {
      inflight += g->nblocks;
      g->ready_since = -1.0;
      continue;
    }
    moved = true;
    finish_group(sv, size_t(std::find(sv->active.begin(), sv->active.end(), g) - sv->active.begin()), rc);
  }
  return moved;
}

}  // namespace
}  // namespace bcm

using namespace bcm;

static int active_requests(const bcm_server* sv) {
  int k = 0;
  for (const Group* g : sv->active) k += int(g->reqs.size());
  return k;
}

extern "C" int bcm_server_create(const bcm_server_params* p, bcm_server_t* out) {
  if (!p || !out) return BCM_ERR_ARG;
  *out = nullptr;
  if (!(p->hbm_budget > 0.0) || p->max_active < 1 || (p->n != 8 && p->n != 16 && p->n != 32) || p->max_iter < 1 ||
      p->count_pre < 0 || p->count_pre > 16 || p->check_every < 0 || p->batch_blocks < 0 || p->window_blocks < 0 ||
      !(p->max_wait >= 0.0))
    return BCM_ERR_ARG;
  bcm_server* sv = new (std::nothrow) bcm_server();
  if (!sv) return BCM_ERR_NOMEM;
  sv->p = *p;
  sv->t0 = Clock::now();
  *out = sv;
  return BCM_OK;
}

extern "C" int bcm_server_destroy(bcm_server_t sv) {
  if (!sv) return BCM_OK;
  for (Group* g : sv->active) release(sv, g);
  for (Request* r : sv->queue) delete r;
  delete sv;
  return BCM_OK;
}

extern "C" double bcm_server_clock(bcm_server_t sv) { return sv ? sv->now() : -1.0; }

static int submit(bcm_server_t sv, const bcm_block_info* blocks, int nblocks, const double* rhs, uint64_t seed,
                  double tol, int precision, double* x_out, int64_t tag, Request* proto) {
  if (!sv || !blocks || nblocks <= 0 || !(tol > 0.0) || precision < BCM_PREC_FP64 || precision > BCM_PREC_FP32)
    return BCM_ERR_ARG;
  for (int b = 0; b < nblocks; ++b)
    for (int f = 0; f < 6; ++f)
      for (int q = 0; q < 4; ++q)
        if (blocks[b].nb_id[f][q] < -1 || blocks[b].nb_id[f][q] >= nblocks) return BCM_ERR_ARG;
  // A tag names one pending request: the caller's buffers are found by it.
  for (const Request* q : sv->queue)
    if (q->tag == tag) return BCM_ERR_ARG;
  for (const Group* g : sv->active)
    for (const Request* q : g->reqs)
      if (q->tag == tag) return BCM_ERR_ARG;
  if (proto && proto->restart)
    BCM_TRY(load2_header(proto->r_dir.c_str(), proto->r_label.c_str(), proto->r_rank, proto->r_step, sv->p.n, 1,
                         nblocks, nullptr));
  Request* r = new (std::nothrow) Request();
  if (!r) return BCM_ERR_NOMEM;
  if (proto) *r = *proto;
  r->tag = tag;
  r->blocks.assign(blocks, blocks + nblocks);
  r->rhs_host = rhs;
  r->seed = seed;
  r->tol = tol;
  r->precision = precision;
  r->x_out = x_out;
  r->t_submit = sv->now();
  r->bytes = request_bytes(sv, nblocks, precision);
  sv->queue.push_back(r);
  admit_ready(sv);
  return BCM_OK;
}

extern "C" int bcm_server_submit(bcm_server_t sv, const bcm_block_info* blocks, int nblocks, const double* rhs,
                                 uint64_t seed, double tol, int precision, double* x_out, int64_t tag) {
  return submit(sv, blocks, nblocks, rhs, seed, tol, precision, x_out, tag, nullptr);
}

extern "C" int bcm_server_submit_restart(bcm_server_t sv, const bcm_block_info* blocks, int nblocks,
                                         const double* rhs, uint64_t seed, double tol, int precision,
                                         const char* dir, const char* label, int rank, int step, double* x_out,
                                         int64_t tag) {
  if (!dir || !label) return BCM_ERR_ARG;
  Request proto;
  proto.restart = true;
  proto.r_dir = dir;
  proto.r_label = label;
  proto.r_rank = rank;
  proto.r_step = step;
  return submit(sv, blocks, nblocks, rhs, seed, tol, precision, x_out, tag, &proto);
}

extern "C" int bcm_server_checkpoint(bcm_server_t sv, int64_t tag, const char* dir, const char* label, int rank,
                                     int step, int* iters) {
  if (!sv || !dir || !label) return BCM_ERR_ARG;
  for (const Request* q : sv->queue)
    if (q->tag == tag) return BCM_ERR_STATE;  // not admitted: no iterate yet
  for (Group* g : sv->active)
    for (size_t s = 0; s < g->reqs.size(); ++s) {
      if (g->reqs[s]->tag != tag) continue;
      bcm_mesh* m = g->mesh;
      const int nb = int(g->reqs[s]->blocks.size());
      void* y = nullptr;
      BCM_TRY(dev_alloc(&y, sizeof(double) * m->bvol * size_t(nb)));
      int rc = krylov_snapshot(m, int(s), static_cast<double*>(y), iters);
      if (rc == BCM_OK) rc = dump2_raw(m, y, nb, 8, dir, label, rank, step);
      (void)hipStreamSynchronize(m->stream);
      dev_free(y);
      return rc;
    }
  return BCM_ERR_ARG;
}

extern "C" int bcm_server_progress(bcm_server_t sv, double max_seconds, int* n_active, int* n_queued) {
  if (!sv || !(max_seconds >= 0.0)) return BCM_ERR_ARG;
  const double t_end = sv->now() + max_seconds;
  do {
    const bool moved = sv->p.window_blocks > 0 ? pass_windowed(sv) : pass_all(sv);
    if (moved) admit_ready(sv);
    if (sv->active.empty() && sv->queue.empty()) break;
    if (!moved) std::this_thread::sleep_for(std::chrono::microseconds(20));
  } while (sv->now() < t_end);
  if (n_active) *n_active = active_requests(sv);
  if (n_queued) *n_queued = int(sv->queue.size());
  return BCM_OK;
}

extern "C" int bcm_server_poll(bcm_server_t sv, bcm_server_result* out, int capacity) {
  if (!sv || (capacity > 0 && !out) || capacity < 0) return -BCM_ERR_ARG;
  int k = 0;
  while (k < capacity && !sv->done.empty()) {
    out[k++] = sv->done.front();
    sv->done.pop_front();
  }
  return k;
}

extern "C" int bcm_server_stats(bcm_server_t sv, double* bytes_in_use, int* n_active, int* n_queued) {
  if (!sv) return BCM_ERR_ARG;
  if (bytes_in_use) *bytes_in_use = double(sv->in_use);
  if (n_active) *n_active = active_requests(sv);
  if (n_queued) *n_queued = int(sv->queue.size());
  return BCM_OK;
}
