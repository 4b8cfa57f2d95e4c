// Host-side parallel loops of the mesh builders (bcm_tree.cpp: the octree
// and its block table; bcm_mesh.hip: the face lists and device tables of
// bcm_mesh_create).  Plain C++ threads; no device code.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <system_error>
#include <thread>
#include <vector>

namespace bcm {

// Threads of the host builders (bcm_tree_set_threads; 0: one per hardware
// thread, at most 64).  Defined in bcm_tree.cpp.
extern std::atomic<int> g_host_threads;

inline int host_threads() {
  const int want = g_host_threads.load();
  const int t = want > 0 ? want : int(std::thread::hardware_concurrency());
  return std::max(1, std::min(t, 64));
}

// Chunks for n items of at least `grain` each (1: run serially).
inline int chunks_for(size_t n, size_t grain) {
  if (n < 2 * grain) return 1;
  return int(std::min<size_t>(size_t(host_threads()), n / grain));
}

// f(begin, end, chunk) over [0, n) in nchunk contiguous chunks, chunk 0 on
// the calling thread.  A thread that cannot be started runs its chunk on
// the calling thread afterwards.  The chunks partition [0, n), so the
// outcome does not depend on the chunk count as long as f writes only its
// own items (or per-chunk outputs that the caller merges in chunk order).
// An exception thrown by f in any chunk (std::bad_alloc, say) is rethrown
// on the calling thread once every chunk has ended.
template <typename F>
void parallel_chunks(size_t n, int nchunk, F&& f) {
  if (nchunk <= 1) {
    f(size_t(0), n, 0);
    return;
  }
  const size_t per = (n + size_t(nchunk) - 1) / size_t(nchunk);
  std::vector<std::exception_ptr> err(static_cast<size_t>(nchunk));
  auto run = [&f, &err](size_t b, size_t e, int c) {
    try {
      f(b, e, c);
    } catch (...) {
      err[size_t(c)] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  std::vector<int> inline_chunks;
  for (int c = 1; c < nchunk; ++c) {
    const size_t b = size_t(c) * per, e = std::min(n, b + per);
    if (b >= e) continue;
    try {
      th.emplace_back(run, b, e, c);
    } catch (...) {
      inline_chunks.push_back(c);
    }
  }
  run(size_t(0), std::min(n, per), 0);
  for (auto& t : th) t.join();
  for (int c : inline_chunks) run(size_t(c) * per, std::min(n, size_t(c) * per + per), c);
  for (const std::exception_ptr& e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace bcm
