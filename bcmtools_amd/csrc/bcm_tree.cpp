// Host octree builder: leaf-block table for the device mesh.
//
// Restates the reference's mesh construction (Utils/src/BCMOctree.cpp,
// Utils/include/{Pedigree,RootGrid,SimpleDivider,BlockFactory}.h) with a
// flat node array instead of pointer-linked Nodes, plus two additions the
// reference lacks: a shell-refinement divider and an explicit face 2:1
// balancer (the reference only aborts on one direction of imbalance,
// Utils/src/BCMOctree.cpp:462-467, and silently drops neighbours two
// levels coarser, :441-474).
#include "bcm_hip.h"
#include "bcm_host_par.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <random>
#include <system_error>
#include <thread>
#include <utility>
#include <vector>

namespace bcm {
std::atomic<int> g_host_threads{0};
}  // namespace bcm

namespace {

using bcm::chunks_for;

// Same-signature wrapper of bcm::parallel_chunks for the loops below.
template <typename F>
void parallel_chunks(size_t n, size_t /*grain*/, int nchunk, F&& f) {
  bcm::parallel_chunks(n, nchunk, std::forward<F>(f));
}

// Pedigree word: x16 | y16 | z16 | root12 | level4 (Utils/include/Pedigree.h:42-73).
inline uint64_t make_pedigree(unsigned level, unsigned x, unsigned y, unsigned z, unsigned root) {
  return (uint64_t(x) << 48) | (uint64_t(y) << 32) | (uint64_t(z) << 16) |
         (uint64_t(root & 0xfff) << 4) | uint64_t(level & 0xf);
}

struct TNode {
  uint32_t level, x, y, z, root;
  int32_t parent;  // index of the parent node; -1 for a root
  int32_t child0;  // index of the first of 8 consecutive children; -1 for a leaf
  int32_t id;      // block ID of an active leaf, else -1
  bool active;
};

// Hilbert child-visit order and orientation tables.  The block IDs of a
// Hilbert-ordered mesh must equal the reference's
// (Utils/src/BCMOctree.cpp:483-536), so the 24 orientation states are
// kept with the same numbering.
const int kHilbertOrder[24][8] = {
    {0, 1, 3, 2, 6, 7, 5, 4}, {0, 4, 6, 2, 3, 7, 5, 1}, {0, 1, 5, 4, 6, 7, 3, 2},
    {5, 1, 0, 4, 6, 2, 3, 7}, {3, 7, 6, 2, 0, 4, 5, 1}, {6, 7, 3, 2, 0, 1, 5, 4},
    {5, 1, 3, 7, 6, 2, 0, 4}, {0, 4, 5, 1, 3, 7, 6, 2}, {5, 4, 0, 1, 3, 2, 6, 7},
    {5, 4, 6, 7, 3, 2, 0, 1}, {0, 2, 3, 1, 5, 7, 6, 4}, {6, 4, 0, 2, 3, 1, 5, 7},
    {5, 7, 3, 1, 0, 2, 6, 4}, {3, 7, 5, 1, 0, 4, 6, 2}, {6, 4, 5, 7, 3, 1, 0, 2},
    {0, 2, 6, 4, 5, 7, 3, 1}, {6, 2, 0, 4, 5, 1, 3, 7}, {6, 2, 3, 7, 5, 1, 0, 4},
    {3, 2, 0, 1, 5, 4, 6, 7}, {6, 7, 5, 4, 0, 1, 3, 2}, {5, 7, 6, 4, 0, 2, 3, 1},
    {3, 2, 6, 7, 5, 4, 0, 1}, {3, 1, 0, 2, 6, 4, 5, 7}, {3, 1, 5, 7, 6, 4, 0, 2},
};
const int kHilbertTurn[24][8] = {
    {1, 2, 0, 3, 4, 0, 5, 6},       {0, 7, 1, 8, 5, 1, 4, 9},       {15, 0, 2, 22, 20, 2, 19, 23},
    {20, 6, 3, 23, 15, 3, 16, 22},  {22, 13, 4, 12, 11, 4, 1, 20},  {11, 19, 5, 20, 22, 5, 0, 12},
    {9, 3, 6, 2, 21, 6, 17, 0},     {10, 1, 7, 11, 12, 7, 13, 14},  {12, 9, 8, 14, 10, 8, 18, 11},
    {6, 8, 9, 7, 17, 9, 21, 1},     {7, 15, 10, 16, 13, 10, 12, 17}, {5, 14, 11, 9, 0, 11, 22, 8},
    {8, 20, 12, 19, 18, 12, 10, 5}, {18, 4, 13, 5, 8, 13, 7, 19},   {17, 11, 14, 1, 6, 14, 23, 7},
    {2, 10, 15, 18, 19, 15, 20, 21}, {19, 17, 16, 21, 2, 16, 3, 18}, {14, 16, 17, 15, 23, 17, 6, 10},
    {13, 21, 18, 17, 7, 18, 8, 16}, {16, 5, 19, 4, 3, 19, 2, 13},   {3, 12, 20, 13, 16, 20, 15, 4},
    {23, 18, 21, 10, 14, 21, 9, 15}, {4, 23, 22, 6, 1, 22, 11, 3},  {21, 22, 23, 0, 9, 23, 14, 2},
};

// Child slots touching face f of a node, in subface order
// (include/NeighborInfo.h:176-193): the children of my neighbour that lie
// against my face f.
const int kChildOnFace[6][4] = {
    {1, 3, 5, 7}, {0, 2, 4, 6}, {2, 6, 3, 7}, {0, 4, 1, 5}, {4, 5, 6, 7}, {0, 1, 2, 3},
};
// Subface of a child slot seen across face f (include/NeighborInfo.h:196-211).
const int kSubfaceOfChild[3][8] = {
    {0, 0, 1, 1, 2, 2, 3, 3},
    {0, 2, 0, 2, 1, 3, 1, 3},
    {0, 1, 2, 3, 0, 1, 2, 3},
};

enum NodeKind { BRANCH, LEAF_ACTIVE, LEAF_INACTIVE };

// A block record whose default construction leaves the bytes alone: the
// table is sized once and every record is written in full by the thread
// that owns it, so a serial zero fill would only add a pass.
struct BlockRec {
  bcm_block_info v;
  BlockRec() {}
};

}  // namespace

struct bcm_tree {
  bcm_tree_params p;
  bcm_divider_fn user_divide = nullptr;  // BCM_DIV_USER
  void* user = nullptr;
  std::vector<TNode> nodes;
  std::vector<int32_t> root_node;   // node index of each root
  std::vector<int32_t> leaf_nodes;  // node index of each block ID
  std::vector<BlockRec> blocks;

  int nroot() const { return p.root_n[0] * p.root_n[1] * p.root_n[2]; }
  int rix(int r) const { return r % p.root_n[0]; }
  int riy(int r) const { return (r / p.root_n[0]) % p.root_n[1]; }
  int riz(int r) const { return (r / p.root_n[0]) / p.root_n[1]; }

  // RootGrid::isOuterBoundary (Utils/include/RootGrid.h:188-208): by root
  // position only, periodic or not.
  bool root_outer(int r, int face) const {
    switch (face) {
      case BCM_X_M: return rix(r) == 0;
      case BCM_X_P: return rix(r) == p.root_n[0] - 1;
      case BCM_Y_M: return riy(r) == 0;
      case BCM_Y_P: return riy(r) == p.root_n[1] - 1;
      case BCM_Z_M: return riz(r) == 0;
      default: return riz(r) == p.root_n[2] - 1;
    }
  }
  // RootGrid::getNeighborRoot (Utils/include/RootGrid.h:118-172): -1 at a
  // non-periodic outer face.
  int root_neighbor(int r, int face) const {
    int i[3] = {rix(r), riy(r), riz(r)};
    int a = face / 2;
    i[a] += (face & 1) ? 1 : -1;
    if (i[a] < 0 || i[a] >= p.root_n[a]) {
      if (!p.periodic[a]) return -1;
      i[a] = (i[a] + p.root_n[a]) % p.root_n[a];
    }
    return i[0] + p.root_n[0] * (i[1] + p.root_n[1] * i[2]);
  }

  NodeKind classify(const TNode& nd) const {
    const int level = int(nd.level);
    switch (p.divider) {
      case BCM_DIV_USER: {  // the caller's Divider::operator() (Utils/include/Divider.h:40-46)
        const int k = user_divide(user, make_pedigree(nd.level, nd.x, nd.y, nd.z, nd.root));
        return k == BCM_NODE_BRANCH ? BRANCH : (k == BCM_NODE_LEAF_ACTIVE ? LEAF_ACTIVE : LEAF_INACTIVE);
      }
      case BCM_DIV_FLAT:  // FlatDivider (Utils/include/SimpleDivider.h:35-50)
        return level >= p.max_level ? LEAF_ACTIVE : BRANCH;
      case BCM_DIV_SIMPLE: {  // SimpleDivider, all six directions (:53-151)
        if (level < p.min_level) return BRANCH;
        if (level >= p.max_level) return LEAF_ACTIVE;
        const unsigned m = (1u << level) - 1;
        const int r = int(nd.root);
        if (nd.x == 0 && root_outer(r, BCM_X_M)) return BRANCH;
        if (nd.y == 0 && root_outer(r, BCM_Y_M)) return BRANCH;
        if (nd.z == 0 && root_outer(r, BCM_Z_M)) return BRANCH;
        if (nd.x == m && root_outer(r, BCM_X_P)) return BRANCH;
        if (nd.y == m && root_outer(r, BCM_Y_P)) return BRANCH;
        if (nd.z == m && root_outer(r, BCM_Z_P)) return BRANCH;
        return LEAF_ACTIVE;
      }
      default: {  // shell refinement: split a node whose box meets a shell
        if (level < p.min_level) return BRANCH;
        if (level >= p.max_level) return LEAF_ACTIVE;
        const double d = 1.0 / double(1u << level);
        const int r = int(nd.root);
        const double lo[3] = {p.root_origin[0] + (rix(r) + nd.x * d) * p.root_length,
                              p.root_origin[1] + (riy(r) + nd.y * d) * p.root_length,
                              p.root_origin[2] + (riz(r) + nd.z * d) * p.root_length};
        const double len = d * p.root_length;
        for (int s = 0; s < p.n_shells; ++s) {
          const double* sh = p.shells[s];
          double dmin2 = 0.0, dmax2 = 0.0;
          for (int a = 0; a < 3; ++a) {
            const double c = sh[a], a0 = lo[a], a1 = lo[a] + len;
            const double dn = c < a0 ? a0 - c : (c > a1 ? c - a1 : 0.0);
            const double dx = std::max(std::fabs(c - a0), std::fabs(c - a1));
            dmin2 += dn * dn;
            dmax2 += dx * dx;
          }
          if (!(dmin2 <= sh[4] * sh[4] && dmax2 >= sh[3] * sh[3])) continue;
          if (p.shell_mode == 1 && level == p.max_level - 1) {
            // Last split: node centre within shell_margin edges of the shell.
            double d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
              const double c = lo[a] + 0.5 * len - sh[a];
              d2 += c * c;
            }
            const double d = std::sqrt(d2);
            const double gap = d < sh[3] ? sh[3] - d : (d > sh[4] ? d - sh[4] : 0.0);
            if (gap <= p.shell_margin * len) return BRANCH;
            continue;
          }
          return BRANCH;
        }
        return LEAF_ACTIVE;
      }
    }
  }

  void split(int idx) {
    const TNode parent = nodes[idx];
    const int32_t c0 = int32_t(nodes.size());
    for (int c = 0; c < 8; ++c) {
      TNode ch;
      ch.level = parent.level + 1;
      ch.x = parent.x * 2 + (c & 1);
      ch.y = parent.y * 2 + ((c >> 1) & 1);
      ch.z = parent.z * 2 + ((c >> 2) & 1);
      ch.root = parent.root;
      ch.parent = idx;
      ch.child0 = -1;
      ch.id = -1;
      ch.active = true;
      nodes.push_back(ch);
    }
    nodes[idx].child0 = c0;
  }

  // BCMOctree::makeNode (Utils/src/BCMOctree.cpp:111-124), iteratively.
  void grow(int root_idx) {
    std::vector<int> stack(1, root_idx);
    while (!stack.empty()) {
      const int idx = stack.back();
      stack.pop_back();
      const NodeKind k = classify(nodes[idx]);
      if (k == LEAF_ACTIVE) continue;
      if (k == LEAF_INACTIVE) { nodes[idx].active = false; continue; }
      split(idx);
      const int c0 = nodes[idx].child0;
      for (int c = 7; c >= 0; --c) stack.push_back(c0 + c);
    }
  }

  // The same tree as grow() on every root, built level by level: the
  // nodes of one level are classified in parallel, then their children
  // appended.  Node indices differ from the depth-first build, which no
  // output depends on (block IDs follow the child links).  Used for the
  // built-in dividers; a caller's divider keeps grow()'s order of calls.
  void grow_levels() {
    std::vector<int> front(root_node.begin(), root_node.end());
    std::vector<uint8_t> kind;
    while (!front.empty()) {
      kind.assign(front.size(), 0);
      const int nc = chunks_for(front.size(), 2048);
      parallel_chunks(front.size(), 2048, nc, [&](size_t b, size_t e, int) {
        for (size_t i = b; i < e; ++i) kind[i] = uint8_t(classify(nodes[size_t(front[i])]));
      });
      size_t nbr = 0;
      for (size_t i = 0; i < front.size(); ++i) {
        if (kind[i] == BRANCH) ++nbr;
        else if (kind[i] == LEAF_INACTIVE) nodes[size_t(front[i])].active = false;
      }
      const size_t base = nodes.size();
      nodes.resize(base + 8 * nbr);
      std::vector<int> next(8 * nbr);
      size_t k = 0;
      for (size_t i = 0; i < front.size(); ++i) {
        if (kind[i] != BRANCH) continue;
        const TNode parent = nodes[size_t(front[i])];
        const int32_t c0 = int32_t(base + 8 * k);
        nodes[size_t(front[i])].child0 = c0;
        for (int c = 0; c < 8; ++c) {
          TNode& ch = nodes[size_t(c0 + c)];
          ch.level = parent.level + 1;
          ch.x = parent.x * 2 + (c & 1);
          ch.y = parent.y * 2 + ((c >> 1) & 1);
          ch.z = parent.z * 2 + ((c >> 2) & 1);
          ch.root = parent.root;
          ch.parent = front[i];
          ch.child0 = -1;
          ch.id = -1;
          ch.active = true;
          next[8 * k + size_t(c)] = c0 + c;
        }
        ++k;
      }
      front.swap(next);
    }
  }

  // BCMOctree::findNeighborNode (Utils/src/BCMOctree.cpp:277-364): the
  // deepest node at or above my level covering my face neighbour; -1 at
  // a non-periodic outer face.  Inside a root the search climbs to the
  // lowest ancestor that also covers the neighbour cell and descends from
  // there (the reference descends from the root; the node reached is the
  // same, since the root's path to it passes through that ancestor).
  int find_neighbor(int idx, int face) const {
    const TNode& nd = nodes[size_t(idx)];
    const int level = int(nd.level);
    const int m = (1 << level) - 1;
    int c[3] = {int(nd.x), int(nd.y), int(nd.z)};
    const int a = face / 2;
    const int own = c[a];
    c[a] += (face & 1) ? 1 : -1;
    int cur, from;
    if (c[a] < 0 || c[a] > m) {
      const int root = root_neighbor(int(nd.root), face);
      if (root < 0) return -1;
      c[a] = (face & 1) ? 0 : m;
      cur = root_node[size_t(root)];
      from = 1;
    } else {
      cur = idx;
      int sh = 0;
      while ((c[a] >> sh) != (own >> sh)) {
        cur = nodes[size_t(cur)].parent;
        ++sh;
      }
      from = level - sh + 1;
    }
    for (int l = from; l <= level; ++l) {
      if (nodes[size_t(cur)].child0 < 0) break;
      const int sh = level - l;
      const int ijk = ((c[0] >> sh) & 1) + 2 * ((c[1] >> sh) & 1) + 4 * ((c[2] >> sh) & 1);
      cur = nodes[size_t(cur)].child0 + ijk;
    }
    return cur;
  }

  // Split leaves until every face neighbour of an active leaf is at most
  // one level coarser.  Splitting only ever makes neighbours finer, so a
  // leaf that passed a check keeps passing it: after the first pass over
  // every leaf, a pass checks only the leaves that failed the previous one
  // (their split neighbour may still be two levels coarser) and the
  // children just created.  The splits reach the same fixed point as
  // repeated passes over all leaves.
  void balance() {
    std::vector<int> check;
    for (int idx = 0; idx < int(nodes.size()); ++idx) check.push_back(idx);
    for (;;) {
      const int nc = chunks_for(check.size(), 4096);
      std::vector<std::vector<int>> split_part(static_cast<size_t>(nc)), fail_part(static_cast<size_t>(nc));
      parallel_chunks(check.size(), 4096, nc, [&](size_t b, size_t e, int c) {
        std::vector<int>& sp = split_part[size_t(c)];
        std::vector<int>& fl = fail_part[size_t(c)];
        for (size_t i = b; i < e; ++i) {
          const int idx = check[i];
          const TNode& nd = nodes[size_t(idx)];
          if (nd.child0 >= 0 || !nd.active || nd.level < 2) continue;
          bool failed = false;
          for (int f = 0; f < 6; ++f) {
            const int nb = find_neighbor(idx, f);
            if (nb < 0) continue;
            const TNode& nn = nodes[size_t(nb)];
            if (nn.child0 < 0 && int(nn.level) + 1 < int(nd.level)) {
              sp.push_back(nb);
              failed = true;
            }
          }
          if (failed) fl.push_back(idx);
        }
      });
      std::vector<int> to_split;
      check.clear();
      for (int c = 0; c < nc; ++c) {
        to_split.insert(to_split.end(), split_part[size_t(c)].begin(), split_part[size_t(c)].end());
        check.insert(check.end(), fail_part[size_t(c)].begin(), fail_part[size_t(c)].end());
      }
      if (to_split.empty()) return;
      std::sort(to_split.begin(), to_split.end());
      to_split.erase(std::unique(to_split.begin(), to_split.end()), to_split.end());
      nodes.reserve(nodes.size() + 8 * to_split.size());
      for (int idx : to_split) {
        if (nodes[size_t(idx)].child0 < 0) split(idx);
        const int c0 = nodes[size_t(idx)].child0;
        if (!nodes[size_t(idx)].active)
          for (int c = 0; c < 8; ++c) nodes[size_t(c0 + c)].active = false;
        for (int c = 0; c < 8; ++c) check.push_back(c0 + c);
      }
    }
  }

  void pick_z(int idx) {
    std::vector<int> stack(1, idx);
    while (!stack.empty()) {
      const int i = stack.back();
      stack.pop_back();
      if (nodes[i].child0 < 0) {
        if (nodes[i].active) {
          nodes[i].id = int32_t(leaf_nodes.size());
          leaf_nodes.push_back(i);
        }
        continue;
      }
      for (int c = 7; c >= 0; --c) stack.push_back(nodes[i].child0 + c);
    }
  }

  void pick_hilbert(int idx, int orient) {
    if (nodes[idx].child0 < 0) {
      if (nodes[idx].active) {
        nodes[idx].id = int32_t(leaf_nodes.size());
        leaf_nodes.push_back(idx);
      }
      return;
    }
    for (int i = 0; i < 8; ++i)
      pick_hilbert(nodes[idx].child0 + kHilbertOrder[orient][i], kHilbertTurn[orient][i]);
  }

  // BCMOctree::checkOnOuterBoundary (Utils/src/BCMOctree.cpp:366-396).
  bool on_outer(const TNode& nd, int face) const {
    const unsigned m = (1u << nd.level) - 1;
    const unsigned c[3] = {nd.x, nd.y, nd.z};
    const unsigned want = (face & 1) ? m : 0u;
    return c[face / 2] == want && root_outer(int(nd.root), face);
  }

  // One record per leaf; the leaves are independent, so chunks of them
  // are filled in parallel.
  int make_blocks() {
    blocks.resize(leaf_nodes.size());
    std::atomic<int> status{BCM_OK};
    const int nc = chunks_for(leaf_nodes.size(), 2048);
    parallel_chunks(leaf_nodes.size(), 2048, nc, [&](size_t b0, size_t b1, int) {
      if (make_block_range(b0, b1) != BCM_OK) status.store(BCM_ERR_BALANCE);
    });
    return status.load();
  }

  int make_block_range(size_t b0, size_t b1) {
    for (size_t b = b0; b < b1; ++b) {
      const TNode& nd = nodes[leaf_nodes[b]];
      bcm_block_info& bi = blocks[b].v;
      std::memset(&bi, 0, sizeof(bi));
      bi.pedigree = make_pedigree(nd.level, nd.x, nd.y, nd.z, nd.root);
      bi.level = int32_t(nd.level);
      bi.root_id = int32_t(nd.root);
      // BCMOctree::getOrigin (:398-410) and BlockFactory::makeBlock (:89-96):
      // (ix + x/ub) * rootLength + rootOrigin, size (1/ub) * rootLength.
      const int ub = 1 << nd.level;
      const int r = int(nd.root);
      const double o[3] = {rix(r) + double(nd.x) / ub, riy(r) + double(nd.y) / ub,
                           riz(r) + double(nd.z) / ub};
      for (int a = 0; a < 3; ++a) {
        bi.origin[a] = o[a] * p.root_length + p.root_origin[a];
        bi.block_size[a] = (1.0 / ub) * p.root_length;
      }
      const int child_id = int((nd.x & 1) + 2 * (nd.y & 1) + 4 * (nd.z & 1));
      for (int f = 0; f < 6; ++f) {
        bi.nb_level_diff[f] = 0;
        for (int s = 0; s < 4; ++s) bi.nb_id[f][s] = -1;
        bi.nb_subface[f] = 0;
        bi.bc_type[f] = BCM_BC_INNER;
        bi.bc_id[f] = -1;
        // BCMOctree::makeNeighborInfo (Utils/src/BCMOctree.cpp:423-478).
        const int nb = find_neighbor(leaf_nodes[b], f);
        if (nb < 0) continue;
        const TNode& nn = nodes[nb];
        const int diff = int(nn.level) - int(nd.level);
        if (diff == -1) {
          if (!nn.active) continue;
          bi.nb_level_diff[f] = -1;
          bi.nb_id[f][0] = nn.id;
          bi.nb_subface[f] = kSubfaceOfChild[f / 2][child_id];
        } else if (diff == 0) {
          if (nn.child0 < 0) {
            if (!nn.active) continue;
            bi.nb_id[f][0] = nn.id;
          } else {
            bi.nb_level_diff[f] = 1;
            for (int s = 0; s < 4; ++s) {
              const TNode& ch = nodes[nn.child0 + kChildOnFace[f][s]];
              if (ch.child0 >= 0) return BCM_ERR_BALANCE;
              if (!ch.active) continue;
              bi.nb_id[f][s] = ch.id;
            }
          }
        }
        // diff <= -2: no neighbour, as in the reference.
      }
      // Boundary setters (examples/sor2/BoundaryConditionSetter.h,
      // examples/sor1/BoundaryConditionSetter.h:34-89).
      for (int f = 0; f < 6; ++f) {
        if (!on_outer(nd, f)) continue;
        if (p.bc_setter == BCM_BCSET_ALL_DIRICHLET) {
          bi.bc_type[f] = BCM_BC_DIRICHLET;
          bi.bc_id[f] = 0;
        } else if (p.bc_setter == BCM_BCSET_PER_FACE) {
          bi.bc_type[f] = p.face_bc_type[f];
          bi.bc_id[f] = f;
        } else {
          bi.bc_type[f] = BCM_BC_PERIODIC;
          const int axis = p.bc_setter - BCM_BCSET_SOR1_X;
          if (f / 2 == axis) {
            bi.bc_type[f] = BCM_BC_DIRICHLET;
            bi.bc_id[f] = f & 1;
          }
        }
      }
    }
    return BCM_OK;
  }
};

// Leaves in random order (BCMOctree::randomShuffle, Utils/src/BCMOctree.cpp:
// 400-404, which calls std::random_shuffle and so is not reproducible across
// platforms): a shuffle of the Z order by a fixed-seed Mersenne twister.
static void shuffle_leaves(bcm_tree* t) {
  std::mt19937 rng(20111213u);
  std::shuffle(t->leaf_nodes.begin(), t->leaf_nodes.end(), rng);
  for (size_t b = 0; b < t->leaf_nodes.size(); ++b) t->nodes[t->leaf_nodes[b]].id = int32_t(b);
}

static int tree_build(const bcm_tree_params* params, bcm_divider_fn divide, void* user, bcm_tree_t* out) {
  if (!params || !out) return BCM_ERR_ARG;
  *out = nullptr;
  const bcm_tree_params& p = *params;
  if ((p.divider == BCM_DIV_USER) != (divide != nullptr)) return BCM_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (p.root_n[a] < 1) return BCM_ERR_ARG;
  if (p.root_n[0] * p.root_n[1] * p.root_n[2] > 0xfff + 1) return BCM_ERR_ARG;
  if (p.max_level < 0 || p.max_level > 15 || p.min_level < 0 || p.min_level > p.max_level)
    return BCM_ERR_ARG;
  if (p.divider < BCM_DIV_FLAT || p.divider > BCM_DIV_USER) return BCM_ERR_ARG;
  if (p.divider == BCM_DIV_SHELL && (p.n_shells < 1 || p.n_shells > BCM_MAX_SHELLS)) return BCM_ERR_ARG;
  if (p.ordering != BCM_ORDER_Z && p.ordering != BCM_ORDER_HILBERT && p.ordering != BCM_ORDER_RANDOM)
    return BCM_ERR_ARG;
  if (p.bc_setter < BCM_BCSET_ALL_DIRICHLET || p.bc_setter > BCM_BCSET_PER_FACE) return BCM_ERR_ARG;
  if (p.bc_setter == BCM_BCSET_PER_FACE)
    for (int f = 0; f < 6; ++f)
      if (p.face_bc_type[f] != BCM_BC_DIRICHLET && p.face_bc_type[f] != BCM_BC_NEUMANN) return BCM_ERR_ARG;
  if (p.shell_mode < 0 || p.shell_mode > 1 || !(p.shell_margin >= 0.0)) return BCM_ERR_ARG;
  if (!(p.root_length > 0.0)) return BCM_ERR_ARG;
  bcm_tree* t = new (std::nothrow) bcm_tree();
  if (!t) return BCM_ERR_NOMEM;
  try {
    t->p = p;
    t->user_divide = divide;
    t->user = user;
    const int nr = t->nroot();
    t->root_node.resize(nr);
    for (int r = 0; r < nr; ++r) {
      TNode root;
      root.level = root.x = root.y = root.z = 0;
      root.root = uint32_t(r);
      root.parent = -1;
      root.child0 = -1;
      root.id = -1;
      root.active = true;
      t->root_node[r] = int32_t(t->nodes.size());
      t->nodes.push_back(root);
      if (divide) t->grow(t->root_node[r]);
    }
    if (!divide) t->grow_levels();
    if (p.balance) t->balance();
    for (int r = 0; r < nr; ++r) {
      if (p.ordering == BCM_ORDER_HILBERT) t->pick_hilbert(t->root_node[r], 0);
      else t->pick_z(t->root_node[r]);
    }
    if (p.ordering == BCM_ORDER_RANDOM) shuffle_leaves(t);
    const int rc = t->make_blocks();
    if (rc != BCM_OK) {
      delete t;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete t;
    return BCM_ERR_NOMEM;
  }
  *out = t;
  return BCM_OK;
}

extern "C" int bcm_tree_build(const bcm_tree_params* params, bcm_tree_t* out) {
  return tree_build(params, nullptr, nullptr, out);
}

extern "C" int bcm_tree_build_user(const bcm_tree_params* params, bcm_divider_fn divide, void* user,
                                   bcm_tree_t* out) {
  if (!divide) return BCM_ERR_ARG;
  return tree_build(params, divide, user, out);
}

extern "C" int bcm_tree_set_threads(int nthreads) {
  if (nthreads < 0) return BCM_ERR_ARG;
  bcm::g_host_threads.store(nthreads);
  return BCM_OK;
}

extern "C" int bcm_tree_num_blocks(bcm_tree_t tree) {
  if (!tree) return -BCM_ERR_ARG;
  return int(tree->blocks.size());
}

extern "C" int bcm_tree_get_blocks(bcm_tree_t tree, bcm_block_info* out, int capacity) {
  if (!tree || !out || capacity < int(tree->blocks.size())) return BCM_ERR_ARG;
  static_assert(sizeof(BlockRec) == sizeof(bcm_block_info), "records are copied as one array");
  std::memcpy(out, tree->blocks.data(), tree->blocks.size() * sizeof(bcm_block_info));
  return BCM_OK;
}

extern "C" int bcm_tree_destroy(bcm_tree_t tree) {
  delete tree;
  return BCM_OK;
}
