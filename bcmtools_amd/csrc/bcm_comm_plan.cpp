// Per-peer communication buffers of one rank: the layout of the reference's
// CommBuffer (src/CommBuffer.cpp:55-94) over the face lists that
// BlockManager builds for it (setFaceLists, setSeparateFaceLists and
// setSeparateLevelDiffFaceLists, src/BlockManager.cpp:471-650; the slots of
// setSendBufferPointers / setRecvBufferPointers, :671-711).  Host code
// only; the faces it returns drive bcm_comm_pack / bcm_comm_unpack.
//
// Rules restated here:
// - a rank's faces are visited block by block in ascending block index,
//   faces 0..5, and for a finer neighbour (level difference +1) subfaces
//   0..3; a face whose neighbour lives on another rank adds one slot to
//   that peer's send list and one to its receive list;
// - each receive list is then ordered as the sender filled its send list:
//   by neighbour block * 24 + opposite face * 4 + neighbour subface
//   (BlockManager::RecvFaceComp, include/BlockManager.h:134-148);
// - a slot takes its face's bytes rounded up to a multiple of 8; the
//   peers' buffers are kept per rank in ascending order (std::map);
// - slot sizes (VCUpdater's buffer sizes, include/Scalar3DUpdater.h): a
//   same-level face n^2 vc values; sent to a coarser neighbour (F2C) or
//   received from a finer one n^2 vc / 4; sent to a finer neighbour (C2F,
//   interpolated on the sender) or received from a coarser one n^2 vc.
#include <algorithm>
#include <cstdint>
#include <new>
#include <vector>

#include "bcm_hip.h"

namespace {

struct Slot {
  int32_t peer;
  int64_t key;  // receive-list order
  bcm_comm_face face;
};

}  // namespace

extern "C" int bcm_comm_plan(const bcm_block_info* blocks, int32_t nblocks, int32_t n, int32_t vc, int32_t dtype,
                             const int32_t* block_rank, int32_t rank, int32_t direction, int32_t axis,
                             int32_t level_diff, bcm_comm_peer* peers, int32_t max_peers, int32_t* npeers,
                             bcm_comm_face* faces, int64_t max_faces, int64_t* nfaces, int64_t* total_bytes) {
  if (nblocks < 0 || (nblocks > 0 && (!blocks || !block_rank)) || n < 2 || (n & 1) || vc < 1 || vc > n / 2 ||
      (dtype != BCM_F64 && dtype != BCM_F32) || rank < 0 ||
      (direction != BCM_COMM_SEND && direction != BCM_COMM_RECV) || axis < -1 || axis > 2 ||
      (level_diff != BCM_COMM_ANY_LEVEL && (level_diff < -1 || level_diff > 1)) || max_peers < 0 || max_faces < 0)
    return BCM_ERR_ARG;
  for (int32_t b = 0; b < nblocks; ++b) {
    if (block_rank[b] < 0) return BCM_ERR_ARG;
    for (int f = 0; f < BCM_NUM_FACE; ++f) {
      const int ld = blocks[b].nb_level_diff[f];
      if (ld < -1 || ld > 1) return BCM_ERR_ARG;
      for (int s = 0; s < 4; ++s)
        if (blocks[b].nb_id[f][s] < -1 || blocks[b].nb_id[f][s] >= nblocks) return BCM_ERR_ARG;
      if (ld == -1 && (blocks[b].nb_subface[f] < 0 || blocks[b].nb_subface[f] > 3)) return BCM_ERR_ARG;
    }
  }
  const int64_t esize = dtype == BCM_F64 ? 8 : 4;
  const int64_t full = int64_t(n) * n * vc;
  std::vector<Slot> slots;
  try {
    for (int32_t b = 0; b < nblocks; ++b) {
      if (block_rank[b] != rank) continue;
      const bcm_block_info& bi = blocks[b];
      for (int f = 0; f < BCM_NUM_FACE; ++f) {
        if (axis >= 0 && f / 2 != axis) continue;
        const int ld = bi.nb_level_diff[f];
        if (level_diff != BCM_COMM_ANY_LEVEL && ld != level_diff) continue;
        const int nsub = ld == 1 ? 4 : 1;
        for (int s = 0; s < nsub; ++s) {
          const int32_t nb = bi.nb_id[f][s];
          if (nb < 0 || block_rank[nb] == rank) continue;  // no neighbour, or a local copy
          Slot sl;
          sl.peer = block_rank[nb];
          // The neighbour subface is meaningful for a coarser neighbour
          // only; the reference's NeighborInfo holds 0 otherwise.
          sl.key = int64_t(nb) * (BCM_NUM_FACE * 4) + int64_t(f ^ 1) * 4 + (ld == -1 ? bi.nb_subface[f] : 0);
          sl.face.block = b;
          sl.face.face = f;
          sl.face.offset = 0;
          if (direction == BCM_COMM_SEND) {
            sl.face.kind = ld == 0 ? BCM_COMM_SAME : ld == -1 ? BCM_COMM_F2C : BCM_COMM_C2F;
            sl.face.subface = ld == 1 ? s : 0;
          } else {
            sl.face.kind = ld == 0 ? BCM_COMM_SAME : ld == -1 ? BCM_COMM_C2F : BCM_COMM_F2C;
            sl.face.subface = ld == 1 ? s : 0;
          }
          slots.push_back(sl);
        }
      }
    }
    // Per peer, ascending; send slots in visiting order, receive slots in
    // the sender's order.
    std::stable_sort(slots.begin(), slots.end(), [direction](const Slot& x, const Slot& y) {
      if (x.peer != y.peer) return x.peer < y.peer;
      return direction == BCM_COMM_RECV && x.key < y.key;
    });
  } catch (const std::bad_alloc&) {
    return BCM_ERR_NOMEM;
  }
  int32_t np = 0;
  for (size_t i = 0; i < slots.size(); ++i)
    if (i == 0 || slots[i].peer != slots[i - 1].peer) ++np;
  const bool write = peers != nullptr || faces != nullptr;
  if (write && (!peers || !faces || np > max_peers || int64_t(slots.size()) > max_faces)) return BCM_ERR_ARG;
  int64_t pos = 0;  // bytes
  int32_t p = -1;
  for (size_t i = 0; i < slots.size(); ++i) {
    const bcm_comm_face& fc = slots[i].face;
    if (i == 0 || slots[i].peer != slots[i - 1].peer) {
      ++p;
      if (write) {
        peers[p].rank = slots[i].peer;
        peers[p].nfaces = 0;
        peers[p].first = int64_t(i);
        peers[p].offset = pos;
        peers[p].bytes = 0;
      }
    }
    const int64_t count = fc.kind == BCM_COMM_F2C ? full / 4 : full;
    const int64_t slot = (count * esize + 7) / 8 * 8;
    if (write) {
      faces[i] = fc;
      faces[i].offset = pos / esize;
      peers[p].nfaces += 1;
      peers[p].bytes += slot;
    }
    pos += slot;
  }
  if (npeers) *npeers = np;
  if (nfaces) *nfaces = int64_t(slots.size());
  if (total_bytes) *total_bytes = pos;
  return BCM_OK;
}
