// BlockManager facade (include/BlockManager.h, src/BlockManager.cpp) over
// the C ABI of include/bcm_hip.h.
//
// Same calls as the reference: registerBlock / endRegisterBlock,
// setDataClass<D[,U[,T]]>(vc), prepareForVCUpdate(id, tag, method), every
// updateVC* / beginUpdateVC* / endUpdateVC* form, set/getCommunicator and
// printBlockLayoutInfo([file]).  Each data class registered with an updater
// (Scalar3D<T> / Scalar3DUpdater<T>, Vector3D<T> / Vector3DUpdater<T>)
// lives in one contiguous host arena (the reference's setDataClass<D,U,T>
// layout) mirrored by one device field.
//
// Residency:
//  * host-resident (default, drop-in): the host arena is authoritative;
//    updateVC* uploads it, fills the face ghosts on the GPU and downloads
//    it, so a driver that computes on getData() pointers runs unchanged.
//    The arena is page-locked (bcm_host_alloc), so these copies run at the
//    full PCIe rate.
//  * device-resident (setDeviceResident(id, true)): the device field is
//    authoritative, updateVC* is one kernel launch, and the driver calls
//    the device kernels (deviceField(id) with the bcm_* entry points);
//    syncToHost / syncToDevice move the data explicitly.
//
// One process drives one GPU: the communicator has one rank (include/bcm/
// mpi.h), getStartID() is 0, every neighbour is local.  beginUpdateVC*
// performs the whole exchange and endUpdateVC* returns immediately, which
// gives the reference's results because the exchange writes face ghosts
// only.  A per-face VCUpdater call (copyFromNeighbor(face, subface)) comes
// back here as exchangeFace and runs that face's rule on the device.
#ifndef BCM_FACADE_BLOCKMANAGER_H
#define BCM_FACADE_BLOCKMANAGER_H

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <ostream>
#include <vector>

#include "BCMTools.h"
#include "BlockBase.h"
#include "Scalar3D.h"
#include "Scalar3DUpdater.h"
#include "Vector3D.h"
#include "Vector3DUpdater.h"
#include "bcm_hip.h"
#include "mpi.h"

class VCUpdateMethod {
 public:
  enum Type { AtOnce, SeparateXYZ, SeparateXYZ_SeparateLevelDiff };
};

class BlockManager : public FaceExchanger {
  struct FieldRec {
    int vc = 0;
    int dtype = BCM_F64;
    int ncomp = 1;                 // values per cell: 1 (Scalar3D) or 3 (Vector3D)
    bcm_field_t dev = nullptr;     // scalar classes
    bcm_vfield_t vdev = nullptr;   // vector classes
    void* host = nullptr;
    size_t bytes = 0;
    size_t block_bytes = 0;
    bool updatable = false;
    bool prepared = false;
    bool device_resident = false;
    void (*free_host)(void*) = nullptr;
  };

  MPI::Intracomm comm;
  std::vector<BlockBase*> blockList;
  Vec3i size;
  int numBlock = 0;
  std::vector<bcm_block_info> table;
  std::map<int, bcm_mesh_t> meshes;  // by ghost width
  std::map<int, FieldRec> fields;    // by data-class id

  static void check(int rc, const char* where) {
    if (rc != BCM_OK) {
      std::cout << "*** error: " << where << " returned " << rc << std::endl;
      Exit(EX_FAILURE);
    }
  }

  template <typename T>
  static void free_array(void* p) {
    delete[] static_cast<T*>(p);
  }
  static void free_pinned(void* p) { bcm_host_free(p); }

  bcm_mesh_t mesh_for(int vc) {
    auto it = meshes.find(vc);
    if (it != meshes.end()) return it->second;
    if (size.x != size.y || size.x != size.z) {
      std::cout << "*** error: the device path needs cubic blocks" << std::endl;
      Exit(EX_FAILURE);
    }
    bcm_mesh_t m = nullptr;
    check(bcm_mesh_create(table.data(), numBlock, size.x, vc, &m), "bcm_mesh_create");
    meshes[vc] = m;
    return m;
  }

  FieldRec& rec(int id) {
    auto it = fields.find(id);
    if (it == fields.end()) {
      std::cout << "*** error: unknown data class id " << id << std::endl;
      Exit(EX_FAILURE);
    }
    return it->second;
  }

  // The updater type U of a data class D (Scalar3DUpdater<T> or
  // Vector3DUpdater<T>, or none).
  template <typename D, typename U, typename T>
  int add_contiguous(int vc, bool updatable) {
    const int nc = D::ncomp;
    const int nx0 = size.x + 2 * vc, ny0 = size.y + 2 * vc, nz0 = size.z + 2 * vc;
    const size_t bvol = size_t(nx0) * ny0 * nz0 * nc;
    // A class mirrored on the device gets a page-locked arena, so its
    // whole-field copies run at the full PCIe rate; a host-only class an
    // ordinary one.
    T* data = nullptr;
    if (updatable) {
      void* p = nullptr;
      check(bcm_host_alloc(sizeof(T) * bvol * size_t(numBlock), &p), "bcm_host_alloc");
      data = static_cast<T*>(p);
    } else {
      data = new T[bvol * numBlock]();
    }
    int id = -1;
    std::vector<VCUpdater*> updaters;
    for (int b = 0; b < numBlock; ++b) {
      BlockBase* block = blockList[b];
      D* dc = new D(block->getSize(), vc, &data[bvol * b]);
      if (updatable) {
        VCUpdater* u = new U(block->getNeighborInfo(), comm);
        dc->setUpdater(u);
        updaters.push_back(u);
      }
      const int id0 = block->setDataClass(dc);
      if (b == 0) id = id0;
      else if (id0 != id) {
        std::cout << "error: DataClass register inconsistency" << std::endl;
        Exit(EX_FAILURE);
      }
    }
    for (int b = 0; b < int(updaters.size()); ++b) updaters[b]->bindExchange(this, id, b);
    FieldRec r;
    r.vc = vc;
    r.dtype = sizeof(T) == 8 ? BCM_F64 : BCM_F32;
    r.ncomp = nc;
    r.host = data;
    r.block_bytes = bvol * sizeof(T);
    r.bytes = r.block_bytes * numBlock;
    r.updatable = updatable;
    r.free_host = updatable ? &free_pinned : &free_array<T>;
    if (updatable) {
      if (nc == 1) check(bcm_field_create(mesh_for(vc), r.dtype, &r.dev), "bcm_field_create");
      else check(bcm_vfield_create(mesh_for(vc), r.dtype, &r.vdev), "bcm_vfield_create");
    }
    fields[id] = r;
    return id;
  }

  void upload(FieldRec& r) {
    if (r.dev) check(bcm_field_upload(r.dev, r.host, r.bytes), "bcm_field_upload");
    else check(bcm_vfield_upload(r.vdev, r.host, r.bytes), "bcm_vfield_upload");
  }
  void download(FieldRec& r) {
    if (r.dev) check(bcm_field_download(r.dev, r.host, r.bytes), "bcm_field_download");
    else check(bcm_vfield_download(r.vdev, r.host, r.bytes), "bcm_vfield_download");
  }

  FieldRec& prepared(int id) {
    FieldRec& r = rec(id);
    if (!r.updatable || !r.prepared) {
      std::cout << "*** error: data class " << id << " is not prepared for VC update" << std::endl;
      Exit(EX_FAILURE);
    }
    return r;
  }

  void exchange(int id, int mask) {
    FieldRec& r = prepared(id);
    if (!r.device_resident) upload(r);
    if (r.dev) check(bcm_ghost_fill(r.dev, mask), "bcm_ghost_fill");
    else check(bcm_vfield_ghost_fill(r.vdev, mask), "bcm_vfield_ghost_fill");
    if (!r.device_resident) download(r);
  }

  static int axis_bit(int xyz) { return BCM_GF_X << xyz; }

  void commFace(int id, int block, Face face, Subface subface, int kind, void* buffer, size_t bytes, bool pack) {
    FieldRec& r = rec(id);
    if (!r.dev) {
      std::cout << "*** error: buffer rules need a scalar class with an updater" << std::endl;
      Exit(EX_FAILURE);
    }
    char* p = static_cast<char*>(r.host) + r.block_bytes * block;
    if (!r.device_resident)
      check(bcm_field_upload_blocks(r.dev, block, 1, p, r.block_bytes), "bcm_field_upload_blocks");
    bcm_comm_face fc;
    fc.block = block;
    fc.face = int(face);
    fc.subface = int(subface);
    fc.kind = kind;
    fc.offset = 0;
    if (pack) {
      check(bcm_comm_pack_host(r.dev, &fc, 1, buffer, bytes), "bcm_comm_pack_host");
    } else {
      check(bcm_comm_unpack_host(r.dev, &fc, 1, buffer, bytes), "bcm_comm_unpack_host");
      if (!r.device_resident)
        check(bcm_field_download_blocks(r.dev, block, 1, p, r.block_bytes), "bcm_field_download_blocks");
    }
  }

  // src/BlockManager.cpp:118-287 with one rank: every face is intra-node.
  void layoutReport(std::ostream& os) const {
    int lmin = INT_MAX, lmax = 0;
    for (int b = 0; b < numBlock; ++b) {
      lmin = std::min(lmin, blockList[b]->getLevel());
      lmax = std::max(lmax, blockList[b]->getLevel());
    }
    os << "Block Layout Information" << std::endl;
    os << "  Min level: " << lmin << std::endl << "  Max level: " << lmax << std::endl;
    os << "  Number of blocks" << std::endl;
    for (int l = lmin; l <= lmax; ++l) {
      int c = 0;
      for (int b = 0; b < numBlock; ++b) c += blockList[b]->getLevel() == l;
      os << "    L=" << l << ": " << c << std::endl;
    }
    os << "    total: " << numBlock << std::endl;
    os << "  Number of total blocks / node" << std::endl;
    os << "    ave: " << REAL_TYPE(numBlock) << std::endl << "    min: " << numBlock << std::endl
       << "    max: " << numBlock << std::endl << "     sd: " << 0 << std::endl;
    int faces[3] = {0, 0, 0};  // dLevel -1, 0, +1
    for (int b = 0; b < numBlock; ++b) {
      const NeighborInfo* ni = blockList[b]->getNeighborInfo();
      for (int f = 0; f < NUM_FACE; ++f) {
        if (!ni[f].exists()) continue;
        const int ld = ni[f].getLevelDifference();
        if (ld == 1) {
          for (int q = 0; q < NUM_SUBFACE; ++q) faces[2] += ni[f].getID(Subface(q)) >= 0;
        } else {
          faces[ld + 1] += 1;
        }
      }
    }
    os << "  Number of faces" << std::endl << "    intra-node " << std::endl;
    os << "        dLevel=-1: " << faces[0] << std::endl << "        dLevel= 0: " << faces[1] << std::endl
       << "        dLevel=+1: " << faces[2] << std::endl << "            total: " << faces[0] + faces[1] + faces[2]
       << std::endl;
    os << "    inter-node " << std::endl << "        dLevel=-1: 0" << std::endl << "        dLevel= 0: 0" << std::endl
       << "        dLevel=+1: 0" << std::endl << "            total: 0" << std::endl;
    os << "  Number of total inter-node faces / node" << std::endl << "    ave: 0" << std::endl << "    min: 0"
       << std::endl << "    max: 0" << std::endl << "     sd: 0" << std::endl;
  }

 public:
  BlockManager() : comm(MPI::COMM_WORLD) {}
  ~BlockManager() {
    for (auto& kv : fields) {
      if (kv.second.dev) bcm_field_destroy(kv.second.dev);
      if (kv.second.vdev) bcm_vfield_destroy(kv.second.vdev);
      // Host arenas stay valid while blocks may still reference them.
    }
    for (auto& kv : meshes) bcm_mesh_destroy(kv.second);
  }

  // include/BlockManager.h:159-162 (one rank).
  void setCommunicator(const MPI::Intracomm& c) { comm = c; }
  const MPI::Intracomm& getCommunicator() { return comm; }

  static BlockManager& getInstance() {
    static BlockManager instance;
    return instance;
  }

  void registerBlock(BlockBase* block) {
    blockList.push_back(block);
    numBlock++;
  }

  // src/BlockManager.cpp:36-59 plus the device block table.
  void endRegisterBlock() {
    if (numBlock == 0) {
      std::cout << "***error: number of blocks must be > 0" << std::endl;
      Exit(EX_FAILURE);
    }
    size = blockList[0]->getSize();
    table.assign(numBlock, bcm_block_info());
    for (int b = 0; b < numBlock; ++b) {
      BlockBase* blk = blockList[b];
      if (blk->getSize() != size) {
        std::cout << "***error: block size inconsistent" << std::endl;
        Exit(EX_FAILURE);
      }
      bcm_block_info& bi = table[b];
      std::memset(&bi, 0, sizeof(bi));
      bi.level = blk->getLevel();
      for (int a = 0; a < 3; ++a) {
        bi.origin[a] = blk->getOrigin()[a];
        bi.block_size[a] = blk->getBlockSize()[a];
      }
      const NeighborInfo* ni = blk->getNeighborInfo();
      for (int f = 0; f < NUM_FACE; ++f) {
        bi.nb_level_diff[f] = ni[f].getLevelDifference();
        for (int s = 0; s < NUM_SUBFACE; ++s) bi.nb_id[f][s] = ni[f].getID(Subface(s));
        bi.nb_subface[f] = ni[f].getNeighborSubface();
        bi.bc_type[f] = blk->boundaryType(f);
        bi.bc_id[f] = blk->boundaryID(f);
      }
    }
  }

  const Vec3i& getSize() const { return size; }
  int getStartID() const { return 0; }
  int getNumBlock() const { return numBlock; }
  BlockBase* getBlock(int localID) {
    if (localID < 0 || localID >= numBlock) {
      std::cout << "***error: illegal local block ID" << std::endl;
      Exit(EX_FAILURE);
    }
    return blockList[localID];
  }

  // include/BlockManager.h:195-196.
  void printBlockLayoutInfo() { layoutReport(std::cout); }
  void printBlockLayoutInfo(const char* filename) {
    std::ofstream ofs(filename, std::ios::out);
    if (!ofs) {
      std::cout << "*** error: cannot open " << filename << std::endl;
      Exit(EX_FAILURE);
    }
    layoutReport(ofs);
  }

  // include/BlockManager.h:202-283.  A class without an updater is host-only.
  template <typename D>
  int setDataClass(int vc) {
    return add_contiguous<D, Scalar3DUpdater<typename D::value_type>, typename D::value_type>(vc, false);
  }
  template <typename D, typename U>
  int setDataClass(int vc) {
    return add_contiguous<D, U, typename D::value_type>(vc, true);
  }
  template <typename D, typename U, typename T>
  int setDataClass(int vc) {
    return add_contiguous<D, U, T>(vc, true);
  }

  // FaceExchanger: one face of one block, for VCUpdater::copyFromNeighbor.
  // A host-resident class moves the blocks the rule reads and writes.
  void exchangeFace(int id, int block, Face face, Subface subface) {
    FieldRec& r = prepared(id);
    const NeighborInfo& ni = blockList[block]->getNeighborInfo()[face];
    const int src = ni.getLevelDifference() == 1 ? ni.getID(subface) : ni.getID();
    if (src < 0) {
      std::cout << "*** error: face " << int(face) << " of block " << block << " has no neighbour" << std::endl;
      Exit(EX_FAILURE);
    }
    char* host = static_cast<char*>(r.host);
    auto up = [&](int b) {
      char* p = host + r.block_bytes * b;
      if (r.dev) check(bcm_field_upload_blocks(r.dev, b, 1, p, r.block_bytes), "bcm_field_upload_blocks");
      else check(bcm_vfield_upload_blocks(r.vdev, b, 1, p, r.block_bytes), "bcm_vfield_upload_blocks");
    };
    if (!r.device_resident) {
      up(block);
      if (src != block) up(src);
    }
    if (r.dev) check(bcm_ghost_fill_face(r.dev, block, int(face), int(subface)), "bcm_ghost_fill_face");
    else check(bcm_vfield_ghost_fill_face(r.vdev, block, int(face), int(subface)), "bcm_vfield_ghost_fill_face");
    if (!r.device_resident) {
      char* p = host + r.block_bytes * block;
      if (r.dev) check(bcm_field_download_blocks(r.dev, block, 1, p, r.block_bytes), "bcm_field_download_blocks");
      else check(bcm_vfield_download_blocks(r.vdev, block, 1, p, r.block_bytes), "bcm_vfield_download_blocks");
    }
  }

  // FaceExchanger: one face of one block into or out of a host buffer, for
  // the VCUpdater buffer rules.  A host-resident class moves the block.
  void packFace(int id, int block, Face face, Subface subface, int kind, void* buffer, size_t bytes) {
    commFace(id, block, face, subface, kind, buffer, bytes, true);
  }
  void unpackFace(int id, int block, Face face, Subface subface, int kind, const void* buffer, size_t bytes) {
    commFace(id, block, face, subface, kind, const_cast<void*>(buffer), bytes, false);
  }

  void prepareForVCUpdate(int id, int /*tag*/, VCUpdateMethod::Type /*method*/) { rec(id).prepared = true; }
  void prepareForVCUpdate(int id, int tag, bool separate = false) {
    prepareForVCUpdate(id, tag, separate ? VCUpdateMethod::SeparateXYZ : VCUpdateMethod::AtOnce);
  }

  // include/BlockManager.h:311-726.
  void updateVC(int id) { exchange(id, BCM_GF_ALL); }
  void updateVC(int id, int xyz) { exchange(id, BCM_GF_ALL | axis_bit(xyz)); }
  void updateVC_Flat(int id, int xyz) { exchange(id, BCM_GF_FLAT | axis_bit(xyz)); }
  void updateVC_F2C(int id, int xyz) { exchange(id, BCM_GF_F2C | axis_bit(xyz)); }
  void updateVC_C2F(int id, int xyz) { exchange(id, BCM_GF_C2F | axis_bit(xyz)); }
  void updateVC_X(int id) { updateVC(id, 0); }
  void updateVC_Y(int id) { updateVC(id, 1); }
  void updateVC_Z(int id) { updateVC(id, 2); }
  void updateVC_X_Flat(int id) { updateVC_Flat(id, 0); }
  void updateVC_Y_Flat(int id) { updateVC_Flat(id, 1); }
  void updateVC_Z_Flat(int id) { updateVC_Flat(id, 2); }
  void updateVC_X_F2C(int id) { updateVC_F2C(id, 0); }
  void updateVC_Y_F2C(int id) { updateVC_F2C(id, 1); }
  void updateVC_Z_F2C(int id) { updateVC_F2C(id, 2); }
  void updateVC_X_C2F(int id) { updateVC_C2F(id, 0); }
  void updateVC_Y_C2F(int id) { updateVC_C2F(id, 1); }
  void updateVC_Z_C2F(int id) { updateVC_C2F(id, 2); }
  void beginUpdateVC(int id) { updateVC(id); }
  void beginUpdateVC(int id, int xyz) { updateVC(id, xyz); }
  void beginUpdateVC_X(int id) { updateVC(id, 0); }
  void beginUpdateVC_Y(int id) { updateVC(id, 1); }
  void beginUpdateVC_Z(int id) { updateVC(id, 2); }
  void beginUpdateVC_Flat(int id, int xyz) { updateVC_Flat(id, xyz); }
  void beginUpdateVC_F2C(int id, int xyz) { updateVC_F2C(id, xyz); }
  void beginUpdateVC_C2F(int id, int xyz) { updateVC_C2F(id, xyz); }
  void beginUpdateVC_X_Flat(int id) { updateVC_Flat(id, 0); }
  void beginUpdateVC_Y_Flat(int id) { updateVC_Flat(id, 1); }
  void beginUpdateVC_Z_Flat(int id) { updateVC_Flat(id, 2); }
  void beginUpdateVC_X_F2C(int id) { updateVC_F2C(id, 0); }
  void beginUpdateVC_Y_F2C(int id) { updateVC_F2C(id, 1); }
  void beginUpdateVC_Z_F2C(int id) { updateVC_F2C(id, 2); }
  void beginUpdateVC_X_C2F(int id) { updateVC_C2F(id, 0); }
  void beginUpdateVC_Y_C2F(int id) { updateVC_C2F(id, 1); }
  void beginUpdateVC_Z_C2F(int id) { updateVC_C2F(id, 2); }
  void endUpdateVC(int) {}
  void endUpdateVC(int, int) {}
  void endUpdateVC_X(int) {}
  void endUpdateVC_Y(int) {}
  void endUpdateVC_Z(int) {}
  void endUpdateVC_Flat(int, int) {}
  void endUpdateVC_F2C(int, int) {}
  void endUpdateVC_C2F(int, int) {}
  void endUpdateVC_X_Flat(int) {}
  void endUpdateVC_Y_Flat(int) {}
  void endUpdateVC_Z_Flat(int) {}
  void endUpdateVC_X_F2C(int) {}
  void endUpdateVC_Y_F2C(int) {}
  void endUpdateVC_Z_F2C(int) {}
  void endUpdateVC_X_C2F(int) {}
  void endUpdateVC_Y_C2F(int) {}
  void endUpdateVC_Z_C2F(int) {}

  // Device path.
  bcm_field_t deviceField(int id) { return rec(id).dev; }
  bcm_vfield_t deviceVectorField(int id) { return rec(id).vdev; }
  bcm_mesh_t deviceMesh(int id) { return mesh_for(rec(id).vc); }
  void setDeviceResident(int id, bool on) {
    FieldRec& r = rec(id);
    if (!r.dev && !r.vdev) {
      std::cout << "*** error: data class " << id << " has no device copy (no updater)" << std::endl;
      Exit(EX_FAILURE);
    }
    if (on && !r.device_resident) upload(r);
    if (!on && r.device_resident) download(r);
    r.device_resident = on;
  }
  void syncToHost(int id) { download(rec(id)); }
  void syncToDevice(int id) { upload(rec(id)); }
  // The local index of a registered block (-1 if it is not one).
  int indexOf(const BlockBase* block) const {
    for (int b = 0; b < numBlock; ++b)
      if (blockList[b] == block) return b;
    return -1;
  }
  const std::vector<bcm_block_info>& blockTable() const { return table; }

 private:
  BlockManager(const BlockManager&);
  BlockManager& operator=(const BlockManager&);
};

#endif
