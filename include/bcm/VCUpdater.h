// VCUpdater (include/VCUpdater.h:53-247): the ghost ("virtual cell")
// exchange of one data class on one block, face by face.
//
// The public dispatchers pick the rule by level difference, as the
// reference does: same level -> copyFromNeighbor(face), finer neighbour ->
// ...F2C(face, subface), coarser neighbour -> ...C2F(face, my subface on
// it).  A concrete updater (Scalar3DUpdater, Vector3DUpdater) implements
// the private rules.  In this build the rules run on the GPU through the
// BlockManager that registered the data class (bindExchange): a local rule
// fills the face from the neighbour's block, a buffer rule packs the face
// into, or unpacks it from, the host buffer a communication layer
// attached (getSend/RecvBufferPointerSetter) in the reference's order.
//
// BlockManager::updateVC* never goes through here: it exchanges every face
// of every block in one device launch (bcm_ghost_fill).
#ifndef BCM_FACADE_VCUPDATER_H
#define BCM_FACADE_VCUPDATER_H

#include <cstddef>
#include <iostream>

#include "BCMTools.h"
#include "DataClass.h"
#include "NeighborInfo.h"
#include "PointerSetter.h"
#include "mpi.h"

// The facade's hook from an updater to its BlockManager: exchange the
// ghosts of one face of one block of one data class on the device.
class FaceExchanger {
 public:
  virtual ~FaceExchanger() {}
  virtual void exchangeFace(int dataClassID, int blockID, Face face, Subface subface) = 0;
  // One face into / out of a host communication buffer (bcm_comm_pack_host,
  // bcm_comm_unpack_host; kind BCM_COMM_SAME, _F2C or _C2F).
  virtual void packFace(int dataClassID, int blockID, Face face, Subface subface, int kind, void* buffer,
                        size_t bytes) = 0;
  virtual void unpackFace(int dataClassID, int blockID, Face face, Subface subface, int kind, const void* buffer,
                          size_t bytes) = 0;
};

class VCUpdater {
  const NeighborInfo* neighborInfo;
  const MPI::Comm& comm;
  int myrank;
  FaceExchanger* exchanger;
  int dataClassID;
  int blockID;

  static void fail(const char* what) {
    std::cout << "*** error: " << what << std::endl;
    Exit(EX_FAILURE);
  }

 public:
  VCUpdater(const NeighborInfo* ni, const MPI::Comm& comm_ = MPI::COMM_WORLD)
      : neighborInfo(ni), comm(comm_), myrank(comm_.Get_rank()), exchanger(0), dataClassID(-1), blockID(-1) {}
  virtual ~VCUpdater() {}

  virtual void setDataClass(DataClass* dataClass) = 0;
  const NeighborInfo* getNeighborInfo() const { return neighborInfo; }

  // Set by BlockManager::setDataClass: where the local rules run.
  void bindExchange(FaceExchanger* ex, int dataClass, int block) {
    exchanger = ex;
    dataClassID = dataClass;
    blockID = block;
  }

  // include/VCUpdater.h:65-82.
  void copyFromNeighbor(Face face, Subface subface) {
    const int d = neighborInfo[face].getLevelDifference();
    if (d == 0) copyFromNeighbor(face);
    else if (d == 1) copyFromNeighborF2C(face, subface);
    else if (d == -1) copyFromNeighborC2F(face, neighborInfo[face].getNeighborSubface());
    else fail("VCUpdater::copyFromNeighbor: level difference out of range");
  }
  // include/VCUpdater.h:85-103.
  void copyToCommBuffer(Face face, Subface subface) {
    const int d = neighborInfo[face].getLevelDifference();
    if (d == 0) copyToCommBuffer(face);
    else if (d == 1) copyToCommBufferC2F(face, subface);
    else if (d == -1) copyToCommBufferF2C(face, neighborInfo[face].getNeighborSubface());
    else fail("VCUpdater::copyToCommBuffer: level difference out of range");
  }
  // include/VCUpdater.h:106-124.
  void copyFromCommBuffer(Face face, Subface subface) {
    const int d = neighborInfo[face].getLevelDifference();
    if (d == 0) copyFromCommBuffer(face);
    else if (d == 1) copyFromCommBufferF2C(face, subface);
    else if (d == -1) copyFromCommBufferC2F(face, neighborInfo[face].getNeighborSubface());
    else fail("VCUpdater::copyFromCommBuffer: level difference out of range");
  }

  void setNeighbor(Face face, DataClass* dataClass) { setNeighbor(face, Subface(0), dataClass); }
  virtual void setNeighbor(Face face, Subface subface, DataClass* dataClass) = 0;
  void clearNeighbor() {
    for (int f = 0; f < NUM_FACE; ++f)
      for (int s = 0; s < NUM_SUBFACE; ++s) clearNeighbor(Face(f), Subface(s));
  }
  virtual void clearNeighbor(Face face, Subface subface) = 0;

  // include/VCUpdater.h:149-190.
  size_t getSendBufferByteSize(Face face, Subface subface) const {
    const int d = neighborInfo[face].getLevelDifference();
    if (d == 1) return getSendBufferByteSizeC2F(face, subface);
    if (d == -1) return getSendBufferByteSizeF2C(face, neighborInfo[face].getNeighborSubface());
    return getSendBufferByteSize(face);
  }
  size_t getRecvBufferByteSize(Face face, Subface subface) const {
    const int d = neighborInfo[face].getLevelDifference();
    if (d == 1) return getRecvBufferByteSizeF2C(face, subface);
    if (d == -1) return getRecvBufferByteSizeC2F(face, neighborInfo[face].getNeighborSubface());
    return getRecvBufferByteSize(face);
  }
  virtual PointerSetterBase* getSendBufferPointerSetter(Face face, Subface subface) = 0;
  virtual PointerSetterBase* getRecvBufferPointerSetter(Face face, Subface subface) = 0;

 protected:
  // A local rule: the face's ghosts on the device, through the BlockManager.
  void exchangeOnDevice(Face face, Subface subface) {
    if (!exchanger) fail("VCUpdater: the data class was not registered through BlockManager::setDataClass");
    exchanger->exchangeFace(dataClassID, blockID, face, subface);
  }
  // A buffer rule: the face to or from a host buffer, on the device.
  void packOnDevice(Face face, Subface subface, int kind, void* buffer, size_t bytes) {
    if (!exchanger) fail("VCUpdater: the data class was not registered through BlockManager::setDataClass");
    exchanger->packFace(dataClassID, blockID, face, subface, kind, buffer, bytes);
  }
  void unpackOnDevice(Face face, Subface subface, int kind, const void* buffer, size_t bytes) {
    if (!exchanger) fail("VCUpdater: the data class was not registered through BlockManager::setDataClass");
    exchanger->unpackFace(dataClassID, blockID, face, subface, kind, buffer, bytes);
  }
  // Where a subface starts on the face's first and second tangential axis
  // (1 = from the midpoint), for user updaters that index subfaces
  // themselves (include/VCUpdater.h:251-273).
  static int subfaceOrigin0(Subface subface) {
    if (subface < SF_00 || subface > SF_11) Exit(EX_FAILURE);
    return int(subface) & 1;
  }
  static int subfaceOrigin1(Subface subface) {
    if (subface < SF_00 || subface > SF_11) Exit(EX_FAILURE);
    return int(subface) >> 1;
  }
  // A buffer rule: a second rank's data.
  static void needsSecondRank(const char* what) {
    std::cout << "*** error: " << what << " moves data between MPI ranks; this build runs one rank on one GPU"
              << std::endl;
    Exit(EX_FAILURE);
  }

 private:
  virtual void copyFromNeighbor(Face face) = 0;
  virtual void copyFromNeighborF2C(Face face, Subface subface) = 0;
  virtual void copyFromNeighborC2F(Face face, Subface subface) = 0;
  virtual void copyToCommBuffer(Face face) = 0;
  virtual void copyToCommBufferF2C(Face face, Subface subface) = 0;
  virtual void copyToCommBufferC2F(Face face, Subface subface) = 0;
  virtual void copyFromCommBuffer(Face face) = 0;
  virtual void copyFromCommBufferF2C(Face face, Subface subface) = 0;
  virtual void copyFromCommBufferC2F(Face face, Subface subface) = 0;
  virtual size_t getSendBufferByteSize(Face face) const = 0;
  virtual size_t getSendBufferByteSizeF2C(Face face, Subface subface) const = 0;
  virtual size_t getSendBufferByteSizeC2F(Face face, Subface subface) const = 0;
  virtual size_t getRecvBufferByteSize(Face face) const = 0;
  virtual size_t getRecvBufferByteSizeF2C(Face face, Subface subface) const = 0;
  virtual size_t getRecvBufferByteSizeC2F(Face face, Subface subface) const = 0;
};

inline UpdatableDataClass::~UpdatableDataClass() { delete updater; }
inline void UpdatableDataClass::setUpdater(VCUpdater* u) {
  updater = u;
  updater->setDataClass(this);
}

#endif
