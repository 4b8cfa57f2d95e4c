// Vector3DUpdater<T>: the VCUpdater of a Vector3D<T>.  The reference
// declares only a private constructor (include/Vector3DUpdater.h:35-45), so
// vector fields cannot be exchanged there; here each face applies the
// scalar rules (copy, F2C, C2F) to every component on the GPU
// (bcm_vfield_ghost_fill, bcm_vfield_ghost_fill_face).  Buffer sizes are
// three times Scalar3DUpdater's.
#ifndef BCM_FACADE_VECTOR3DUPDATER_H
#define BCM_FACADE_VECTOR3DUPDATER_H

#include "VCUpdater.h"
#include "Vector3D.h"

template <typename T>
class Vector3DUpdater : public VCUpdater {
  Vector3D<T>* dataClass;
  Vector3D<T>* neighbor[NUM_FACE][NUM_SUBFACE];
  T* sendBuffer[NUM_FACE][NUM_SUBFACE];
  T* recvBuffer[NUM_FACE][NUM_SUBFACE];
  int nx, ny, nz, vc;

 public:
  Vector3DUpdater(const NeighborInfo* ni, const MPI::Comm& comm = MPI::COMM_WORLD)
      : VCUpdater(ni, comm), dataClass(0), nx(0), ny(0), nz(0), vc(0) {
    for (int f = 0; f < NUM_FACE; ++f)
      for (int s = 0; s < NUM_SUBFACE; ++s) {
        neighbor[f][s] = 0;
        sendBuffer[f][s] = recvBuffer[f][s] = 0;
      }
  }
  void setDataClass(DataClass* dc) {
    dataClass = dynamic_cast<Vector3D<T>*>(dc);
    nx = dc->getSizeX();
    ny = dc->getSizeY();
    nz = dc->getSizeZ();
    vc = dc->getVCSize();
  }
  void setNeighbor(Face face, Subface subface, DataClass* dc) {
    neighbor[face][subface] = dynamic_cast<Vector3D<T>*>(dc);
  }
  void clearNeighbor(Face face, Subface subface) { neighbor[face][subface] = 0; }
  PointerSetterBase* getSendBufferPointerSetter(Face face, Subface subface) {
    return new PointerSetter<T>(&sendBuffer[face][subface]);
  }
  PointerSetterBase* getRecvBufferPointerSetter(Face face, Subface subface) {
    return new PointerSetter<T>(&recvBuffer[face][subface]);
  }

 private:
  size_t getCommBufferSize(Face face) const {
    if (face == X_M || face == X_P) return 3 * size_t(ny) * nz * vc;
    if (face == Y_M || face == Y_P) return 3 * size_t(nz) * nx * vc;
    return 3 * size_t(nx) * ny * vc;
  }
  void copyFromNeighbor(Face face) { exchangeOnDevice(face, Subface(0)); }
  void copyFromNeighborF2C(Face face, Subface subface) { exchangeOnDevice(face, subface); }
  void copyFromNeighborC2F(Face face, Subface) { exchangeOnDevice(face, Subface(0)); }
  void copyToCommBuffer(Face) { needsSecondRank("Vector3DUpdater::copyToCommBuffer"); }
  void copyToCommBufferF2C(Face, Subface) { needsSecondRank("Vector3DUpdater::copyToCommBufferF2C"); }
  void copyToCommBufferC2F(Face, Subface) { needsSecondRank("Vector3DUpdater::copyToCommBufferC2F"); }
  void copyFromCommBuffer(Face) { needsSecondRank("Vector3DUpdater::copyFromCommBuffer"); }
  void copyFromCommBufferF2C(Face, Subface) { needsSecondRank("Vector3DUpdater::copyFromCommBufferF2C"); }
  void copyFromCommBufferC2F(Face, Subface) { needsSecondRank("Vector3DUpdater::copyFromCommBufferC2F"); }
  size_t getSendBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getSendBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getSendBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getRecvBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
};

#endif
