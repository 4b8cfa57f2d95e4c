// Scalar3DUpdater<T> (include/Scalar3DUpdater.h:64-904): the VCUpdater of a
// Scalar3D<T>.  Its local rules -- same-level copy, fine-to-coarse average
// and coarse-to-fine trilinear interpolation of one face -- run on the GPU
// (bcm_ghost_fill_face, the kernels of the whole-mesh exchange); the buffer
// sizes follow include/Scalar3DUpdater.h:84-177.
#ifndef BCM_FACADE_SCALAR3DUPDATER_H
#define BCM_FACADE_SCALAR3DUPDATER_H

#include "Scalar3D.h"
#include "VCUpdater.h"

template <typename T>
class Scalar3DUpdater : public VCUpdater {
  Scalar3D<T>* dataClass;
  Scalar3D<T>* neighbor[NUM_FACE][NUM_SUBFACE];
  T* sendBuffer[NUM_FACE][NUM_SUBFACE];
  T* recvBuffer[NUM_FACE][NUM_SUBFACE];
  int nx, ny, nz, vc;

 public:
  Scalar3DUpdater(const NeighborInfo* ni, const MPI::Comm& comm = MPI::COMM_WORLD)
      : VCUpdater(ni, comm), dataClass(0), nx(0), ny(0), nz(0), vc(0) {
    for (int f = 0; f < NUM_FACE; ++f)
      for (int s = 0; s < NUM_SUBFACE; ++s) {
        neighbor[f][s] = 0;
        sendBuffer[f][s] = recvBuffer[f][s] = 0;
      }
  }

  void setDataClass(DataClass* dc) {
    dataClass = dynamic_cast<Scalar3D<T>*>(dc);
    nx = dc->getSizeX();
    ny = dc->getSizeY();
    nz = dc->getSizeZ();
    vc = dc->getVCSize();
  }
  void setNeighbor(Face face, Subface subface, DataClass* dc) {
    neighbor[face][subface] = dynamic_cast<Scalar3D<T>*>(dc);
  }
  void clearNeighbor(Face face, Subface subface) { neighbor[face][subface] = 0; }
  PointerSetterBase* getSendBufferPointerSetter(Face face, Subface subface) {
    return new PointerSetter<T>(&sendBuffer[face][subface]);
  }
  PointerSetterBase* getRecvBufferPointerSetter(Face face, Subface subface) {
    return new PointerSetter<T>(&recvBuffer[face][subface]);
  }

 private:
  // Cells of one face's ghost layer (include/Scalar3DUpdater.h:162-177).
  size_t getCommBufferSize(Face face) const {
    if (face == X_M || face == X_P) return size_t(ny) * nz * vc;
    if (face == Y_M || face == Y_P) return size_t(nz) * nx * vc;
    return size_t(nx) * ny * vc;
  }
  void copyFromNeighbor(Face face) { exchangeOnDevice(face, Subface(0)); }
  void copyFromNeighborF2C(Face face, Subface subface) { exchangeOnDevice(face, subface); }
  void copyFromNeighborC2F(Face face, Subface) { exchangeOnDevice(face, Subface(0)); }
  void copyToCommBuffer(Face) { needsSecondRank("Scalar3DUpdater::copyToCommBuffer"); }
  void copyToCommBufferF2C(Face, Subface) { needsSecondRank("Scalar3DUpdater::copyToCommBufferF2C"); }
  void copyToCommBufferC2F(Face, Subface) { needsSecondRank("Scalar3DUpdater::copyToCommBufferC2F"); }
  void copyFromCommBuffer(Face) { needsSecondRank("Scalar3DUpdater::copyFromCommBuffer"); }
  void copyFromCommBufferF2C(Face, Subface) { needsSecondRank("Scalar3DUpdater::copyFromCommBufferF2C"); }
  void copyFromCommBufferC2F(Face, Subface) { needsSecondRank("Scalar3DUpdater::copyFromCommBufferC2F"); }
  size_t getSendBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getSendBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getSendBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getRecvBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
};

#endif
