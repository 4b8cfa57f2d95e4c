// Scalar3DUpdater<T> (include/Scalar3DUpdater.h:64-904): the VCUpdater of a
// Scalar3D<T>.  Its local rules -- same-level copy, fine-to-coarse average
// and coarse-to-fine trilinear interpolation of one face -- run on the GPU
// (bcm_ghost_fill_face, the kernels of the whole-mesh exchange), and so do
// its buffer rules (bcm_comm_pack_host / bcm_comm_unpack_host); the buffer
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
  // The buffer rules (include/Scalar3DUpdater.h:377-515, 711-901), on the
  // device: the buffers the communication layer attached, in the
  // reference's element order.  A face without a buffer is left alone.
  void copyToCommBuffer(Face face) {
    if (T* b = sendBuffer[face][0]) packOnDevice(face, SF_00, BCM_COMM_SAME, b, getSendBufferByteSize(face));
  }
  void copyToCommBufferF2C(Face face, Subface subface) {
    if (T* b = sendBuffer[face][0]) packOnDevice(face, subface, BCM_COMM_F2C, b, getSendBufferByteSizeF2C(face, subface));
  }
  void copyToCommBufferC2F(Face face, Subface subface) {
    if (T* b = sendBuffer[face][subface])
      packOnDevice(face, subface, BCM_COMM_C2F, b, getSendBufferByteSizeC2F(face, subface));
  }
  void copyFromCommBuffer(Face face) {
    if (T* b = recvBuffer[face][0]) unpackOnDevice(face, SF_00, BCM_COMM_SAME, b, getRecvBufferByteSize(face));
  }
  void copyFromCommBufferF2C(Face face, Subface subface) {
    if (T* b = recvBuffer[face][subface])
      unpackOnDevice(face, subface, BCM_COMM_F2C, b, getRecvBufferByteSizeF2C(face, subface));
  }
  void copyFromCommBufferC2F(Face face, Subface) { copyFromCommBuffer(face); }
  size_t getSendBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getSendBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getSendBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSize(Face face) const { return sizeof(T) * getCommBufferSize(face); }
  size_t getRecvBufferByteSizeF2C(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face) / 4; }
  size_t getRecvBufferByteSizeC2F(Face face, Subface) const { return sizeof(T) * getCommBufferSize(face); }
};

#endif
