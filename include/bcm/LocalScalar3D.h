// LocalScalar3D<T> of examples/SolverTest (LocalScalar3D.h, BlockScalar3D.h)
// held on the GPU: one device-resident data class of the BlockManager plus
// the field's boundary conditions (one type and value per domain face).
//
// GetBlockData(block) returns the block's array in device memory, the
// argument SolverTest's per-block Fortran kernels take (include/
// bcm_fortran.h); the whole-field operations (UpdateVirtualCells,
// ImposeBoundaryCondition, Fill) are single device calls.  The host copy
// moves only on syncToHost / syncToDevice.
//
// Boundary types are BlockScalar3D's (BlockScalar3D.h:31-38): DIRICHLET,
// NEUMANN, PERIODIC, INNER; a NEUMANN value is scaled by each block's
// cellSize.x as BlockScalar3D.h:98-100 does.  POISEUILLE_U / POISEUILLE_P
// (bcax.f90) are per-block only here: call bc_x3_poiseuille_*_ on
// GetBlockData.  The ghost exchange of every updaterType is the library's
// (Scalar3DUpdater's rules); Scalar3DUpdater1's sup.f90 boxes, whose F2C
// sum runs in another order, are available per block through
// sup_copy_from_neighbor_*_.
//
// LocalScalar3D<int> is the mask class of the *_Mask solvers: a device
// field holding 0/1 (SetupMask sets 1 on interiors, blas.f90:700-751).
#ifndef BCM_FACADE_LOCALSCALAR3D_H
#define BCM_FACADE_LOCALSCALAR3D_H

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "BlockManager.h"
#include "Scalar3D.h"
#include "Scalar3DUpdater.h"
#include "bcm_hip.h"

template <typename T>
class LocalScalar3D {
  // Device fields hold FP64 or FP32 values; the mask class holds 0/1 as FP64.
  typedef typename std::conditional<std::is_same<T, float>::value, float, double>::type Elem;

 public:
  // BlockScalar3D::BoundaryType (BlockScalar3D.h:31-38).
  enum BoundaryType { DIRICHLET, NEUMANN, PERIODIC, INNER, POISEUILLE_U = 100, POISEUILLE_P = 101 };

  LocalScalar3D(BlockManager& blockManager, int vc, std::string updateMethod, int* boundaryType,
                T* boundaryValue, int updaterType = 1)
      : bm(&blockManager), vc(vc) {
    (void)updaterType;
    id = blockManager.setDataClass<Scalar3D<Elem>, Scalar3DUpdater<Elem> >(vc);
    method = BCM_VC_AT_ONCE;
    if (updateMethod == "SeparateXYZ") method = BCM_VC_SEPARATE_XYZ;
    else if (updateMethod == "SeparateXYZ_SeparateLevelDiff") method = BCM_VC_SEPARATE_XYZ_SEPARATE_LEVEL_DIFF;
    blockManager.prepareForVCUpdate(id, 100, VCUpdateMethod::AtOnce);
    blockManager.setDeviceResident(id, true);
    for (int f = 0; f < NUM_FACE; ++f) {
      const int t = boundaryType ? boundaryType[f] : INNER;
      switch (t) {
        case DIRICHLET: types[f] = BCM_BC_DIRICHLET; break;
        case NEUMANN: types[f] = BCM_BC_NEUMANN; break;
        case PERIODIC: types[f] = BCM_BC_PERIODIC; break;
        case INNER: types[f] = BCM_BC_INNER; break;
        default:
          std::cout << "*** error: LocalScalar3D: boundary type " << t
                    << " (Poiseuille) is per block only: use bc_x3_poiseuille_*_ on GetBlockData" << std::endl;
          Exit(EX_FAILURE);
      }
      values[f] = boundaryValue ? double(boundaryValue[f]) : 0.0;
    }
  }

  int GetVC() { return vc; }
  int GetID() { return id; }
  bcm_field_t field() const { return bm->deviceField(id); }

  // The block's (n+2vc)^3 array in device memory.
  Elem* GetBlockData(BlockBase* block) {
    const int b = bm->indexOf(block);
    if (b < 0) {
      std::cout << "*** error: LocalScalar3D::GetBlockData: unknown block" << std::endl;
      Exit(EX_FAILURE);
    }
    const Vec3i& s = block->getSize();
    const size_t bvol = size_t(s.x + 2 * vc) * (s.y + 2 * vc) * (s.z + 2 * vc);
    return static_cast<Elem*>(bcm_field_device_ptr(field())) + bvol * size_t(b);
  }

  // LocalScalar3D.h:199-223.
  void UpdateVirtualCells(BlockManager&) { check(bcm_update_virtual_cells(field(), method), "UpdateVirtualCells"); }
  // LocalScalar3D.h:225-257: exchange, then each block's outer faces.
  void ImposeBoundaryCondition(BlockManager& blockManager) {
    UpdateVirtualCells(blockManager);
    check(bcm_impose_face_bc(field(), types, values), "ImposeBoundaryCondition");
  }
  // LocalScalar3D.h:259-284: the conditions folded into a system.
  void ImposeBoundaryCondition(BlockManager&, LocalScalar3D<T>* plsAp, LocalScalar3D<T>* plsAw,
                               LocalScalar3D<T>* plsAe, LocalScalar3D<T>* plsAs, LocalScalar3D<T>* plsAn,
                               LocalScalar3D<T>* plsAb, LocalScalar3D<T>* plsAt, LocalScalar3D<T>* plsb) {
    const bcm_field_t A[7] = {plsAp->field(), plsAw->field(), plsAe->field(), plsAs->field(),
                              plsAn->field(), plsAb->field(), plsAt->field()};
    check(bcm_fold_face_bc(A, plsb->field(), types, values), "ImposeBoundaryCondition(A, b)");
  }
  // LocalScalar3D.h:286-298 (an INNER face keeps its value).
  void ResetBoundaryConditionValue(BlockManager&, int i_face, T boundaryValue) {
    if (i_face >= 0 && i_face < NUM_FACE && types[i_face] != BCM_BC_INNER) values[i_face] = double(boundaryValue);
  }
  // LocalScalar3D.h:351-372: every value, ghosts included.
  void Fill(BlockManager&, T value, T deviation = T(0)) {
    (void)deviation;
    check(bcm_fill(field(), double(value)), "Fill");
  }
  // blas.f90:700-751 through Solver.cpp:999-1027 (the mask class).
  void SetupMask(BlockManager&) { check(bcm_setup_mask(field()), "SetupMask"); }

  // LocalScalar3D.cpp:24-82 (sum, max, min, |max|, |min| over interiors).
  void CalcStats(BlockManager& blockManager) {
    bcm_field_t f = field();
    const int nb = blockManager.getNumBlock();
    const Vec3i s = blockManager.getSize();
    const int nx0 = s.x + 2 * vc, ny0 = s.y + 2 * vc, nz0 = s.z + 2 * vc;
    std::vector<Elem> h(size_t(nx0) * ny0 * nz0 * nb);
    check(bcm_field_download(f, h.data(), h.size() * sizeof(Elem)), "CalcStats");
    bool first = true;
    sum_g = 0;
    for (int b = 0; b < nb; ++b) {
      double sum_b = 0.0;
      for (int k = 0; k < s.z; ++k)
        for (int j = 0; j < s.y; ++j)
          for (int i = 0; i < s.x; ++i) {
            const double v = double(h[size_t(b) * nx0 * ny0 * nz0 + (i + vc) + nx0 * ((j + vc) + size_t(ny0) * (k + vc))]);
            sum_b += v;
            if (first || v > max_g) max_g = v;
            if (first || v < min_g) min_g = v;
            if (first || std::fabs(v) > absmax_g) absmax_g = std::fabs(v);
            if (first || std::fabs(v) < absmin_g) absmin_g = std::fabs(v);
            first = false;
          }
      sum_g += sum_b;
    }
  }
  double GetSum() const { return sum_g; }
  double GetMax() const { return max_g; }
  double GetMin() const { return min_g; }
  double GetAbsMax() const { return absmax_g; }
  double GetAbsMin() const { return absmin_g; }

  // LocalScalar3D.cpp:84-324: restart files under ./BIN of rank 0, with the
  // boundary conditions imposed before a dump and after a load.
  void Dump(BlockManager& blockManager, const int step, const char* label) {
    ImposeBoundaryCondition(blockManager);
    check(bcm_field_dump(field(), "./BIN", label, 0, step), "Dump");
  }
  void Load(BlockManager& blockManager, const int step, const char* label) {
    check(bcm_field_load(field(), "./BIN", label, 0, step), "Load");
    ImposeBoundaryCondition(blockManager);
  }
  void Dump2(BlockManager& blockManager, const int step, const char* label) {
    ImposeBoundaryCondition(blockManager);
    check(bcm_field_dump2(field(), "./BIN", label, 0, step), "Dump2");
  }
  void Load2(BlockManager& blockManager, const int step, const char* label) {
    check(bcm_field_load2(field(), "./BIN", label, 0, step), "Load2");
    ImposeBoundaryCondition(blockManager);
  }

  // LocalScalar3D.h:132-137: the two objects trade their data classes.
  friend void LSSwap(LocalScalar3D<T>& x1, LocalScalar3D<T>& x0) {
    std::swap(x1.id, x0.id);
  }

 private:
  static void check(int rc, const char* where) {
    if (rc != BCM_OK) {
      std::cout << "*** error: LocalScalar3D::" << where << " returned " << rc << std::endl;
      Exit(EX_FAILURE);
    }
  }

  BlockManager* bm;
  int vc;
  int id = -1;
  int method = BCM_VC_AT_ONCE;
  int32_t types[NUM_FACE];
  double values[NUM_FACE];
  double sum_g = 0, max_g = 0, min_g = 0, absmax_g = 0, absmin_g = 0;
};

#endif
