// ILS3D of examples/SolverTest (ILS3D.h) over device-resident
// LocalScalar3D fields: the same methods, arguments and stopping rules,
// each whole-field step one call of the library (include/bcm_hip.h)
// instead of a loop over blocks.  Scalars come back to the host where the
// reference reads them (the dot products and norms that steer the
// iteration), so every iteration makes the reference's decisions from the
// same quantities.
//
// The algorithms are the reference's, step for step: Jacobi
// (ILS3D.h:478-529), CG (:531-661), right-preconditioned BiCGSTAB
// (:663-844) with IsConverged (:41-63), and the *_Mask forms (:846-1372).
// The library's fused solvers (bcm_bicgstab, bcm_solve) compute the same
// iterations in one call each; this class is the drop-in for driver code
// written against ILS3D's interface.
#ifndef BCM_FACADE_ILS3D_H
#define BCM_FACADE_ILS3D_H

#include <cfloat>
#include <cmath>
#include <iostream>

#include "BlockManager.h"
#include "LocalScalar3D.h"
#include "bcm_hip.h"

class ILS3D {
  typedef LocalScalar3D<double> LS;
  typedef LocalScalar3D<int> Mask;

  static void ok(int rc, const char* where) {
    if (rc != BCM_OK) {
      std::cout << "*** error: ILS3D::" << where << " returned " << rc << std::endl;
      Exit(EX_FAILURE);
    }
  }
  struct A7 {
    bcm_field_t f[7];
    A7(LS* Ap, LS* Aw, LS* Ae, LS* As, LS* An, LS* Ab, LS* At)
        : f{Ap->field(), Aw->field(), Ae->field(), As->field(), An->field(), Ab->field(), At->field()} {}
  };

  // ILS3D.h:41-63: converged when |rr|/|bb| (or |rr| for a zero b) is at
  // most epsilon^2; a NaN or infinite residual ends the program.
  bool IsConverged(BlockManager&, double& residual, double& rr, double& bb, double& epsilon, int count,
                   int countMax) {
    (void)count;
    (void)countMax;
    residual = (std::fabs(bb) > FLT_MIN) ? std::fabs(rr) / std::fabs(bb) : std::fabs(rr);
    if (residual <= epsilon * epsilon) return true;
    if (std::isnan(residual)) {
      std::cout << bb << " " << rr << std::endl;
      Exit(0);
    }
    if (std::isinf(residual)) Exit(0);
    return false;
  }

 public:
  ILS3D() {}

  // ILS3D.h:65-126: x0 <- jacobi(x); swap; exchange and conditions.
  void Jacobi_Smoother(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn,
                       LS* plsAb, LS* plsAt, LS* plsb, LS* plsx0, double omega) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_jacobi7(plsx0->field(), plsx->field(), A.f, plsb->field(), omega), "Jacobi_Smoother");
    LSSwap(*plsx, *plsx0);
    plsx->ImposeBoundaryCondition(bm);
  }
  void CalcAx(BlockManager&, LS* plsAx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb, LS* plsAt,
              LS* plsx) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_calc_ax(plsAx->field(), A.f, plsx->field()), "CalcAx");
  }
  void CalcR(BlockManager&, LS* plsr, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb, LS* plsAt,
             LS* plsx, LS* plsb) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_calc_r(plsr->field(), A.f, plsx->field(), plsb->field()), "CalcR");
  }
  void CalcR2(BlockManager&, double& rr, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
              LS* plsAt, LS* plsx, LS* plsb) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_calc_r2(A.f, plsx->field(), plsb->field(), &rr), "CalcR2");
  }
  void DOT(BlockManager&, double& xy, LS* plsx, LS* plsy) { ok(bcm_dot(plsx->field(), plsy->field(), &xy), "DOT"); }
  void AXPY(BlockManager&, LS* plsy, LS* plsx, double a) { ok(bcm_axpy(plsy->field(), plsx->field(), a), "AXPY"); }
  void XPAY(BlockManager&, LS* plsy, LS* plsx, double a) { ok(bcm_xpay(plsy->field(), plsx->field(), a), "XPAY"); }
  void AXPYZ(BlockManager&, LS* plsz, LS* plsx, LS* plsy, double a) {
    ok(bcm_axpyz(plsz->field(), plsx->field(), plsy->field(), a), "AXPYZ");
  }
  void AXPBYPZ(BlockManager&, LS* plsz, LS* plsx, LS* plsy, double a, double b) {
    ok(bcm_axpbypz(plsz->field(), plsx->field(), plsy->field(), a, b), "AXPBYPZ");
  }
  void Fill(BlockManager&, LS* plsx, double value) { ok(bcm_fill(plsx->field(), value), "Fill"); }
  void Copy(BlockManager&, LS* plsy, LS* plsx) { ok(bcm_copy(plsy->field(), plsx->field()), "Copy"); }

  void Jacobi_PreConditioner(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn,
                             LS* plsAb, LS* plsAt, LS* plsb, LS* plsx0, double omega, int countPreConditioner) {
    for (int c = 0; c < countPreConditioner; ++c)
      Jacobi_Smoother(bm, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsx0, omega);
  }

  // ILS3D.h:478-529.
  void Jacobi(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
              LS* plsAt, LS* plsb, LS* plsx0, double omega, int countMax, double epsilon, int& count,
              double& residual) {
    double bb = 0.0;
    DOT(bm, bb, plsb, plsb);
    for (count = 1; count <= countMax; ++count) {
      Jacobi_Smoother(bm, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsx0, omega);
      double rr = 0.0;
      CalcR2(bm, rr, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsx, plsb);
      if (IsConverged(bm, residual, rr, bb, epsilon, count, countMax)) break;
    }
    plsx->ImposeBoundaryCondition(bm);
  }

  // ILS3D.h:531-661: Jacobi-preconditioned conjugate gradients.
  void CG(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb, LS* plsAt,
          LS* plsb, LS* plsr, LS* plsp, LS* plsq, LS* plsz, LS* plsx0, double omega, int countPreConditioner,
          int countMax, double epsilon, int& count, double& residual) {
    double bb = 0.0;
    DOT(bm, bb, plsb, plsb);
    if (std::fabs(bb) < FLT_MIN) {
      residual = 0.0;
      count = 0;
      return;
    }
    CalcR(bm, plsr, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsx, plsb);
    plsr->ImposeBoundaryCondition(bm);
    double rr0 = 1.0;
    for (count = 1; count <= countMax; ++count) {
      Fill(bm, plsz, 0.0);
      Jacobi_PreConditioner(bm, plsz, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsr, plsx0, omega,
                            countPreConditioner);
      double rr1 = 0.0;
      DOT(bm, rr1, plsr, plsz);
      if (std::fabs(rr1) < FLT_MIN) {
        residual = rr1;
        count = 0;
        break;
      }
      const double beta = rr1 / rr0;
      if (count == 1) Copy(bm, plsp, plsz);
      else XPAY(bm, plsp, plsz, beta);
      plsp->ImposeBoundaryCondition(bm);
      CalcAx(bm, plsq, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsp);
      double qp = 0.0;
      DOT(bm, qp, plsq, plsp);
      const double alpha = rr1 / qp;
      AXPY(bm, plsx, plsp, alpha);
      AXPY(bm, plsr, plsq, -alpha);
      plsr->ImposeBoundaryCondition(bm);
      double rr = 0.0;
      DOT(bm, rr, plsr, plsr);
      if (IsConverged(bm, residual, rr, bb, epsilon, count, countMax)) break;
      rr0 = rr1;
    }
    plsx->ImposeBoundaryCondition(bm);
  }

  // ILS3D.h:663-844: BiCGSTAB, right-preconditioned by Jacobi sweeps.
  void BiCGSTAB(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
                LS* plsAt, LS* plsb, LS* plsr, LS* plsr0, LS* plsp, LS* plsp_, LS* plsq_, LS* plss, LS* plss_,
                LS* plst_, LS* plsx0, double omega, int countPreConditioner, int countMax, double epsilon,
                int& count, double& residual) {
    bicgstab(bm, nullptr, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsr, plsr0, plsp, plsp_,
             plsq_, plss, plss_, plst_, plsx0, omega, countPreConditioner, countMax, epsilon, count, residual);
  }

  // ILS3D.h:846-899: masked sweeps, unmasked residual norm.
  void Jacobi_Mask(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
                   LS* plsAt, LS* plsb, LS* plsx0, Mask* plsMaskId, double omega, int countMax, double epsilon,
                   int& count, double& residual) {
    double bb = 0.0;
    DOT(bm, bb, plsb, plsb);
    for (count = 1; count <= countMax; ++count) {
      Jacobi_Smoother_Mask(bm, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsx0, plsMaskId,
                           omega);
      double rr = 0.0;
      CalcR2(bm, rr, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsx, plsb);
      if (IsConverged(bm, residual, rr, bb, epsilon, count, countMax)) break;
    }
    plsx->ImposeBoundaryCondition(bm);
  }

  // ILS3D.h:901-1097: the dot products, preconditioner sweeps and A p_,
  // A s_ masked; the initial residual and the vector updates not.
  void BiCGSTAB_Mask(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
                     LS* plsAt, LS* plsb, LS* plsr, LS* plsr0, LS* plsp, LS* plsp_, LS* plsq_, LS* plss, LS* plss_,
                     LS* plst_, LS* plsx0, Mask* plsMaskId, double omega, int countPreConditioner, int countMax,
                     double epsilon, int& count, double& residual) {
    bicgstab(bm, plsMaskId, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsr, plsr0, plsp, plsp_,
             plsq_, plss, plss_, plst_, plsx0, omega, countPreConditioner, countMax, epsilon, count, residual);
  }

  void DOT_Mask(BlockManager&, double& xy, LS* plsx, LS* plsy, Mask* plsMaskId) {
    ok(bcm_dot_mask(plsx->field(), plsy->field(), plsMaskId->field(), &xy), "DOT_Mask");
  }
  void Jacobi_PreConditioner_Mask(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs,
                                  LS* plsAn, LS* plsAb, LS* plsAt, LS* plsb, LS* plsx0, Mask* plsMaskId,
                                  double omega, int countPreConditioner) {
    for (int c = 0; c < countPreConditioner; ++c)
      Jacobi_Smoother_Mask(bm, plsx, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsb, plsx0, plsMaskId,
                           omega);
  }
  // ILS3D.h:1175-1239.
  void Jacobi_Smoother_Mask(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn,
                            LS* plsAb, LS* plsAt, LS* plsb, LS* plsx0, Mask* plsMaskId, double omega) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_jacobi7_mask(plsx0->field(), plsx->field(), A.f, plsb->field(), plsMaskId->field(), omega),
       "Jacobi_Smoother_Mask");
    LSSwap(*plsx, *plsx0);
    plsx->ImposeBoundaryCondition(bm);
  }
  // ILS3D.h:1241-1325: x0 <- jacobi(x), exchange x0, x <- jacobi(x0).
  void Jacobi_Smoother_Mask_2(BlockManager& bm, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn,
                              LS* plsAb, LS* plsAt, LS* plsb, LS* plsx0, Mask* plsMaskId, double omega) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_jacobi7_mask(plsx0->field(), plsx->field(), A.f, plsb->field(), plsMaskId->field(), omega),
       "Jacobi_Smoother_Mask_2");
    plsx0->UpdateVirtualCells(bm);
    ok(bcm_jacobi7_mask(plsx->field(), plsx0->field(), A.f, plsb->field(), plsMaskId->field(), omega),
       "Jacobi_Smoother_Mask_2");
    plsx->ImposeBoundaryCondition(bm);
  }
  void CalcAx_Mask(BlockManager&, LS* plsAx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn, LS* plsAb,
                   LS* plsAt, LS* plsx, Mask* plsMaskId) {
    A7 A(plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt);
    ok(bcm_calc_ax_mask(plsAx->field(), A.f, plsx->field(), plsMaskId->field()), "CalcAx_Mask");
  }

 private:
  // BiCGSTAB and BiCGSTAB_Mask share their steps; with a mask the dot
  // products, the preconditioner and A p_, A s_ are the masked forms.
  void bicgstab(BlockManager& bm, Mask* m, LS* plsx, LS* plsAp, LS* plsAw, LS* plsAe, LS* plsAs, LS* plsAn,
                LS* plsAb, LS* plsAt, LS* plsb, LS* plsr, LS* plsr0, LS* plsp, LS* plsp_, LS* plsq_, LS* plss,
                LS* plss_, LS* plst_, LS* plsx0, double omega, int countPre, int countMax, double epsilon,
                int& count, double& residual) {
    auto dot = [&](double& v, LS* a, LS* b) {
      if (m) DOT_Mask(bm, v, a, b, m);
      else DOT(bm, v, a, b);
    };
    auto precondition = [&](LS* x, LS* rhs) {
      if (m) Jacobi_PreConditioner_Mask(bm, x, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, rhs, plsx0, m,
                                        omega, countPre);
      else Jacobi_PreConditioner(bm, x, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, rhs, plsx0, omega,
                                 countPre);
    };
    auto ax = [&](LS* out, LS* x) {
      if (m) CalcAx_Mask(bm, out, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, x, m);
      else CalcAx(bm, out, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, x);
    };
    double bb = 0.0;
    dot(bb, plsb, plsb);
    if (std::fabs(bb) < FLT_MIN) {
      residual = 0.0;
      count = 0;
      return;
    }
    CalcR(bm, plsr, plsAp, plsAw, plsAe, plsAs, plsAn, plsAb, plsAt, plsx, plsb);
    Copy(bm, plsr0, plsr);
    double rr0 = 1.0, alpha = 0.0, gamma = 1.0;
    for (count = 1; count <= countMax; ++count) {
      double rr1 = 0.0;
      dot(rr1, plsr, plsr0);
      if (std::fabs(rr1) < FLT_MIN) {
        residual = rr1;
        count = 0;
        break;
      }
      if (count == 1) {
        Copy(bm, plsp, plsr);
      } else {
        const double beta = rr1 / rr0 * alpha / gamma;
        AXPY(bm, plsp, plsq_, -gamma);
        XPAY(bm, plsp, plsr, beta);
      }
      plsp->ImposeBoundaryCondition(bm);
      Fill(bm, plsp_, 0.0);
      precondition(plsp_, plsp);
      ax(plsq_, plsp_);
      double q_r0 = 0.0;
      dot(q_r0, plsq_, plsr0);
      alpha = rr1 / q_r0;
      AXPYZ(bm, plss, plsq_, plsr, -alpha);
      plss->ImposeBoundaryCondition(bm);
      Fill(bm, plss_, 0.0);
      precondition(plss_, plss);
      ax(plst_, plss_);
      double t_s = 0.0, t_t_ = 0.0;
      dot(t_s, plst_, plss);
      dot(t_t_, plst_, plst_);
      gamma = t_s / t_t_;
      AXPBYPZ(bm, plsx, plsp_, plss_, alpha, gamma);
      AXPYZ(bm, plsr, plst_, plss, -gamma);
      double rr = 0.0;
      dot(rr, plsr, plsr);
      if (IsConverged(bm, residual, rr, bb, epsilon, count, countMax)) break;
      rr0 = rr1;
    }
    plsx->ImposeBoundaryCondition(bm);
  }
};

#endif
