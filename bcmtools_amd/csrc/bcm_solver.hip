// Iterative solvers on the device mesh: ILS3D::Jacobi and ILS3D::BiCGSTAB
// (examples/SolverTest/ILS3D.h:41-844) over stored coefficients, the same
// BiCGSTAB over the matrix-free Poisson operator, and the mixed-precision
// solver (FP32 inner BiCGSTAB inside FP64 iterative refinement).
//
// The host drives the iteration; every vector operation is a kernel on the
// mesh stream and each scalar (dot product, norm) is reduced on the device
// by a fixed tree, then read back, as ILS3D::DOT reads its MPI allreduce.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <utility>

#include "bcm_internal.h"

namespace bcm {

namespace {

// IsConverged (ILS3D.h:41-63): residual = |rr|/|bb| compared with eps^2.
int is_converged(double rr, double bb, double eps, double* residual, bool* done) {
  const double res = (std::fabs(bb) > FLT_MIN) ? std::fabs(rr) / std::fabs(bb) : std::fabs(rr);
  *residual = res;
  *done = false;
  if (std::isnan(res) || std::isinf(res)) return BCM_ERR_NAN;
  if (res <= eps * eps) *done = true;
  return BCM_OK;
}

// Workspace slot: one field-sized device buffer kept on the mesh.
void* ws_get(bcm_mesh* m, int slot, size_t bytes, int* rc) {
  *rc = BCM_OK;
  if (int(m->ws.size()) <= slot) m->ws.resize(slot + 1, nullptr);
  if (!m->ws[slot]) {
    void* p = nullptr;
    *rc = dev_alloc(&p, bytes);
    if (*rc != BCM_OK) return nullptr;
    if (hipMemsetAsync(p, 0, bytes, m->stream) != hipSuccess) {
      dev_free(p);
      *rc = BCM_ERR_HIP;
      return nullptr;
    }
    m->ws[slot] = p;
  }
  return m->ws[slot];
}

// Operator abstraction: stored 7 arrays (field handles) or matrix-free Poisson.
struct PoissonOp {
  bcm_mesh* m;
  template <typename T> int ax(T* out, const T* x) const { return spmv_poisson_raw<T>(m, out, x); }
  template <typename T> int r(T* out, const T* x, const T* b) const {
    return residual_poisson_raw<T>(m, out, x, b, false, 0);
  }
  template <typename T> int jacobi(T* x1, const T* x0, const T* b, double w) const {
    return jacobi_poisson_raw<T>(m, x1, x0, b, w);
  }
};

struct StoredOp {
  bcm_mesh* m;
  const bcm_field_t* A;
  int dtype;
  bcm_field_t mask = nullptr;  // ILS3D's *_Mask forms: Jacobi sweeps and A x on mask == 1 cells
  // The C ABI works on field handles; wrap raw pointers in temporaries.
  bcm_field tmp(void* p) const {
    bcm_field f;
    f.mesh = m;
    f.dtype = dtype;
    f.data = p;
    f.count = size_t(m->nblocks) * m->bvol;
    f.bytes = f.count * (dtype == BCM_F64 ? 8 : 4);
    return f;
  }
  template <typename T> int ax(T* out, const T* x) const {
    bcm_field o = tmp(out), xi = tmp(const_cast<T*>(x));
    return mask ? bcm_calc_ax_mask(&o, A, &xi, mask) : bcm_calc_ax(&o, A, &xi);
  }
  template <typename T> int r(T* out, const T* x, const T* b) const {
    bcm_field o = tmp(out), xi = tmp(const_cast<T*>(x)), bi = tmp(const_cast<T*>(b));
    return bcm_calc_r(&o, A, &xi, &bi);
  }
  template <typename T> int jacobi(T* x1, const T* x0, const T* b, double w) const {
    bcm_field o = tmp(x1), xi = tmp(const_cast<T*>(x0)), bi = tmp(const_cast<T*>(b));
    return mask ? bcm_jacobi7_mask(&o, &xi, A, &bi, mask, w) : bcm_jacobi7(&o, &xi, A, &bi, w);
  }
  template <typename T> const T* mask_data() const { return mask ? static_cast<const T*>(mask->data) : nullptr; }
};

template <typename T>
struct Bc {
  bcm_mesh* m;
  bool mirror;
  // ImposeBoundaryCondition (LocalScalar3D.h:228-258): ghost fill, then the
  // outer-face condition (the Dirichlet mirror when requested; with the
  // folded operator the outer ghosts carry zero weight).
  int impose(T* x) const {
    BCM_TRY(ghost_fill_raw<T>(m, x, BCM_GF_ALL));
    if (mirror) {
      bcm_field f;
      f.mesh = m;
      f.dtype = sizeof(T) == 8 ? BCM_F64 : BCM_F32;
      f.data = x;
      f.count = size_t(m->nblocks) * m->bvol;
      f.bytes = f.count * sizeof(T);
      // Values are already on the device (uploaded by the caller's entry point).
      BCM_TRY(dirichlet_device(m, &f));
    }
    return BCM_OK;
  }
};

template <typename T>
int dot_host(bcm_mesh* m, const T* x, const T* y, double* out, const T* mask = nullptr) {
  BCM_TRY(dot_raw<T>(m, x, y, 0, mask));
  return fetch_result(m, 1, out);
}

// Jacobi_Smoother + Jacobi_PreConditioner (ILS3D.h:65-126, 452-475):
// x0 = J(x); swap(x, x0); impose BC on x.
template <typename T, typename Op>
int jacobi_pre(const Op& op, const Bc<T>& bc, T*& x, T*& x0, const T* b, double omega, int count) {
  for (int c = 0; c < count; ++c) {
    BCM_TRY(op.template jacobi<T>(x0, x, b, omega));
    std::swap(x, x0);
    BCM_TRY(bc.impose(x));
  }
  return BCM_OK;
}

// ILS3D::BiCGSTAB (ILS3D.h:663-844).  w = {r, r0, p, p_, q_, s, s_, t_, x0}.
// Pointers in w may be permuted by the preconditioner's swaps.  With a mask
// (the operator's), ILS3D::BiCGSTAB_Mask (ILS3D.h:901-1097): the dot
// products, the preconditioner sweeps and A p_, A s_ visit only mask == 1
// cells; the initial residual and the vector updates are unmasked.
//
// xbest (the FP32 inner solve of iterative refinement, not in ILS3D): the
// iterate with the lowest recurrence residual so far is kept in xbest, and
// when the iteration ends without meeting eps, x returns that iterate.  The
// iteration also ends once the recurrence residual has set no new minimum
// for kInnerStall iterations after the first kInnerMin (it has reached the
// accuracy its arithmetic attains, about cond(A) eps32 relative, and would
// only drift; the first iterations are exempt because BiCGSTAB's residual
// often rises well above its start before it falls), before an
// update whose alpha or gamma is not finite (a breakdown: q.r0 or t.t of
// zero), and at a non-finite residual.  Each of these returns
// BCM_ERR_NOCONV with the best iterate in x.
constexpr int kInnerStall = 100;
constexpr int kInnerMin = 150;

template <typename T, typename Op>
int bicgstab(bcm_mesh* m, const Op& op, const Bc<T>& bc, T* x, const T* b, T* w[9], double omega, int count_pre,
             int count_max, double eps, int* count_out, double* residual, const T* mask = nullptr,
             T* xbest = nullptr) {
  const bool guard = xbest != nullptr;
  T *r = w[0], *r0 = w[1], *p = w[2], *p_ = w[3], *q_ = w[4], *s = w[5], *s_ = w[6], *t_ = w[7], *x0 = w[8];
  double bb = 0.0;
  BCM_TRY(dot_host<T>(m, b, b, &bb, mask));
  if (std::fabs(bb) < FLT_MIN) {
    *residual = 0.0;
    *count_out = 0;
    return BCM_OK;
  }
  BCM_TRY(op.template r<T>(r, x, b));
  BCM_TRY(blas_raw<T>(m, OP_COPY_INTERIOR, r0, r, nullptr, 0.0, 0.0));
  double rr0 = 1.0, alpha = 0.0, gamma = 1.0;
  double best_rr = HUGE_VAL;
  int since_best = 0;
  bool best_is_x = true;  // x = 0 at the start is the only iterate so far
  if (guard) BCM_TRY(blas_raw<T>(m, OP_COPY_INTERIOR, xbest, x, nullptr, 0.0, 0.0));
  int count = 1;
  int status = BCM_ERR_NOCONV;
  for (count = 1; count <= count_max; ++count) {
    double rr1 = 0.0;
    BCM_TRY(dot_host<T>(m, r, r0, &rr1, mask));
    if (std::fabs(rr1) < FLT_MIN) {
      *residual = rr1;
      count = 0;
      status = BCM_OK;
      break;
    }
    if (count == 1) {
      BCM_TRY(blas_raw<T>(m, OP_COPY_INTERIOR, p, r, nullptr, 0.0, 0.0));
    } else {
      const double beta = rr1 / rr0 * alpha / gamma;
      BCM_TRY(blas_raw<T>(m, OP_AXPY, p, q_, nullptr, -gamma, 0.0));
      BCM_TRY(blas_raw<T>(m, OP_XPAY, p, r, nullptr, beta, 0.0));
    }
    BCM_TRY(bc.impose(p));
    BCM_TRY(fill_raw<T>(m, p_, 0.0));
    BCM_TRY(jacobi_pre<T>(op, bc, p_, x0, p, omega, count_pre));
    BCM_TRY(op.template ax<T>(q_, p_));
    double q_r0 = 0.0;
    BCM_TRY(dot_host<T>(m, q_, r0, &q_r0, mask));
    alpha = rr1 / q_r0;
    if (guard && !std::isfinite(alpha)) {
      --count;
      break;
    }
    BCM_TRY(blas_raw<T>(m, OP_AXPYZ, s, q_, r, -alpha, 0.0));
    BCM_TRY(bc.impose(s));
    BCM_TRY(fill_raw<T>(m, s_, 0.0));
    BCM_TRY(jacobi_pre<T>(op, bc, s_, x0, s, omega, count_pre));
    BCM_TRY(op.template ax<T>(t_, s_));
    double ts[2];
    BCM_TRY(dot_raw<T>(m, t_, s, 0, mask));
    BCM_TRY(dot_raw<T>(m, t_, t_, 1, mask));
    BCM_TRY(fetch_result(m, 2, ts));
    gamma = ts[0] / ts[1];
    if (guard && !std::isfinite(gamma)) {
      --count;
      break;
    }
    BCM_TRY(blas_raw<T>(m, OP_AXPBYPZ, x, p_, s_, alpha, gamma));
    BCM_TRY(blas_raw<T>(m, OP_AXPYZ, r, t_, s, -gamma, 0.0));
    double rr = 0.0;
    BCM_TRY(dot_host<T>(m, r, r, &rr, mask));
    bool done = false;
    const int crc = is_converged(rr, bb, eps, residual, &done);
    if (crc != BCM_OK) {
      if (!guard) return crc;
      best_is_x = false;
      break;
    }
    if (done) {
      status = BCM_OK;
      best_is_x = true;
      break;
    }
    if (guard) {
      if (rr < best_rr) {
        best_rr = rr;
        since_best = 0;
        best_is_x = true;
        BCM_TRY(blas_raw<T>(m, OP_COPY_INTERIOR, xbest, x, nullptr, 0.0, 0.0));
      } else {
        best_is_x = false;
        if (++since_best >= kInnerStall && count >= kInnerMin) break;
      }
    }
    rr0 = rr1;
  }
  if (guard && !best_is_x) BCM_TRY(blas_raw<T>(m, OP_COPY_INTERIOR, x, xbest, nullptr, 0.0, 0.0));
  if (count > count_max) count = count_max;
  BCM_TRY(bc.impose(x));
  *count_out = count;
  w[0] = r; w[1] = r0; w[2] = p; w[3] = p_; w[4] = q_; w[5] = s; w[6] = s_; w[7] = t_; w[8] = x0;
  return status;
}

}  // namespace

void* workspace(bcm_mesh* m, int slot, size_t bytes, int* rc) { return ws_get(m, slot, bytes, rc); }

}  // namespace bcm

using namespace bcm;

// ILS3D::Jacobi (ILS3D.h:478-529); with a mask ILS3D::Jacobi_Mask
// (ILS3D.h:846-899), where only the sweep is masked.
static int jacobi_solve_impl(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, bcm_field_t x0,
                             const bcm_krylov_params* p, const double* bvalues, int nbvalues, int* count,
                             double* residual, bcm_field_t mask) {
  if (!x || !b || !x0 || !A || !p || !count || !residual) return BCM_ERR_ARG;
  if (x->mesh != b->mesh || x->mesh != x0->mesh || x->dtype != b->dtype || x->dtype != x0->dtype) return BCM_ERR_ARG;
  for (int i = 0; i < 7; ++i)
    if (!A[i] || A[i]->mesh != x->mesh || A[i]->dtype != x->dtype) return BCM_ERR_ARG;
  if (mask && (mask->mesh != x->mesh || mask->dtype != x->dtype)) return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  const bool mirror = nbvalues > 0;
  if (mirror) BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  double bb = 0.0;
  BCM_TRY(bcm_dot(b, b, &bb));
  int status = BCM_ERR_NOCONV;
  int c = 1;
  for (c = 1; c <= p->count_max; ++c) {
    if (mask) BCM_TRY(bcm_jacobi7_mask(x0, x, A, b, mask, p->omega));
    else BCM_TRY(bcm_jacobi7(x0, x, A, b, p->omega));
    std::swap(x->data, x0->data);
    BCM_TRY(bcm_ghost_fill(x, BCM_GF_ALL));
    if (mirror) BCM_TRY(dirichlet_device(m, x));
    double rr = 0.0;
    BCM_TRY(bcm_calc_r2(A, x, b, &rr));
    const double res = (std::fabs(bb) > FLT_MIN) ? std::fabs(rr) / std::fabs(bb) : std::fabs(rr);
    *residual = res;
    if (std::isnan(res) || std::isinf(res)) return BCM_ERR_NAN;
    if (res <= p->epsilon * p->epsilon) {
      status = BCM_OK;
      break;
    }
  }
  *count = c > p->count_max ? p->count_max : c;
  BCM_TRY(bcm_ghost_fill(x, BCM_GF_ALL));
  if (mirror) BCM_TRY(dirichlet_device(m, x));
  return status;
}

extern "C" int bcm_jacobi_solve(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, bcm_field_t x0,
                                const bcm_krylov_params* p, const double* bvalues, int nbvalues, int* count,
                                double* residual) {
  return jacobi_solve_impl(x, A, b, x0, p, bvalues, nbvalues, count, residual, nullptr);
}

extern "C" int bcm_jacobi_solve_mask(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, bcm_field_t x0,
                                     bcm_field_t mask, const bcm_krylov_params* p, const double* bvalues,
                                     int nbvalues, int* count, double* residual) {
  if (!mask) return BCM_ERR_ARG;
  return jacobi_solve_impl(x, A, b, x0, p, bvalues, nbvalues, count, residual, mask);
}

static int bicgstab_impl(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, const bcm_field_t work[9],
                         const bcm_krylov_params* p, const double* bvalues, int nbvalues, int* count,
                         double* residual, bcm_field_t mask) {
  if (!x || !b || !A || !work || !p || !count || !residual) return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  for (int i = 0; i < 7; ++i)
    if (!A[i] || A[i]->mesh != m || A[i]->dtype != x->dtype) return BCM_ERR_ARG;
  for (int i = 0; i < 9; ++i)
    if (!work[i] || work[i]->mesh != m || work[i]->dtype != x->dtype) return BCM_ERR_ARG;
  if (b->mesh != m || b->dtype != x->dtype) return BCM_ERR_ARG;
  if (mask && (mask->mesh != m || mask->dtype != x->dtype)) return BCM_ERR_ARG;
  const bool mirror = nbvalues > 0;
  if (mirror) BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  StoredOp op{m, A, x->dtype, mask};
  int rc;
  if (x->dtype == BCM_F64) {
    double* w[9];
    for (int i = 0; i < 9; ++i) w[i] = static_cast<double*>(work[i]->data);
    rc = bicgstab<double>(m, op, Bc<double>{m, mirror}, static_cast<double*>(x->data),
                          static_cast<const double*>(b->data), w, p->omega, p->count_pre, p->count_max, p->epsilon,
                          count, residual, op.mask_data<double>());
    for (int i = 0; i < 9; ++i) work[i]->data = w[i];
  } else {
    float* w[9];
    for (int i = 0; i < 9; ++i) w[i] = static_cast<float*>(work[i]->data);
    rc = bicgstab<float>(m, op, Bc<float>{m, mirror}, static_cast<float*>(x->data),
                         static_cast<const float*>(b->data), w, p->omega, p->count_pre, p->count_max, p->epsilon,
                         count, residual, op.mask_data<float>());
    for (int i = 0; i < 9; ++i) work[i]->data = w[i];
  }
  return rc;
}

extern "C" int bcm_bicgstab(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, const bcm_field_t work[9],
                            const bcm_krylov_params* p, const double* bvalues, int nbvalues, int* count,
                            double* residual) {
  return bicgstab_impl(x, A, b, work, p, bvalues, nbvalues, count, residual, nullptr);
}

extern "C" int bcm_bicgstab_mask(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, const bcm_field_t work[9],
                                 bcm_field_t mask, const bcm_krylov_params* p, const double* bvalues, int nbvalues,
                                 int* count, double* residual) {
  if (!mask) return BCM_ERR_ARG;
  return bicgstab_impl(x, A, b, work, p, bvalues, nbvalues, count, residual, mask);
}

// Workspace slots on the mesh.
// Slots: b; 9 FP64 BiCGSTAB vectors; 12 FP32 vectors of the mixed solver;
// the FP64 trial iterate and the FP64 correction's residual, correction and
// best iterate.  The SOR workspaces use slots 30 and up.
enum {
  WS_B64 = 0,
  WS_F64_FIRST = 1,
  WS_F32_FIRST = 12,
  WS_X64_TRIAL = WS_F32_FIRST + 12,
  WS_R64,
  WS_D64,
  WS_BEST64
};
static_assert(WS_BEST64 < WS_SOR_CA_F64, "workspace slots overlap");

extern "C" int bcm_solve_fp64(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
                              const bcm_ir_params* p, int* iters, double* relres) {
  if (!x || !rhs || !p || !iters || !relres || x->mesh != rhs->mesh || x->dtype != BCM_F64 ||
      rhs->dtype != BCM_F64)
    return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  const size_t bytes = size_t(m->nblocks) * m->bvol * sizeof(double);
  int rc = BCM_OK;
  double* b = static_cast<double*>(ws_get(m, WS_B64, bytes, &rc));
  if (rc != BCM_OK) return rc;
  double* w[9];
  for (int i = 0; i < 9; ++i) {
    w[i] = static_cast<double*>(ws_get(m, WS_F64_FIRST + i, bytes, &rc));
    if (rc != BCM_OK) return rc;
  }
  BCM_TRY(assemble_poisson_b_raw(m, b, static_cast<const double*>(rhs->data)));
  BCM_TRY(ghost_fill_raw<double>(m, static_cast<double*>(x->data), BCM_GF_ALL));
  // The recurrence residual can drift below the true one; restart from the
  // current x until the true FP64 residual meets the tolerance.
  *iters = 0;
  for (int restart = 0; restart < 8; ++restart) {
    double res = 0.0;
    int cnt = 0;
    rc = bicgstab<double>(m, PoissonOp{m}, Bc<double>{m, false}, static_cast<double*>(x->data), b, w, p->omega,
                          p->count_pre, p->inner_max - *iters, 0.5 * p->tol, &cnt, &res);
    for (int i = 0; i < 9; ++i) m->ws[WS_F64_FIRST + i] = w[i];
    if (rc != BCM_OK && rc != BCM_ERR_NOCONV) return rc;
    *iters += cnt;
    // True FP64 relative residual ||b - A x|| / ||b||.
    double v[2];
    BCM_TRY(residual_poisson_raw<double>(m, w[0], static_cast<const double*>(x->data), b, true, 0));
    BCM_TRY(dot_raw<double>(m, b, b, 1));
    BCM_TRY(fetch_result(m, 2, v));
    *relres = v[1] > 0.0 ? std::sqrt(v[0] / v[1]) : std::sqrt(v[0]);
    if (*relres <= p->tol) return BCM_OK;
    if (*iters >= p->inner_max || cnt == 0) break;
  }
  return BCM_ERR_NOCONV;
}

extern "C" int bcm_solve_mixed(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
                               const bcm_ir_params* p, int* outer, int* inner_total, double* relres) {
  if (!x || !rhs || !p || !outer || !inner_total || !relres || x->mesh != rhs->mesh || x->dtype != BCM_F64 ||
      rhs->dtype != BCM_F64 || p->outer_max < 1)
    return BCM_ERR_ARG;
  bcm_mesh* m = x->mesh;
  BCM_TRY(set_bvalues(m, bvalues, nbvalues));
  const size_t n = size_t(m->nblocks) * m->bvol;
  int rc = BCM_OK;
  double* b = static_cast<double*>(ws_get(m, WS_B64, n * 8, &rc));
  if (rc != BCM_OK) return rc;
  float* w[12];  // r, r0, p, p_, q_, s, s_, t_, x0, rhs32, d32, best d32
  for (int i = 0; i < 12; ++i) {
    w[i] = static_cast<float*>(ws_get(m, WS_F32_FIRST + i, n * 4, &rc));
    if (rc != BCM_OK) return rc;
  }
  double* xd = static_cast<double*>(x->data);
  BCM_TRY(assemble_poisson_b_raw(m, b, static_cast<const double*>(rhs->data)));
  double bb = 0.0;
  BCM_TRY(dot_host<double>(m, b, b, &bb));
  *outer = 0;
  *inner_total = 0;
  *relres = 0.0;
  if (bb == 0.0) {
    BCM_TRY(fill_raw<double>(m, xd, 0.0));
    return BCM_OK;
  }
  const PoissonOp op32{m};
  const Bc<float> bc32{m, false};
  float* r32 = w[9];
  float* d32 = w[10];
  // The refined iterate alternates between x and an FP64 workspace: a
  // correction is applied into the other buffer and kept only when it
  // lowers the FP64 residual.  The FP32 inner solve returns its best iterate
  // by its recurrence residual, which can drift away from the true one by
  // about cond(A) eps32 times the largest residual on its way.  A correction
  // that does not lower the FP64 residual is discarded; the remaining steps
  // then solve their corrections in FP64 (unless p->fp32_strict, which
  // stops with BCM_ERR_NOCONV and the best x so far).
  double* xt = static_cast<double*>(ws_get(m, WS_X64_TRIAL, n * 8, &rc));
  if (rc != BCM_OK) return rc;
  double* cur = xd;
  double* alt = xt;
  bool corr64 = false;
  int status = BCM_ERR_NOCONV;
  // FP64 residual r = b - A x (x ghosts refreshed), rounded to FP32.
  auto residual = [&](double* v, double* rel) -> int {
    BCM_TRY(ghost_fill_raw<double>(m, v, BCM_GF_ALL));
    BCM_TRY(residual_d2f_raw(m, r32, v, b, 0));
    double rr = 0.0;
    BCM_TRY(fetch_result(m, 1, &rr));
    *rel = std::sqrt(rr / bb);
    return BCM_OK;
  };
  BCM_TRY(residual(cur, relres));
  if (!std::isfinite(*relres)) return BCM_ERR_NAN;
  for (int it = 0; it < p->outer_max && *relres > p->tol; ++it) {
    if (p->fp32_only && it == 1) break;
    const double itol = p->fp32_only ? 0.5 * p->tol : p->inner_tol;
    int cnt = 0;
    double res = 0.0;
    BCM_HIP_TRY(hipMemcpyAsync(alt, cur, n * 8, hipMemcpyDeviceToDevice, m->stream));
    if (!corr64) {
      // Correction A d = r in FP32, Jacobi-preconditioned BiCGSTAB.
      BCM_TRY(fill_raw<float>(m, d32, 0.0));
      float* ww[9] = {w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8]};
      rc = bicgstab<float>(m, op32, bc32, d32, r32, ww, p->omega, p->count_pre, p->inner_max, itol, &cnt, &res,
                           nullptr, w[11]);
      for (int i = 0; i < 9; ++i) w[i] = ww[i];
      for (int i = 0; i < 9; ++i) m->ws[WS_F32_FIRST + i] = w[i];
      if (rc != BCM_OK && rc != BCM_ERR_NOCONV) return rc;
      BCM_TRY(convert_raw(m, CONV_ADD_F2D, alt, d32));
    } else {
      // The same correction in FP64: r = b - A x, then A d = r.
      double* w64[9];
      for (int i = 0; i < 9; ++i) {
        w64[i] = static_cast<double*>(ws_get(m, WS_F64_FIRST + i, n * 8, &rc));
        if (rc != BCM_OK) return rc;
      }
      double* r64 = static_cast<double*>(ws_get(m, WS_R64, n * 8, &rc));
      if (rc != BCM_OK) return rc;
      double* d64 = static_cast<double*>(ws_get(m, WS_D64, n * 8, &rc));
      if (rc != BCM_OK) return rc;
      double* best64 = static_cast<double*>(ws_get(m, WS_BEST64, n * 8, &rc));
      if (rc != BCM_OK) return rc;
      BCM_TRY(residual_poisson_raw<double>(m, r64, cur, b, false, 0));
      BCM_TRY(fill_raw<double>(m, d64, 0.0));
      rc = bicgstab<double>(m, PoissonOp{m}, Bc<double>{m, false}, d64, r64, w64, p->omega, p->count_pre,
                            p->inner_max, p->inner_tol, &cnt, &res, nullptr, best64);
      for (int i = 0; i < 9; ++i) m->ws[WS_F64_FIRST + i] = w64[i];
      if (rc != BCM_OK && rc != BCM_ERR_NOCONV) return rc;
      BCM_TRY(blas_raw<double>(m, OP_AXPY, alt, d64, nullptr, 1.0, 0.0));
    }
    *inner_total += cnt > 0 ? cnt : 0;
    *outer = it + 1;
    // Trial x' = x + d in FP64, kept if its FP64 residual is lower.
    double rel_alt = 0.0;
    BCM_TRY(residual(alt, &rel_alt));
    if (std::isfinite(rel_alt) && rel_alt < *relres) {
      std::swap(cur, alt);
      *relres = rel_alt;
    } else {
      // Rejected: r32 must hold the residual of the kept iterate again.
      BCM_TRY(residual(cur, relres));
      if (corr64 || p->fp32_strict || p->fp32_only) break;
      corr64 = true;
    }
  }
  if (*relres <= p->tol) status = BCM_OK;
  if (cur != xd) BCM_HIP_TRY(hipMemcpyAsync(xd, cur, n * 8, hipMemcpyDeviceToDevice, m->stream));
  BCM_TRY(ghost_fill_raw<double>(m, xd, BCM_GF_ALL));
  return status;
}
