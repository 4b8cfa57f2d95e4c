/*
 * bcm_oracle.c -- CPU restatement of the BCMTools block data path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
 * reported CPU baseline.  It is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the
 * checker; the product (libbcm_hip.so) never links or calls it.
 *
 * Every function restates the reference algorithm with its loop structure,
 * operand order and parallel structure (OpenMP over faces for ghost fill,
 * over blocks for the SolverTest kernels, inside a block for SOR), compiled
 * with -ffp-contract=off so FP64 results are bit-exact to the reference
 * build (g++ -O3 without -march, hence no FMA).
 *
 * Pinning (which check fixes each function):
 *   ghost fill (copy/F2C/C2F), Dirichlet mirror, sor1 sweeps: golden
 *     fixtures written by the reference library itself (oracle/_ref,
 *     tests/golden), plus closed forms (copy identity, F2C of a constant,
 *     C2F reproduces trilinear fields, extrapolation weights).
 *   sor2 sweep: shares the colour schedule with sor1 (pinned); its update
 *     formula is pinned by the discrete fixed point and by convergence to
 *     the closed-form solution sin(2 pi x) sin(2 pi y) sin(2 pi z).
 *   jacobi7 / calc_ax / calc_r / calc_r2 / BLAS-1 / BiCGSTAB: the
 *     reference kernels are Fortran (blas.f90) and no Fortran compiler
 *     exists here, so they are pinned by independent numpy brute force
 *     with the same operand order, by fixed-point and symmetry
 *     invariants, and by a direct sparse solve (tests/test_oracle.py).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bcm_hip.h"

/* Index3DS (include/Index3D.h:35-59). */
typedef struct {
  int n, vc, nx0;
  long bvol;
} geo_t;

static inline long IDX(const geo_t* g, int i, int j, int k) {
  return (long)i + (long)g->nx0 * j + (long)g->nx0 * g->nx0 * k + (long)(1 + g->nx0 + g->nx0 * g->nx0) * g->vc;
}

static geo_t mkgeo(int n, int vc) {
  geo_t g;
  g.n = n;
  g.vc = vc;
  g.nx0 = n + 2 * vc;
  g.bvol = (long)g.nx0 * g.nx0 * g.nx0;
  return g;
}

/* VCUpdater::subfaceOrigin0/1 (include/VCUpdater.h:252-273). */
static int so0(int sf) { return sf & 1; }
static int so1(int sf) { return sf >> 1; }

#define DEFINE_GHOST(T, SUF)                                                                                  \
  /* Scalar3D::copyFromDataClass_0 (include/Scalar3D.h:256-275). */                                          \
  static void copy_box_##SUF(const geo_t* g, T* d, int i0, int j0, int k0, const T* s, int i1, int j1, int k1, \
                             int nx, int ny, int nz) {                                                        \
    for (int k = 0; k < nz; ++k)                                                                              \
      for (int j = 0; j < ny; ++j)                                                                            \
        for (int i = 0; i < nx; ++i) d[IDX(g, i0 + i, j0 + j, k0 + k)] = s[IDX(g, i1 + i, j1 + j, k1 + k)];   \
  }                                                                                                           \
  /* Scalar3DUpdater::copyFromNeighbor (include/Scalar3DUpdater.h:321-347). */                               \
  static void copy_from_neighbor_##SUF(const geo_t* g, T* d, const T* s, int face) {                          \
    const int n = g->n, vc = g->vc;                                                                           \
    switch (face) {                                                                                           \
      case BCM_X_M: copy_box_##SUF(g, d, -vc, 0, 0, s, n - vc, 0, 0, vc, n, n); break;                         \
      case BCM_X_P: copy_box_##SUF(g, d, n, 0, 0, s, 0, 0, 0, vc, n, n); break;                                \
      case BCM_Y_M: copy_box_##SUF(g, d, 0, -vc, 0, s, 0, n - vc, 0, n, vc, n); break;                         \
      case BCM_Y_P: copy_box_##SUF(g, d, 0, n, 0, s, 0, 0, 0, n, vc, n); break;                                \
      case BCM_Z_M: copy_box_##SUF(g, d, 0, 0, -vc, s, 0, 0, n - vc, n, n, vc); break;                         \
      default: copy_box_##SUF(g, d, 0, 0, n, s, 0, 0, 0, n, n, vc); break;                                     \
    }                                                                                                         \
  }                                                                                                           \
  /* interpolateF2C (include/Scalar3DUpdater.h:190-198). */                                                  \
  static T f2c_##SUF(const geo_t* g, const T* f, int I, int J, int K) {                                       \
    int i = 2 * I, j = 2 * J, k = 2 * K;                                                                      \
    return (T)(0.125 * (f[IDX(g, i, j, k)] + f[IDX(g, i + 1, j, k)] + f[IDX(g, i, j + 1, k)] +                \
                        f[IDX(g, i + 1, j + 1, k)] + f[IDX(g, i, j, k + 1)] + f[IDX(g, i + 1, j, k + 1)] +    \
                        f[IDX(g, i, j + 1, k + 1)] + f[IDX(g, i + 1, j + 1, k + 1)]));                        \
  }                                                                                                           \
  /* interpolateC2F + linearInterpolate (include/Scalar3DUpdater.h:221-266). */                              \
  static T c2f_##SUF(const geo_t* g, const T* c, int i, int j, int k) {                                       \
    int n = g->n, I, J, K;                                                                                    \
    double r, s, t;                                                                                           \
    I = i / 2 - 1 + i % 2; if (I < 0) I = 0; if (I > n - 2) I = n - 2; r = -0.25 + 0.5 * i - (double)I;        \
    J = j / 2 - 1 + j % 2; if (J < 0) J = 0; if (J > n - 2) J = n - 2; s = -0.25 + 0.5 * j - (double)J;        \
    K = k / 2 - 1 + k % 2; if (K < 0) K = 0; if (K > n - 2) K = n - 2; t = -0.25 + 0.5 * k - (double)K;        \
    return (T)((1.0 - t) * ((1.0 - s) * ((1.0 - r) * c[IDX(g, I, J, K)] + r * c[IDX(g, I + 1, J, K)]) +       \
                            s * ((1.0 - r) * c[IDX(g, I, J + 1, K)] + r * c[IDX(g, I + 1, J + 1, K)])) +      \
               t * ((1.0 - s) * ((1.0 - r) * c[IDX(g, I, J, K + 1)] + r * c[IDX(g, I + 1, J, K + 1)]) +       \
                    s * ((1.0 - r) * c[IDX(g, I, J + 1, K + 1)] + r * c[IDX(g, I + 1, J + 1, K + 1)])));      \
  }                                                                                                           \
  /* copyFromNeighborF2C_0 (include/Scalar3DUpdater.h:519-612). */                                           \
  static void f2c_face_##SUF(const geo_t* g, T* c, const T* f, int face, int sf) {                            \
    const int n = g->n, vc = g->vc, h = n / 2;                                                                \
    int a0 = h * so0(sf), a1 = h * so1(sf);                                                                   \
    switch (face) {                                                                                           \
      case BCM_X_M: for (int k = 0; k < h; k++) for (int j = 0; j < h; j++) for (int i = 0; i < vc; i++)      \
          c[IDX(g, i - vc, j + a0, k + a1)] = f2c_##SUF(g, f, i + h - vc, j, k); break;                       \
      case BCM_X_P: for (int k = 0; k < h; k++) for (int j = 0; j < h; j++) for (int i = 0; i < vc; i++)      \
          c[IDX(g, i + n, j + a0, k + a1)] = f2c_##SUF(g, f, i, j, k); break;                                 \
      case BCM_Y_M: for (int k = 0; k < h; k++) for (int j = 0; j < vc; j++) for (int i = 0; i < h; i++)      \
          c[IDX(g, i + a1, j - vc, k + a0)] = f2c_##SUF(g, f, i, j + h - vc, k); break;                       \
      case BCM_Y_P: for (int k = 0; k < h; k++) for (int j = 0; j < vc; j++) for (int i = 0; i < h; i++)      \
          c[IDX(g, i + a1, j + n, k + a0)] = f2c_##SUF(g, f, i, j, k); break;                                 \
      case BCM_Z_M: for (int k = 0; k < vc; k++) for (int j = 0; j < h; j++) for (int i = 0; i < h; i++)      \
          c[IDX(g, i + a0, j + a1, k - vc)] = f2c_##SUF(g, f, i, j, k + h - vc); break;                       \
      default: for (int k = 0; k < vc; k++) for (int j = 0; j < h; j++) for (int i = 0; i < h; i++)           \
          c[IDX(g, i + a0, j + a1, k + n)] = f2c_##SUF(g, f, i, j, k); break;                                 \
    }                                                                                                         \
  }                                                                                                           \
  /* copyFromNeighborC2F_0 (include/Scalar3DUpdater.h:615-708). */                                           \
  static void c2f_face_##SUF(const geo_t* g, T* f, const T* c, int face, int sf) {                            \
    const int n = g->n, vc = g->vc;                                                                           \
    int a0 = n * so0(sf), a1 = n * so1(sf);                                                                   \
    switch (face) {                                                                                           \
      case BCM_X_M: for (int K = 0; K < n; K++) for (int J = 0; J < n; J++) for (int I = 0; I < vc; I++)      \
          f[IDX(g, I - vc, J, K)] = c2f_##SUF(g, c, I + 2 * n - vc, J + a0, K + a1); break;                   \
      case BCM_X_P: for (int K = 0; K < n; K++) for (int J = 0; J < n; J++) for (int I = 0; I < vc; I++)      \
          f[IDX(g, I + n, J, K)] = c2f_##SUF(g, c, I, J + a0, K + a1); break;                                 \
      case BCM_Y_M: for (int K = 0; K < n; K++) for (int J = 0; J < vc; J++) for (int I = 0; I < n; I++)      \
          f[IDX(g, I, J - vc, K)] = c2f_##SUF(g, c, I + a1, J + 2 * n - vc, K + a0); break;                   \
      case BCM_Y_P: for (int K = 0; K < n; K++) for (int J = 0; J < vc; J++) for (int I = 0; I < n; I++)      \
          f[IDX(g, I, J + n, K)] = c2f_##SUF(g, c, I + a1, J, K + a0); break;                                 \
      case BCM_Z_M: for (int K = 0; K < vc; K++) for (int J = 0; J < n; J++) for (int I = 0; I < n; I++)      \
          f[IDX(g, I, J, K - vc)] = c2f_##SUF(g, c, I + a0, J + a1, K + 2 * n - vc); break;                   \
      default: for (int K = 0; K < vc; K++) for (int J = 0; J < n; J++) for (int I = 0; I < n; I++)           \
          f[IDX(g, I, J, K + n)] = c2f_##SUF(g, c, I + a0, J + a1, K); break;                                 \
    }                                                                                                         \
  }                                                                                                           \
  /* BlockManager::setFaceLists + copyVCFromNeighbor (src/BlockManager.cpp:471-530, 715-743)                 \
   * with the VCUpdater::copyFromNeighbor dispatch (include/VCUpdater.h:65-82).                              \
   * mask: BCM_GF_* kind bits and axis bits, as bcm_ghost_fill. */                                           \
  void oracle_ghost_fill_##SUF(const bcm_block_info* blk, int nb, int n, int vc, T* data, int mask) {         \
    geo_t g = mkgeo(n, vc);                                                                                   \
    int axes = (mask >> 3) & 7;                                                                               \
    if (!axes) axes = 7;                                                                                      \
    long nface = (long)nb * 6 * 4;                                                                            \
    _Pragma("omp parallel for schedule(dynamic, 64)")                                                         \
    for (long t = 0; t < nface; ++t) {                                                                        \
      int b = (int)(t / 24), face = (int)(t / 4 % 6), sf = (int)(t % 4);                                      \
      const bcm_block_info* bi = &blk[b];                                                                     \
      if (!(axes & (1 << (face / 2)))) continue;                                                              \
      int ld = bi->nb_level_diff[face];                                                                       \
      int nbid = bi->nb_id[face][sf];                                                                         \
      if (ld == 0 || ld == -1) { if (sf != 0) continue; nbid = bi->nb_id[face][0]; }                           \
      if (nbid < 0) continue;                                                                                 \
      T* d = data + (long)b * g.bvol;                                                                         \
      const T* s = data + (long)nbid * g.bvol;                                                                \
      if (ld == 0) { if (mask & BCM_GF_FLAT) copy_from_neighbor_##SUF(&g, d, s, face); }                      \
      else if (ld == 1) { if (mask & BCM_GF_F2C) f2c_face_##SUF(&g, d, s, face, sf); }                        \
      else { if (mask & BCM_GF_C2F) c2f_face_##SUF(&g, d, s, face, bi->nb_subface[face]); }                    \
    }                                                                                                         \
  }                                                                                                           \
  /* Solver::setDirichletBoundaryInBlock (examples/sor1/Solver.cpp:232-294). */                              \
  void oracle_dirichlet_##SUF(const bcm_block_info* blk, int nb, int n, int vc, T* data, const double* vals) { \
    geo_t g = mkgeo(n, vc);                                                                                   \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      T* f = data + (long)b * g.bvol;                                                                         \
      for (int face = 0; face < 6; ++face) {                                                                  \
        if (blk[b].bc_type[face] != BCM_BC_DIRICHLET) continue;                                               \
        double v = vals[blk[b].bc_id[face]];                                                                  \
        switch (face) {                                                                                       \
          case BCM_X_M: for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j)                                \
              f[IDX(&g, -1, j, k)] = (T)(2.0 * v - f[IDX(&g, 0, j, k)]); break;                               \
          case BCM_X_P: for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j)                                \
              f[IDX(&g, n, j, k)] = (T)(2.0 * v - f[IDX(&g, n - 1, j, k)]); break;                            \
          case BCM_Y_M: for (int k = 0; k < n; ++k) for (int i = 0; i < n; ++i)                                \
              f[IDX(&g, i, -1, k)] = (T)(2.0 * v - f[IDX(&g, i, 0, k)]); break;                               \
          case BCM_Y_P: for (int k = 0; k < n; ++k) for (int i = 0; i < n; ++i)                                \
              f[IDX(&g, i, n, k)] = (T)(2.0 * v - f[IDX(&g, i, n - 1, k)]); break;                            \
          case BCM_Z_M: for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i)                                \
              f[IDX(&g, i, j, -1)] = (T)(2.0 * v - f[IDX(&g, i, j, 0)]); break;                               \
          default: for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i)                                     \
              f[IDX(&g, i, j, n)] = (T)(2.0 * v - f[IDX(&g, i, j, n - 1)]); break;                            \
        }                                                                                                     \
      }                                                                                                       \
    }                                                                                                         \
  }

DEFINE_GHOST(double, d)
DEFINE_GHOST(float, f)

/* BlockScalar3D::ImposeBoundaryCondition (examples/SolverTest/BlockScalar3D.h:
 * 199-206, value scaling :98-100) with the per-face routines of
 * examples/SolverTest/bcx.f90: bc_x1_d/bc_x3_d (:18-76) and the Y/Z forms,
 * bc_x1_n/bc_x3_n (:198-256) and the Y/Z forms.  For each outer face, with
 * u the interior cell next to it and (p, q) the two in-face indices:
 *   Dirichlet: ghost1 = 2*xc - u, ghost2 = xc, xc = v;
 *   Neumann:   ghost1 = u + s*xc, ghost2 = u + s*(xc*2), xc = v*cellSize.x,
 *              s = +1 on a plus face, -1 on a minus face.
 * ghost2 exists only for vc >= 2 (the Fortran writes it unconditionally). */
#define DEFINE_IMPOSE_BC(T, SUF)                                                                              \
  void oracle_impose_bc_##SUF(const bcm_block_info* blk, int nb, int n, int vc, T* data, const double* vals) {  \
    geo_t g = mkgeo(n, vc);                                                                                   \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      T* f = data + (long)b * g.bvol;                                                                         \
      double cell_x = blk[b].block_size[0] / n;                                                               \
      for (int face = 0; face < 6; ++face) {                                                                  \
        int type = blk[b].bc_type[face];                                                                      \
        if (type != BCM_BC_DIRICHLET && type != BCM_BC_NEUMANN) continue;                                     \
        double xc = vals[blk[b].bc_id[face]];                                                                 \
        if (type == BCM_BC_NEUMANN) xc = xc * cell_x;                                                         \
        int axis = face / 2, plus = face & 1;                                                                 \
        for (int q = 0; q < n; ++q)                                                                           \
          for (int p = 0; p < n; ++p) {                                                                       \
            int c[3], c1[3], c2[3];                                                                           \
            int t0 = axis == 0 ? 1 : 0, t1 = axis == 2 ? 1 : 2;                                               \
            c[t0] = p; c[t1] = q;                                                                             \
            c[axis] = plus ? n - 1 : 0;                                                                       \
            memcpy(c1, c, sizeof c); memcpy(c2, c, sizeof c);                                                 \
            c1[axis] = plus ? n : -1;                                                                         \
            c2[axis] = plus ? n + 1 : -2;                                                                     \
            double u = f[IDX(&g, c[0], c[1], c[2])];                                                          \
            double g1, g2;                                                                                    \
            if (type == BCM_BC_DIRICHLET) { g1 = 2.0 * xc - u; g2 = xc; }                                     \
            else if (plus) { g1 = u + xc; g2 = u + xc * 2.0; }                                                \
            else { g1 = u - xc; g2 = u - xc * 2.0; }                                                          \
            f[IDX(&g, c1[0], c1[1], c1[2])] = (T)g1;                                                          \
            if (vc >= 2) f[IDX(&g, c2[0], c2[1], c2[2])] = (T)g2;                                             \
          }                                                                                                   \
      }                                                                                                       \
    }                                                                                                         \
  }

DEFINE_IMPOSE_BC(double, d)
DEFINE_IMPOSE_BC(float, f)

/* ------------------------------------------------------------------ */
/* Red-black SOR                                                        */
/* ------------------------------------------------------------------ */
/* sor1 calcSorInBlock (examples/sor1/Solver.cpp:132-153), coefficients from
 * :98-106; sor2 calcSorInBlock (examples/sor2/Solver.cpp:136-160),
 * coefficients from :99-102.  s == NULL selects the sor1 form. */
void oracle_sor_sweep_d(const bcm_block_info* blk, int nb, int n, int vc, double* data, const double* src,
                        double omega) {
  geo_t g = mkgeo(n, vc);
  for (int b = 0; b < nb; ++b) {
    double* f = data + (long)b * g.bvol;
    const double* s = src ? src + (long)b * g.bvol : NULL;
    double dx = blk[b].block_size[0] / n, dy = blk[b].block_size[1] / n, dz = blk[b].block_size[2] / n;
    double cx = 1.0 / (dx * dx), cy = 1.0 / (dy * dy), cz = 1.0 / (dz * dz);
    if (!s) {
      double c = 2.0 * (cx + cy + cz);
      cx = omega * cx / c;
      cy = omega * cy / c;
      cz = omega * cz / c;
      double c0 = 1.0 - omega;
      for (int col = 0; col < 2; col++) {
#pragma omp parallel for schedule(static, 1)
        for (int k = 0; k < n; k++)
          for (int j = 0; j < n; j++)
            for (int i = (col + j + k) % 2; i < n; i += 2)
              f[IDX(&g, i, j, k)] = c0 * f[IDX(&g, i, j, k)] + cx * (f[IDX(&g, i - 1, j, k)] + f[IDX(&g, i + 1, j, k)]) +
                                    cy * (f[IDX(&g, i, j - 1, k)] + f[IDX(&g, i, j + 1, k)]) +
                                    cz * (f[IDX(&g, i, j, k - 1)] + f[IDX(&g, i, j, k + 1)]);
      }
    } else {
      double c0 = 1.0 / (2.0 * (cx + cy + cz));
      for (int col = 0; col < 2; col++) {
#pragma omp parallel for schedule(static, 1)
        for (int k = 0; k < n; k++)
          for (int j = 0; j < n; j++)
            for (int i = (col + j + k) % 2; i < n; i += 2) {
              double fNew = c0 * (+cx * (f[IDX(&g, i - 1, j, k)] + f[IDX(&g, i + 1, j, k)]) +
                                  cy * (f[IDX(&g, i, j - 1, k)] + f[IDX(&g, i, j + 1, k)]) +
                                  cz * (f[IDX(&g, i, j, k - 1)] + f[IDX(&g, i, j, k + 1)]) - s[IDX(&g, i, j, k)]);
              f[IDX(&g, i, j, k)] = omega * fNew + (1.0 - omega) * f[IDX(&g, i, j, k)];
            }
      }
    }
  }
}

/* Solver::run (examples/sor1/Solver.cpp:74-129, examples/sor2/Solver.cpp:
 * 77-133): per outer iteration, on every block `inner` times (nLoopInner)
 * SOR + in-block BC, the ghosts of other blocks staying as the last exchange
 * left them; then beginUpdateVC, setBoundaryCondition, endUpdateVC.  A sweep
 * writes only its block's interior and the in-block BC only its block's
 * Dirichlet ghosts, so all blocks' inner loops in turn equal the reference's
 * block-by-block order. */
void oracle_sor_iterate_inner_d(const bcm_block_info* blk, int nb, int n, int vc, double* data, const double* src,
                                double omega, const double* vals, int iters, int inner) {
  for (int it = 0; it < iters; ++it) {
    for (int in = 0; in < inner; ++in) {
      oracle_sor_sweep_d(blk, nb, n, vc, data, src, omega);
      oracle_dirichlet_d(blk, nb, n, vc, data, vals);
    }
    oracle_ghost_fill_d(blk, nb, n, vc, data, BCM_GF_ALL);
    oracle_dirichlet_d(blk, nb, n, vc, data, vals);
  }
}

/* nLoopInner == 1. */
void oracle_sor_iterate_d(const bcm_block_info* blk, int nb, int n, int vc, double* data, const double* src,
                          double omega, const double* vals, int iters) {
  oracle_sor_iterate_inner_d(blk, nb, n, vc, data, src, omega, vals, iters, 1);
}

/* ------------------------------------------------------------------ */
/* SolverTest per-block kernels (examples/SolverTest/blas.f90)          */
/* A[7] = {Ap, Aw, Ae, As, An, Ab, At}; OpenMP over blocks as ILS3D.    */
/* ------------------------------------------------------------------ */
#define DEFINE_BLAS(T, SUF)                                                                                   \
  /* jacobi_smoother (blas.f90:452-502); with a mask, jacobi_smoother_mask                                    \
   * (blas.f90:915-967): only cells whose mask is 1 are written. */                                         \
  void oracle_jacobi7_mask_##SUF(int nb, int n, int vc, T* x1, const T* x0, T* const A[7], const T* bv,        \
                                 double w, const T* mask) {                                                   \
    geo_t g = mkgeo(n, vc);                                                                                   \
    const T omega = (T)w;                                                                                     \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        T r = bv[o] - A[1][o] * x0[o - 1] - A[2][o] * x0[o + 1] - A[3][o] * x0[o - g.nx0] -                   \
              A[4][o] * x0[o + g.nx0] - A[5][o] * x0[o - (long)g.nx0 * g.nx0] -                               \
              A[6][o] * x0[o + (long)g.nx0 * g.nx0];                                                          \
        T xt = r / A[0][o];                                                                                   \
        x1[o] = x0[o] - omega * (x0[o] - xt);                                                                 \
      }                                                                                                       \
    }                                                                                                         \
  }                                                                                                           \
  /* calc_ax (blas.f90:556-600); calc_ax_mask (:994-1030). */                                                \
  void oracle_calc_ax_mask_##SUF(int nb, int n, int vc, T* ax, T* const A[7], const T* x, const T* mask) {    \
    geo_t g = mkgeo(n, vc);                                                                                   \
    const long sy = g.nx0, sz = (long)g.nx0 * g.nx0;                                                          \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        ax[o] = +A[1][o] * x[o - 1] + A[2][o] * x[o + 1] + A[3][o] * x[o - sy] + A[4][o] * x[o + sy] +        \
                A[5][o] * x[o - sz] + A[6][o] * x[o + sz] + A[0][o] * x[o];                                   \
      }                                                                                                       \
    }                                                                                                         \
  }                                                                                                           \
  /* calc_r (blas.f90:602-648); calc_r_mask (:1032-1070). */                                                 \
  void oracle_calc_r_mask_##SUF(int nb, int n, int vc, T* r, T* const A[7], const T* x, const T* bv,           \
                                const T* mask) {                                                              \
    geo_t g = mkgeo(n, vc);                                                                                   \
    const long sy = g.nx0, sz = (long)g.nx0 * g.nx0;                                                          \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        r[o] = bv[o] - A[1][o] * x[o - 1] - A[2][o] * x[o + 1] - A[3][o] * x[o - sy] - A[4][o] * x[o + sy] -  \
               A[5][o] * x[o - sz] - A[6][o] * x[o + sz] - A[0][o] * x[o];                                    \
      }                                                                                                       \
    }                                                                                                         \
  }                                                                                                           \
  /* calc_r2 + CalcR2 (blas.f90:650-698, ILS3D.h:168-217): per-block sum in T,                               \
   * blocks summed in block order (the reference's OpenMP order varies). */                                  \
  double oracle_calc_r2_mask_##SUF(int nb, int n, int vc, T* const A[7], const T* x, const T* bv,              \
                                   const T* mask) {                                                           \
    geo_t g = mkgeo(n, vc);                                                                                   \
    const long sy = g.nx0, sz = (long)g.nx0 * g.nx0;                                                          \
    double* part = (double*)malloc(sizeof(double) * (size_t)nb);                                              \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      T rr = 0;                                                                                               \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        T r = bv[o] - A[1][o] * x[o - 1] - A[2][o] * x[o + 1] - A[3][o] * x[o - sy] - A[4][o] * x[o + sy] -   \
              A[5][o] * x[o - sz] - A[6][o] * x[o + sz] - A[0][o] * x[o];                                     \
        rr = rr + r * r;                                                                                      \
      }                                                                                                       \
      part[b] = rr;                                                                                           \
    }                                                                                                         \
    double s = 0.0;                                                                                           \
    for (int b = 0; b < nb; ++b) s += part[b];                                                                \
    free(part);                                                                                               \
    return s;                                                                                                 \
  }                                                                                                           \
  /* dot (blas.f90:379-412) + ILS3D::DOT (ILS3D.h:272-311). */                                               \
  double oracle_dot_mask_##SUF(int nb, int n, int vc, const T* x, const T* y, const T* mask) {                 \
    geo_t g = mkgeo(n, vc);                                                                                   \
    double* part = (double*)malloc(sizeof(double) * (size_t)nb);                                              \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      T xy = 0;                                                                                               \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        xy = xy + x[o] * y[o];                                                                                \
      }                                                                                                       \
      part[b] = xy;                                                                                           \
    }                                                                                                         \
    double s = 0.0;                                                                                           \
    for (int b = 0; b < nb; ++b) s += part[b];                                                                \
    free(part);                                                                                               \
    return s;                                                                                                 \
  }                                                                                                           \
  /* axpy / xpay / axpyz / axpbypz (blas.f90:242-377), interior cells. */                                    \
  void oracle_blas_mask_##SUF(int nb, int n, int vc, int op, T* z, const T* x, const T* y, double a,           \
                              double bc, const T* mask) {                                                     \
    geo_t g = mkgeo(n, vc);                                                                                   \
    const T ta = (T)a, tb = (T)bc;                                                                            \
    _Pragma("omp parallel for schedule(static)")                                                              \
    for (int b = 0; b < nb; ++b) {                                                                            \
      long o0 = (long)b * g.bvol;                                                                             \
      for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {                  \
        long o = o0 + IDX(&g, i, j, k);                                                                       \
        if (mask && mask[o] != (T)1) continue;                                                                \
        if (op == 0) z[o] = z[o] + ta * x[o];                                                                 \
        else if (op == 1) z[o] = x[o] + ta * z[o];                                                            \
        else if (op == 2) z[o] = ta * x[o] + y[o];                                                            \
        else if (op == 3) z[o] = ta * x[o] + tb * y[o] + z[o];                                                \
        else z[o] = x[o];                                                                                     \
      }                                                                                                       \
    }                                                                                                         \
  }                                                                                                           \
  void oracle_jacobi7_##SUF(int nb, int n, int vc, T* x1, const T* x0, T* const A[7], const T* bv, double w) { \
    oracle_jacobi7_mask_##SUF(nb, n, vc, x1, x0, A, bv, w, NULL);                                             \
  }                                                                                                           \
  void oracle_calc_ax_##SUF(int nb, int n, int vc, T* ax, T* const A[7], const T* x) {                        \
    oracle_calc_ax_mask_##SUF(nb, n, vc, ax, A, x, NULL);                                                     \
  }                                                                                                           \
  void oracle_calc_r_##SUF(int nb, int n, int vc, T* r, T* const A[7], const T* x, const T* bv) {             \
    oracle_calc_r_mask_##SUF(nb, n, vc, r, A, x, bv, NULL);                                                   \
  }                                                                                                           \
  double oracle_calc_r2_##SUF(int nb, int n, int vc, T* const A[7], const T* x, const T* bv) {                 \
    return oracle_calc_r2_mask_##SUF(nb, n, vc, A, x, bv, NULL);                                              \
  }                                                                                                           \
  double oracle_dot_##SUF(int nb, int n, int vc, const T* x, const T* y) {                                    \
    return oracle_dot_mask_##SUF(nb, n, vc, x, y, NULL);                                                      \
  }                                                                                                           \
  void oracle_blas_##SUF(int nb, int n, int vc, int op, T* z, const T* x, const T* y, double a, double bc) {   \
    oracle_blas_mask_##SUF(nb, n, vc, op, z, x, y, a, bc, NULL);                                              \
  }

DEFINE_BLAS(double, d)
DEFINE_BLAS(float, f)

/* Poisson assembly without cut cells (examples/SolverTest/bcut.f90:2848-2865):
 * Aw..At = 1/h^2, Ap = -(l0+l1)-(l2+l3)-(l4+l5); then the outer-face folds
 * in X_M..Z_P order (examples/SolverTest/BlockScalar3D.h:207-214):
 * Dirichlet bc_A*_d (bca.f90:12-202): Ap -= Anb; b -= Anb*xc*2; Anb = 0;
 * Neumann bc_A*_n (bca.f90:204-394): Ap += Anb; b += Anb*xc on a minus face,
 * b -= Anb*xc on a plus face, xc = v*cellSize.x; Anb = 0.
 * A may be NULL (b only). */
void oracle_poisson_assemble_d(const bcm_block_info* blk, int nb, int n, int vc, double* const A[7], double* bout,
                               const double* rhs, const double* vals) {
  geo_t g = mkgeo(n, vc);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    double dx = blk[b].block_size[0] / n, dy = blk[b].block_size[1] / n, dz = blk[b].block_size[2] / n;
    double l[6] = {1.0 / (dx * dx), 1.0 / (dx * dx), 1.0 / (dy * dy), 1.0 / (dy * dy), 1.0 / (dz * dz),
                   1.0 / (dz * dz)};
    long o0 = (long)b * g.bvol;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          long o = o0 + IDX(&g, i, j, k);
          double a[6];
          for (int f = 0; f < 6; ++f) a[f] = l[f];
          double ap = -(a[0] + a[1]) - (a[2] + a[3]) - (a[4] + a[5]);
          double bv = rhs[o];
          int at[6] = {i == 0, i == n - 1, j == 0, j == n - 1, k == 0, k == n - 1};
          for (int f = 0; f < 6; ++f) {
            if (!at[f]) continue;
            if (blk[b].bc_type[f] == BCM_BC_DIRICHLET) {
              ap = ap - a[f];
              bv = bv - a[f] * vals[blk[b].bc_id[f]] * 2.0;
              a[f] = 0.0;
            } else if (blk[b].bc_type[f] == BCM_BC_NEUMANN) {
              double xc = vals[blk[b].bc_id[f]] * dx;
              ap = ap + a[f];
              if (f & 1) bv = bv - a[f] * xc;
              else bv = bv + a[f] * xc;
              a[f] = 0.0;
            }
          }
          bout[o] = bv;
          if (A) {
            A[0][o] = ap;
            for (int f = 0; f < 6; ++f) A[f + 1][o] = a[f];
          }
        }
  }
}

/* ILS3D::BiCGSTAB (examples/SolverTest/ILS3D.h:663-844) with the Jacobi
 * preconditioner (:65-126, :452-475), FP64, stored A, ghost fill as the
 * boundary step (the operator carries the folded Dirichlet condition).
 * w[9] = {r, r0, p, p_, q_, s, s_, t_, x0}.  Returns 0 when converged. */
static void swapp(double** a, double** b) {
  double* t = *a;
  *a = *b;
  *b = t;
}

static void jacobi_pre(const bcm_block_info* blk, int nb, int n, int vc, double** x, double** x0,
                       double* const A[7], const double* bv, double omega, int count, const double* mask) {
  for (int c = 0; c < count; ++c) {
    oracle_jacobi7_mask_d(nb, n, vc, *x0, *x, A, bv, omega, mask);
    swapp(x, x0);
    oracle_ghost_fill_d(blk, nb, n, vc, *x, BCM_GF_ALL);
  }
}

/* With a mask, ILS3D::BiCGSTAB_Mask (ILS3D.h:901-1097): the dot products, the
 * preconditioner sweeps and the products A p_, A s_ visit only the cells
 * whose mask is 1; the initial residual and the vector updates do not. */
static int bicgstab_impl(const bcm_block_info* blk, int nb, int n, int vc, double* x, double* const A[7],
                         const double* bv, double* w[9], double omega, int count_pre, int count_max, double eps,
                         int* count_out, double* residual, const double* mask) {
  double *r = w[0], *r0 = w[1], *p = w[2], *p_ = w[3], *q_ = w[4], *s = w[5], *s_ = w[6], *t_ = w[7], *x0 = w[8];
  size_t total = (size_t)nb * (size_t)mkgeo(n, vc).bvol;
  double bb = oracle_dot_mask_d(nb, n, vc, bv, bv, mask);
  if (fabs(bb) < FLT_MIN) {
    *residual = 0.0;
    *count_out = 0;
    return 0;
  }
  oracle_calc_r_d(nb, n, vc, r, A, x, bv);
  oracle_blas_d(nb, n, vc, 4, r0, r, NULL, 0.0, 0.0);
  double rr0 = 1.0, alpha = 0.0, gamma = 1.0;
  int count, status = 1;
  for (count = 1; count <= count_max; ++count) {
    double rr1 = oracle_dot_mask_d(nb, n, vc, r, r0, mask);
    if (fabs(rr1) < FLT_MIN) {
      *residual = rr1;
      count = 0;
      status = 0;
      break;
    }
    if (count == 1) {
      oracle_blas_d(nb, n, vc, 4, p, r, NULL, 0.0, 0.0);
    } else {
      double beta = rr1 / rr0 * alpha / gamma;
      oracle_blas_d(nb, n, vc, 0, p, q_, NULL, -gamma, 0.0);
      oracle_blas_d(nb, n, vc, 1, p, r, NULL, beta, 0.0);
    }
    oracle_ghost_fill_d(blk, nb, n, vc, p, BCM_GF_ALL);
    memset(p_, 0, total * sizeof(double));
    jacobi_pre(blk, nb, n, vc, &p_, &x0, A, p, omega, count_pre, mask);
    oracle_calc_ax_mask_d(nb, n, vc, q_, A, p_, mask);
    double q_r0 = oracle_dot_mask_d(nb, n, vc, q_, r0, mask);
    alpha = rr1 / q_r0;
    oracle_blas_d(nb, n, vc, 2, s, q_, r, -alpha, 0.0);
    oracle_ghost_fill_d(blk, nb, n, vc, s, BCM_GF_ALL);
    memset(s_, 0, total * sizeof(double));
    jacobi_pre(blk, nb, n, vc, &s_, &x0, A, s, omega, count_pre, mask);
    oracle_calc_ax_mask_d(nb, n, vc, t_, A, s_, mask);
    double t_s = oracle_dot_mask_d(nb, n, vc, t_, s, mask);
    double t_t = oracle_dot_mask_d(nb, n, vc, t_, t_, mask);
    gamma = t_s / t_t;
    oracle_blas_d(nb, n, vc, 3, x, p_, s_, alpha, gamma);
    oracle_blas_d(nb, n, vc, 2, r, t_, s, -gamma, 0.0);
    double rr = oracle_dot_mask_d(nb, n, vc, r, r, mask);
    double res = (fabs(bb) > FLT_MIN) ? fabs(rr) / fabs(bb) : fabs(rr);
    *residual = res;
    if (isnan(res) || isinf(res)) return 2;
    if (res <= eps * eps) {
      status = 0;
      break;
    }
    rr0 = rr1;
  }
  if (count > count_max) count = count_max;
  oracle_ghost_fill_d(blk, nb, n, vc, x, BCM_GF_ALL);
  *count_out = count;
  w[0] = r; w[1] = r0; w[2] = p; w[3] = p_; w[4] = q_; w[5] = s; w[6] = s_; w[7] = t_; w[8] = x0;
  return status;
}

int oracle_bicgstab_d(const bcm_block_info* blk, int nb, int n, int vc, double* x, double* const A[7],
                      const double* bv, double* w[9], double omega, int count_pre, int count_max, double eps,
                      int* count_out, double* residual) {
  return bicgstab_impl(blk, nb, n, vc, x, A, bv, w, omega, count_pre, count_max, eps, count_out, residual, NULL);
}

int oracle_bicgstab_mask_d(const bcm_block_info* blk, int nb, int n, int vc, double* x, double* const A[7],
                           const double* bv, double* w[9], double omega, int count_pre, int count_max, double eps,
                           int* count_out, double* residual, const double* mask) {
  return bicgstab_impl(blk, nb, n, vc, x, A, bv, w, omega, count_pre, count_max, eps, count_out, residual, mask);
}

/* ILS3D::Jacobi (examples/SolverTest/ILS3D.h:478-529), FP64, stored A; with
 * a mask ILS3D::Jacobi_Mask (:846-899), where only the sweep is masked. */
static int jacobi_solve_impl(const bcm_block_info* blk, int nb, int n, int vc, double** x, double* const A[7],
                             const double* bv, double** x0, double omega, int count_max, double eps,
                             int* count_out, double* residual, const double* mask) {
  double bb = oracle_dot_d(nb, n, vc, bv, bv);
  int c, status = 1;
  for (c = 1; c <= count_max; ++c) {
    oracle_jacobi7_mask_d(nb, n, vc, *x0, *x, A, bv, omega, mask);
    swapp(x, x0);
    oracle_ghost_fill_d(blk, nb, n, vc, *x, BCM_GF_ALL);
    double rr = oracle_calc_r2_d(nb, n, vc, A, *x, bv);
    double res = (fabs(bb) > FLT_MIN) ? fabs(rr) / fabs(bb) : fabs(rr);
    *residual = res;
    if (isnan(res) || isinf(res)) return 2;
    if (res <= eps * eps) {
      status = 0;
      break;
    }
  }
  *count_out = c > count_max ? count_max : c;
  oracle_ghost_fill_d(blk, nb, n, vc, *x, BCM_GF_ALL);
  return status;
}

int oracle_jacobi_solve_d(const bcm_block_info* blk, int nb, int n, int vc, double** x, double* const A[7],
                          const double* bv, double** x0, double omega, int count_max, double eps, int* count_out,
                          double* residual) {
  return jacobi_solve_impl(blk, nb, n, vc, x, A, bv, x0, omega, count_max, eps, count_out, residual, NULL);
}

int oracle_jacobi_solve_mask_d(const bcm_block_info* blk, int nb, int n, int vc, double** x, double* const A[7],
                               const double* bv, double** x0, double omega, int count_max, double eps,
                               int* count_out, double* residual, const double* mask) {
  return jacobi_solve_impl(blk, nb, n, vc, x, A, bv, x0, omega, count_max, eps, count_out, residual, mask);
}
