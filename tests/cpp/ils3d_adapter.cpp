// SolverTest's solver interface on the GPU: LocalScalar3D fields
// (include/bcm/LocalScalar3D.h) and ILS3D's methods (include/bcm/ILS3D.h)
// on an adaptive mesh whose domain faces carry Dirichlet and Neumann
// conditions.
//
// The driver assembles the 7-point Laplacian (Ap = -2 sum 1/h^2, Anb = 1/h^2)
// and a right-hand side on the host, folds the boundary conditions into
// them with LocalScalar3D::ImposeBoundaryCondition(A, b), then solves with
// ILS3D::BiCGSTAB, BiCGSTAB_Mask (setup_mask), CG and Jacobi, and with the
// library's bcm_bicgstab on the same fields.  It writes, into the working
// directory, the block table, the folded system and every solution
// (raw float64, block after block) and one line per solver to stats.txt:
//   <name> <count> <residual>
// The checks against the oracle are made by tests/test_gpu_ils3d.py.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "Block.h"
#include "BlockManager.h"
#include "ILS3D.h"
#include "LocalScalar3D.h"
#include "Scalar3D.h"

typedef LocalScalar3D<double> LS;

static void write_file(const char* name, const void* p, size_t bytes) {
  FILE* f = std::fopen(name, "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
    std::printf("cannot write %s\n", name);
    std::exit(3);
  }
  std::fclose(f);
}

// The host arena of a data class, block after block.
static double* host_data(BlockManager& bm, int id) {
  return dynamic_cast<Scalar3D<double>*>(bm.getBlock(0)->getDataClass(id))->getData();
}

int main() {
  const int n = 8, vc = 1;
  bcm_tree_params p;
  std::memset(&p, 0, sizeof(p));
  p.root_n[0] = p.root_n[1] = p.root_n[2] = 1;
  p.root_length = 1.0;
  p.divider = BCM_DIV_SIMPLE;
  p.min_level = 2;
  p.max_level = 3;
  p.ordering = BCM_ORDER_HILBERT;
  p.bc_setter = BCM_BCSET_PER_FACE;
  const int bc_lib[6] = {BCM_BC_DIRICHLET, BCM_BC_DIRICHLET, BCM_BC_NEUMANN,
                         BCM_BC_NEUMANN,   BCM_BC_DIRICHLET, BCM_BC_NEUMANN};
  for (int f = 0; f < 6; ++f) p.face_bc_type[f] = bc_lib[f];
  bcm_tree_t tree = nullptr;
  if (bcm_tree_build(&p, &tree) != BCM_OK) return 2;
  const int nb = bcm_tree_num_blocks(tree);
  std::vector<bcm_block_info> info(nb);
  bcm_tree_get_blocks(tree, info.data(), nb);
  bcm_tree_destroy(tree);
  write_file("blocks.bin", info.data(), info.size() * sizeof(bcm_block_info));

  BlockManager& bm = BlockManager::getInstance();
  for (int b = 0; b < nb; ++b) {
    NeighborInfo* ni = new NeighborInfo[NUM_FACE];
    BoundaryInfo* bi = new BoundaryInfo[NUM_FACE];
    for (int f = 0; f < NUM_FACE; ++f) {
      ni[f].setLevelDifference(info[b].nb_level_diff[f]);
      for (int s = 0; s < NUM_SUBFACE; ++s) ni[f].setID(Subface(s), info[b].nb_id[f][s]);
      ni[f].setNeighborSubface(Subface(info[b].nb_subface[f]));
      bi[f].setType(BoundaryInfo::Type(info[b].bc_type[f]));
      bi[f].setID(info[b].bc_id[f]);
    }
    Vec3r origin(info[b].origin[0], info[b].origin[1], info[b].origin[2]);
    Vec3r bsize(info[b].block_size[0], info[b].block_size[1], info[b].block_size[2]);
    bm.registerBlock(new Block(Vec3i(n, n, n), origin, bsize, info[b].level, ni, bi));
  }
  bm.endRegisterBlock();

  // SolverTest's boundary description: BlockScalar3D types per domain face.
  int types[6] = {LS::DIRICHLET, LS::DIRICHLET, LS::NEUMANN, LS::NEUMANN, LS::DIRICHLET, LS::NEUMANN};
  double values[6] = {1.0, -0.5, 0.25, -0.125, 0.75, 0.5};
  auto field = [&]() { return new LS(bm, vc, "AtOnce", types, values); };
  LS *Ap = field(), *Aw = field(), *Ae = field(), *As = field(), *An = field(), *Ab = field(), *At = field();
  LS *b = field(), *x = field(), *r = field(), *r0 = field(), *pp = field(), *p_ = field(), *q_ = field();
  LS *s = field(), *s_ = field(), *t_ = field(), *x0 = field(), *q = field(), *z = field();
  int mtypes[6] = {LS::INNER, LS::INNER, LS::INNER, LS::INNER, LS::INNER, LS::INNER};
  LocalScalar3D<int> mask(bm, vc, "AtOnce", mtypes, nullptr);
  mask.SetupMask(bm);

  // Laplacian and right-hand side on the host (bcut.f90:2857-2859 form).
  const int w = n + 2 * vc;
  const size_t bvol = size_t(w) * w * w;
  LS* A[7] = {Ap, Aw, Ae, As, An, Ab, At};
  std::vector<double*> hA(7);
  for (int i = 0; i < 7; ++i) hA[i] = host_data(bm, A[i]->GetID());
  double* hb = host_data(bm, b->GetID());
  for (int bk = 0; bk < nb; ++bk) {
    const double h[3] = {info[bk].block_size[0] / n, info[bk].block_size[1] / n, info[bk].block_size[2] / n};
    const double l[3] = {1.0 / (h[0] * h[0]), 1.0 / (h[1] * h[1]), 1.0 / (h[2] * h[2])};
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const size_t o = bk * bvol + (i + vc) + w * ((j + vc) + size_t(w) * (k + vc));
          hA[0][o] = -(l[0] + l[0]) - (l[1] + l[1]) - (l[2] + l[2]);
          hA[1][o] = hA[2][o] = l[0];
          hA[3][o] = hA[4][o] = l[1];
          hA[5][o] = hA[6][o] = l[2];
          const double cx = info[bk].origin[0] + (i + 0.5) * h[0];
          const double cy = info[bk].origin[1] + (j + 0.5) * h[1];
          const double cz = info[bk].origin[2] + (k + 0.5) * h[2];
          hb[o] = cx * cy - 0.5 * cz;
        }
  }
  for (int i = 0; i < 7; ++i) bm.syncToDevice(A[i]->GetID());
  bm.syncToDevice(b->GetID());
  x->ImposeBoundaryCondition(bm, Ap, Aw, Ae, As, An, Ab, At, b);
  for (int i = 0; i < 7; ++i) bm.syncToHost(A[i]->GetID());
  bm.syncToHost(b->GetID());
  {
    std::vector<double> all(7 * nb * bvol);
    for (int i = 0; i < 7; ++i) std::memcpy(&all[i * nb * bvol], hA[i], nb * bvol * sizeof(double));
    write_file("A.bin", all.data(), all.size() * sizeof(double));
    write_file("b.bin", hb, nb * bvol * sizeof(double));
  }

  FILE* st = std::fopen("stats.txt", "w");
  ILS3D ils;
  const double omega = 0.9;
  const int pre = 4;
  auto dump = [&](LS* f, const char* name) {
    bm.syncToHost(f->GetID());
    write_file(name, host_data(bm, f->GetID()), nb * bvol * sizeof(double));
  };
  int count = 0;
  double residual = 0.0;

  ils.Fill(bm, x, 0.0);
  ils.BiCGSTAB(bm, x, Ap, Aw, Ae, As, An, Ab, At, b, r, r0, pp, p_, q_, s, s_, t_, x0, omega, pre, 400, 1e-12, count,
               residual);
  std::fprintf(st, "bicgstab %d %.17g\n", count, residual);
  dump(x, "x_bicgstab.bin");

  ils.Fill(bm, x, 0.0);
  ils.BiCGSTAB_Mask(bm, x, Ap, Aw, Ae, As, An, Ab, At, b, r, r0, pp, p_, q_, s, s_, t_, x0, &mask, omega, pre, 400,
                    1e-12, count, residual);
  std::fprintf(st, "bicgstab_mask %d %.17g\n", count, residual);
  dump(x, "x_bicgstab_mask.bin");

  ils.Fill(bm, x, 0.0);
  ils.CG(bm, x, Ap, Aw, Ae, As, An, Ab, At, b, r, pp, q, z, x0, omega, pre, 1000, 1e-12, count, residual);
  std::fprintf(st, "cg %d %.17g\n", count, residual);
  dump(x, "x_cg.bin");

  ils.Fill(bm, x, 0.0);
  ils.Jacobi(bm, x, Ap, Aw, Ae, As, An, Ab, At, b, x0, 0.8, 60, 1e-14, count, residual);
  std::fprintf(st, "jacobi %d %.17g\n", count, residual);
  dump(x, "x_jacobi.bin");

  // The library's own BiCGSTAB on the same fields.
  {
    ils.Fill(bm, x, 0.0);
    const bcm_field_t Af[7] = {Ap->field(), Aw->field(), Ae->field(), As->field(), An->field(), Ab->field(),
                               At->field()};
    const bcm_field_t work[9] = {r->field(), r0->field(), pp->field(), p_->field(), q_->field(),
                                 s->field(), s_->field(), t_->field(), x0->field()};
    bcm_krylov_params kp;
    kp.omega = omega;
    kp.count_pre = pre;
    kp.count_max = 400;
    kp.epsilon = 1e-12;
    const int rc = bcm_bicgstab(x->field(), Af, b->field(), work, &kp, values, 6, &count, &residual);
    std::fprintf(st, "library_bicgstab %d %.17g %d\n", count, residual, rc);
    dump(x, "x_library.bin");
  }
  std::fclose(st);
  std::printf("blocks %d\nDONE\n", nb);
  return 0;
}
