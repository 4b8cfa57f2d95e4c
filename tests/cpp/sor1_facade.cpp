// A sor1-style driver written against the BlockManager facade
// (include/bcm/), checking the drop-in contract end to end:
//  1. host-resident: the driver's own per-block loops on getData()
//     pointers (the structure of examples/sor1/Solver.cpp:74-153), with
//     BlockManager::beginUpdateVC / endUpdateVC doing the ghost fill on
//     the GPU;
//  2. device-resident: the same iteration as one bcm_sor_iterate call.
// Both must reproduce examples/sor1/log.4:54 (errorMax = 0.00353051) and
// agree bit for bit.
#include <cmath>
#include <cstdio>
#include <vector>

#include "Block.h"
#include "BlockManager.h"
#include "Scalar3D.h"
#include "Scalar3DUpdater.h"

static std::vector<double> g_bval;

static void dirichlet_in_block(int n, double* f, Index3DS idx, const Block* blk) {
  const BoundaryInfo* bi = blk->getBoundaryInfo();
  for (int face = 0; face < NUM_FACE; ++face) {
    if (bi[face].getType() != BoundaryInfo::DIRICHLET) continue;
    const double v = g_bval[bi[face].getID()];
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < n; ++c) {
        switch (face) {
          case X_M: f[idx(-1, a, c)] = 2.0 * v - f[idx(0, a, c)]; break;
          case X_P: f[idx(n, a, c)] = 2.0 * v - f[idx(n - 1, a, c)]; break;
          case Y_M: f[idx(a, -1, c)] = 2.0 * v - f[idx(a, 0, c)]; break;
          case Y_P: f[idx(a, n, c)] = 2.0 * v - f[idx(a, n - 1, c)]; break;
          case Z_M: f[idx(a, c, -1)] = 2.0 * v - f[idx(a, c, 0)]; break;
          default: f[idx(a, c, n)] = 2.0 * v - f[idx(a, c, n - 1)]; break;
        }
      }
  }
}

static void set_bc(BlockManager& bm, int id, int n) {
  for (int b = 0; b < bm.getNumBlock(); ++b) {
    Block* blk = dynamic_cast<Block*>(bm.getBlock(b));
    Scalar3D<double>* f = dynamic_cast<Scalar3D<double>*>(blk->getDataClass(id));
    dirichlet_in_block(n, f->getData(), f->getIndex(), blk);
  }
}

static double error_max(BlockManager& bm, int id, int n) {
  double e = 0.0;
  for (int b = 0; b < bm.getNumBlock(); ++b) {
    BlockBase* blk = bm.getBlock(b);
    Scalar3D<double>* f = dynamic_cast<Scalar3D<double>*>(blk->getDataClass(id));
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double x = blk->getOrigin()[0] + (i + 0.5) * blk->getCellSize()[0];
          e = std::max(e, std::fabs((*f)(i, j, k) - x));
        }
  }
  return e;
}

int main() {
  // examples/sor1/test.conf: simple divider levels 0..3, type x, size 2,
  // vc 1, omega 1.2, 400 outer x 1 inner; b0 = 0, b1 = 1.
  bcm_tree_params p;
  std::memset(&p, 0, sizeof(p));
  p.root_n[0] = p.root_n[1] = p.root_n[2] = 1;
  p.periodic[1] = p.periodic[2] = 1;
  p.root_length = 1.0;
  p.divider = BCM_DIV_SIMPLE;
  p.min_level = 0;
  p.max_level = 3;
  p.ordering = BCM_ORDER_Z;
  p.bc_setter = BCM_BCSET_SOR1_X;
  bcm_tree_t tree = nullptr;
  if (bcm_tree_build(&p, &tree) != BCM_OK) return 2;
  const int nb = bcm_tree_num_blocks(tree);
  std::vector<bcm_block_info> info(nb);
  bcm_tree_get_blocks(tree, info.data(), nb);
  bcm_tree_destroy(tree);
  const int n = 2, vc = 1, iters = 400;
  const double omega = 1.2;
  g_bval = {0.0, 1.0};

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
  bm.printBlockLayoutInfo();

  // 1. host-resident driver.
  const int id = bm.setDataClass<Scalar3D<double>, Scalar3DUpdater<double> >(vc);
  bm.prepareForVCUpdate(id, 100, false);
  bm.beginUpdateVC(id);
  set_bc(bm, id, n);
  bm.endUpdateVC(id);
  for (int it = 0; it < iters; ++it) {
    for (int b = 0; b < bm.getNumBlock(); ++b) {
      Block* blk = dynamic_cast<Block*>(bm.getBlock(b));
      Scalar3D<double>* sf = dynamic_cast<Scalar3D<double>*>(blk->getDataClass(id));
      double* f = sf->getData();
      Index3DS idx = sf->getIndex();
      const Vec3r& cs = blk->getCellSize();
      double cx = 1.0 / (cs.x * cs.x), cy = 1.0 / (cs.y * cs.y), cz = 1.0 / (cs.z * cs.z);
      const double c = 2.0 * (cx + cy + cz);
      cx = omega * cx / c;
      cy = omega * cy / c;
      cz = omega * cz / c;
      const double c0 = 1.0 - omega;
      for (int col = 0; col < 2; col++)
        for (int k = 0; k < n; k++)
          for (int j = 0; j < n; j++)
            for (int i = (col + j + k) % 2; i < n; i += 2)
              f[idx(i, j, k)] = c0 * f[idx(i, j, k)] + cx * (f[idx(i - 1, j, k)] + f[idx(i + 1, j, k)]) +
                                cy * (f[idx(i, j - 1, k)] + f[idx(i, j + 1, k)]) +
                                cz * (f[idx(i, j, k - 1)] + f[idx(i, j, k + 1)]);
      dirichlet_in_block(n, f, idx, blk);
    }
    bm.beginUpdateVC(id);
    set_bc(bm, id, n);
    bm.endUpdateVC(id);
  }
  const double e_host = error_max(bm, id, n);

  // 2. device-resident: the same iteration on the GPU.
  const int id2 = bm.setDataClass<Scalar3D<double>, Scalar3DUpdater<double> >(vc);
  bm.prepareForVCUpdate(id2, 100, false);
  bm.setDeviceResident(id2, true);
  bcm_field_t d = bm.deviceField(id2);
  if (bcm_ghost_fill(d, BCM_GF_ALL) != BCM_OK) return 3;
  if (bcm_dirichlet_mirror(d, g_bval.data(), 2) != BCM_OK) return 3;
  if (bcm_sor_iterate(d, nullptr, omega, g_bval.data(), 2, iters) != BCM_OK) return 3;
  bm.syncToHost(id2);
  const double e_dev = error_max(bm, id2, n);

  // Compare interiors and face ghosts.
  long diffs = 0;
  for (int b = 0; b < nb; ++b) {
    Scalar3D<double>* a = dynamic_cast<Scalar3D<double>*>(bm.getBlock(b)->getDataClass(id));
    Scalar3D<double>* c = dynamic_cast<Scalar3D<double>*>(bm.getBlock(b)->getDataClass(id2));
    for (int k = -vc; k < n + vc; ++k)
      for (int j = -vc; j < n + vc; ++j)
        for (int i = -vc; i < n + vc; ++i) {
          const int out = (i < 0 || i >= n) + (j < 0 || j >= n) + (k < 0 || k >= n);
          if (out <= 1 && (*a)(i, j, k) != (*c)(i, j, k)) diffs++;
        }
  }
  std::printf("errorMax host-resident = %.9g\nerrorMax device-resident = %.9g\ndiffering cells = %ld\n", e_host,
              e_dev, diffs);
  const bool ok = std::fabs(e_host - 0.00353051) < 5e-9 && e_host == e_dev && diffs == 0;
  std::printf("%s\n", ok ? "PASS" : "FAIL");
  return ok ? 0 : 1;
}
