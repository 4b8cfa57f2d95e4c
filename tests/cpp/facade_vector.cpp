// The facade's mesh construction and exchange paths beyond sor1, end to end
// on the GPU:
//  1. BCMOctree with a driver-defined Divider (refine near a sphere
//     surface) in Hilbert order, Partition, BlockFactory and a
//     driver-defined BoundaryConditionSetter (all faces Dirichlet id 0);
//  2. Vector3D<double> / Vector3DUpdater<double>: updateVC on the vector
//     field equals, component by component, updateVC on a Scalar3D holding
//     that component;
//  3. per-face exchange: VCUpdater::copyFromNeighbor(face, subface) called
//     for every face of every block reproduces updateVC, for the scalar and
//     the vector field, host-resident and device-resident; so do the
//     buffer rules for the scalar field, as two ranks would run them: the
//     sender's copyToCommBuffer(face, subface) into a host buffer attached
//     with getSendBufferPointerSetter, the receiver's
//     copyFromCommBuffer(face, subface) from it (getRecvBufferPointerSetter);
//  4. Scalar3D / Vector3D box copies (copyToBuffer, copyFromBuffer,
//     copyFromDataClass) round trip.
// Prints PASS when every check holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "BCMOctree.h"
#include "Block.h"
#include "BlockFactory.h"
#include "BlockManager.h"
#include "BoundaryConditionSetterBase.h"
#include "Partition.h"
#include "RootGrid.h"
#include "Scalar3D.h"
#include "Scalar3DUpdater.h"
#include "Vector3D.h"
#include "Vector3DUpdater.h"

// Refine every node whose box meets the sphere |x - c| = r, to maxLevel.
class SphereSurface : public Divider {
  const RootGrid* grid;
  int minLevel, maxLevel;
  double c[3], r;

 public:
  SphereSurface(const RootGrid* g, int lo, int hi) : grid(g), minLevel(lo), maxLevel(hi), r(0.3) {
    c[0] = c[1] = c[2] = 0.5;
  }
  NodeType operator()(const Pedigree& p) {
    const int level = int(p.getLevel());
    if (level < minLevel) return BRANCH;
    if (level >= maxLevel) return LEAF_ACTIVE;
    const double h = 1.0 / p.getUpperBound();
    const int root = int(p.getRootID());
    const double lo[3] = {grid->rootID2indexX(root) + p.getX() * h, grid->rootID2indexY(root) + p.getY() * h,
                          grid->rootID2indexZ(root) + p.getZ() * h};
    double dmin = 0.0, dmax = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double x0 = lo[a] * 0.5, x1 = (lo[a] + h) * 0.5;  // root length 0.5
      const double dn = c[a] < x0 ? x0 - c[a] : (c[a] > x1 ? c[a] - x1 : 0.0);
      const double dx = std::fmax(std::fabs(c[a] - x0), std::fabs(c[a] - x1));
      dmin += dn * dn;
      dmax += dx * dx;
    }
    return (dmin <= r * r && dmax >= r * r) ? BRANCH : LEAF_ACTIVE;
  }
};

class AllDirichlet : public BoundaryConditionSetterBase {
 public:
  BoundaryInfo* operator()(const Node* node, const BCMOctree* tree) {
    BoundaryInfo* bi = new BoundaryInfo[NUM_FACE];
    for (int f = 0; f < NUM_FACE; ++f)
      if (tree->checkOnOuterBoundary(node, Face(f))) {
        bi[f].setType(BoundaryInfo::DIRICHLET);
        bi[f].setID(0);
      }
    return bi;
  }
};

static int failures = 0;
static void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    ++failures;
  }
}

int main() {
  const int n = 8, vc = 1;
  RootGrid* grid = new RootGrid(2, 2, 2);
  SphereSurface divider(grid, 1, 2);  // two levels: 2:1 by construction
  BCMOctree* tree = new BCMOctree(grid, &divider, BCMOctree::HILBERT);
  Partition partition(1, tree->getNumLeafNode());
  AllDirichlet setter;
  BlockFactory factory(tree, &partition, &setter, Vec3r(0.0, 0.0, 0.0), 0.5, Vec3i(n, n, n));
  BlockManager& bm = BlockManager::getInstance();
  for (int id = partition.getStart(0); id < partition.getEnd(0); ++id)
    bm.registerBlock(factory.makeBlock(tree->getLeafNodeArray()[id]));
  bm.endRegisterBlock();
  bm.printBlockLayoutInfo();
  delete tree;
  const int nb = bm.getNumBlock();
  int levels[16] = {0};
  for (int b = 0; b < nb; ++b) levels[bm.getBlock(b)->getLevel()]++;
  std::printf("blocks %d (level 1: %d, level 2: %d)\n", nb, levels[1], levels[2]);
  expect(levels[1] > 0 && levels[2] > 0, "the sphere divider refines to two levels");

  const int sid = bm.setDataClass<Scalar3D<double>, Scalar3DUpdater<double> >(vc);
  const int vid = bm.setDataClass<Vector3D<double>, Vector3DUpdater<double> >(vc);
  bm.prepareForVCUpdate(sid, 100, false);
  bm.prepareForVCUpdate(vid, 101, false);

  // Interior values: a smooth function per component, distinct per block.
  auto value = [](int b, int i, int j, int k, int l) {
    return std::sin(0.37 * (i + 1) + 0.11 * l) * std::cos(0.23 * (j + 2) - 0.05 * b) + 0.01 * k * (l + 1);
  };
  const int w = n + 2 * vc;
  for (int b = 0; b < nb; ++b) {
    Vector3D<double>* v = dynamic_cast<Vector3D<double>*>(bm.getBlock(b)->getDataClass(vid));
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          for (int l = 0; l < 3; ++l) (*v)(i, j, k, l) = value(b, i, j, k, l);
  }
  bm.updateVC(vid);
  // Component l through the scalar field: the same ghosts.
  for (int l = 0; l < 3; ++l) {
    for (int b = 0; b < nb; ++b) {
      Scalar3D<double>* s = dynamic_cast<Scalar3D<double>*>(bm.getBlock(b)->getDataClass(sid));
      double* d = s->getData();
      for (int c = 0; c < w * w * w; ++c) d[c] = 0.0;
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) (*s)(i, j, k) = value(b, i, j, k, l);
    }
    bm.updateVC(sid);
    bool same = true;
    for (int b = 0; b < nb && same; ++b) {
      Scalar3D<double>* s = dynamic_cast<Scalar3D<double>*>(bm.getBlock(b)->getDataClass(sid));
      Vector3D<double>* v = dynamic_cast<Vector3D<double>*>(bm.getBlock(b)->getDataClass(vid));
      for (int k = -vc; k < n + vc && same; ++k)
        for (int j = -vc; j < n + vc && same; ++j)
          for (int i = -vc; i < n + vc; ++i)
            if (std::memcmp(&(*s)(i, j, k), &(*v)(i, j, k, l), sizeof(double)) != 0) {
              same = false;
              break;
            }
    }
    expect(same, "vector ghost fill equals the scalar fill of each component");
  }

  // Per-face exchange through the VCUpdater interface, against updateVC.
  for (int resident = 0; resident < 2; ++resident) {
    for (int run = 0; run < 3; ++run) {
      const int id = run == 1 ? vid : sid;
      const bool buffers = run == 2;  // the scalar field through the buffer rules
      std::vector<double> want;
      for (int b = 0; b < nb; ++b) {
        UpdatableDataClass* dc = dynamic_cast<UpdatableDataClass*>(bm.getBlock(b)->getDataClass(id));
        const int per = (id == vid ? 3 : 1) * w * w * w;
        const double* d = id == vid ? dynamic_cast<Vector3D<double>*>(dc)->getData()
                                    : dynamic_cast<Scalar3D<double>*>(dc)->getData();
        want.insert(want.end(), d, d + per);
      }
      // Clear every ghost (interiors stay), then exchange face by face.
      for (int b = 0; b < nb; ++b) {
        UpdatableDataClass* dc = dynamic_cast<UpdatableDataClass*>(bm.getBlock(b)->getDataClass(id));
        const int nc = id == vid ? 3 : 1;
        double* d = id == vid ? dynamic_cast<Vector3D<double>*>(dc)->getData()
                              : dynamic_cast<Scalar3D<double>*>(dc)->getData();
        for (int k = -vc; k < n + vc; ++k)
          for (int j = -vc; j < n + vc; ++j)
            for (int i = -vc; i < n + vc; ++i) {
              if (i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n) continue;
              for (int l = 0; l < nc; ++l) d[nc * ((i + vc) + w * ((j + vc) + w * (k + vc))) + l] = 0.0;
            }
      }
      if (resident) bm.setDeviceResident(id, true);
      auto updater = [&](int b) {
        return dynamic_cast<UpdatableDataClass*>(bm.getBlock(b)->getDataClass(id))->getUpdater();
      };
      // Receiver b's face f (subface q of a finer neighbour's face): the
      // sender packs its face toward b into a host buffer, b unpacks it.
      auto through_buffer = [&](int b, int f, int q) {
        const NeighborInfo* ni = bm.getBlock(b)->getNeighborInfo();
        const int ld = ni[f].getLevelDifference();
        const int src = ld == 1 ? ni[f].getID(Subface(q)) : ni[f].getID();
        const int sf = f ^ 1;
        // The sender's subface argument: my position on a coarse sender's
        // face (C2F), its own position on my face (F2C), 0 at one level.
        const int sq = ld == -1 ? ni[f].getNeighborSubface() : (ld == 1 ? int(bm.getBlock(src)->getNeighborInfo()[sf].getNeighborSubface()) : 0);
        VCUpdater* su = updater(src);
        VCUpdater* ru = updater(b);
        std::vector<char> buf(su->getSendBufferByteSize(Face(sf), Subface(sq)));
        expect(buf.size() == ru->getRecvBufferByteSize(Face(f), Subface(q)), "send and receive buffers agree in size");
        PointerSetterBase* ps = su->getSendBufferPointerSetter(Face(sf), Subface(ld == -1 ? sq : 0));
        ps->setPointer(buf.data());
        delete ps;
        su->copyToCommBuffer(Face(sf), Subface(sq));
        PointerSetterBase* pr = ru->getRecvBufferPointerSetter(Face(f), Subface(ld == 1 ? q : 0));
        pr->setPointer(buf.data());
        delete pr;
        ru->copyFromCommBuffer(Face(f), Subface(q));
        ps = su->getSendBufferPointerSetter(Face(sf), Subface(ld == -1 ? sq : 0));
        ps->setPointer(nullptr);
        delete ps;
        pr = ru->getRecvBufferPointerSetter(Face(f), Subface(ld == 1 ? q : 0));
        pr->setPointer(nullptr);
        delete pr;
      };
      for (int b = 0; b < nb; ++b) {
        VCUpdater* u = updater(b);
        const NeighborInfo* ni = bm.getBlock(b)->getNeighborInfo();
        for (int f = 0; f < NUM_FACE; ++f) {
          if (ni[f].getLevelDifference() == 1) {
            for (int q = 0; q < NUM_SUBFACE; ++q)
              if (ni[f].getID(Subface(q)) >= 0) {
                if (buffers) through_buffer(b, f, q);
                else u->copyFromNeighbor(Face(f), Subface(q));
              }
          } else if (ni[f].getID() >= 0) {
            if (buffers) through_buffer(b, f, 0);
            else u->copyFromNeighbor(Face(f), Subface(0));
          }
        }
      }
      if (resident) bm.setDeviceResident(id, false);
      // Outer faces keep their cleared ghosts in both (no neighbour): compare
      // every cell whose face has a neighbour, i.e. all but the outer ghosts.
      size_t off = 0;
      bool same = true;
      for (int b = 0; b < nb; ++b) {
        UpdatableDataClass* dc = dynamic_cast<UpdatableDataClass*>(bm.getBlock(b)->getDataClass(id));
        const int nc = id == vid ? 3 : 1;
        const double* d = id == vid ? dynamic_cast<Vector3D<double>*>(dc)->getData()
                                    : dynamic_cast<Scalar3D<double>*>(dc)->getData();
        const NeighborInfo* ni = bm.getBlock(b)->getNeighborInfo();
        for (int k = -vc; k < n + vc; ++k)
          for (int j = -vc; j < n + vc; ++j)
            for (int i = -vc; i < n + vc; ++i) {
              const int c[3] = {i, j, k};
              int outside = 0, face = -1;
              for (int a = 0; a < 3; ++a)
                if (c[a] < 0 || c[a] >= n) {
                  ++outside;
                  face = 2 * a + (c[a] >= n);
                }
              if (outside != 1) continue;  // interior, edges, corners
              const bool has = ni[face].getLevelDifference() == 1 ? ni[face].getID(Subface(0)) >= 0
                                                                   : ni[face].getID() >= 0;
              if (!has) continue;
              for (int l = 0; l < nc; ++l) {
                const size_t e = size_t(nc) * ((i + vc) + w * ((j + vc) + w * (k + vc))) + l;
                if (std::memcmp(&d[e], &want[off + e], sizeof(double)) != 0) same = false;
              }
            }
        off += size_t(nc) * w * w * w;
      }
      expect(same, buffers ? (resident ? "buffer rules equal updateVC (device-resident)"
                                       : "buffer rules equal updateVC (host-resident)")
                           : (resident ? "per-face exchange equals updateVC (device-resident)"
                                       : "per-face exchange equals updateVC (host-resident)"));
    }
  }

  // Box copies.
  {
    Scalar3D<double>* a = dynamic_cast<Scalar3D<double>*>(bm.getBlock(0)->getDataClass(sid));
    Scalar3D<double>* b = dynamic_cast<Scalar3D<double>*>(bm.getBlock(nb - 1)->getDataClass(sid));
    std::vector<double> buf(3 * 4 * 5);
    a->copyToBuffer(1, 2, 0, 3, 4, 5, buf.data());
    expect(buf[1 + 3 * (2 + 4 * 3)] == (*a)(2, 4, 3), "Scalar3D::copyToBuffer");
    b->copyFromBuffer(0, 0, 1, 3, 4, 5, buf.data());
    expect((*b)(1, 2, 4) == (*a)(2, 4, 3), "Scalar3D::copyFromBuffer");
    b->copyFromDataClass(-1, -1, -1, 0, 0, 0, 2, 2, 2, a);
    expect((*b)(0, 0, 0) == (*a)(1, 1, 1), "Scalar3D::copyFromDataClass");
    Vector3D<double>* va = dynamic_cast<Vector3D<double>*>(bm.getBlock(0)->getDataClass(vid));
    std::vector<double> vb(3 * 2 * 2 * 2);
    va->copyToBuffer(0, 0, 0, 2, 2, 2, vb.data());
    expect(vb[2 + 3 * (1 + 2 * (1 + 2 * 1))] == (*va)(1, 1, 1, 2), "Vector3D::copyToBuffer");
  }

  const MPI::Intracomm& comm = bm.getCommunicator();
  expect(comm.Get_rank() == 0 && comm.Get_size() == 1, "one-rank communicator");
  bm.printBlockLayoutInfo("layout.txt");
  std::printf(failures ? "FAILED (%d)\n" : "PASS\n", failures);
  return failures ? 1 : 0;
}
