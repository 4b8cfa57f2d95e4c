// BCMOctree (Utils/include/BCMOctree.h, Utils/src/BCMOctree.cpp): the
// octree of blocks over a RootGrid, refined by a Divider, its active leaves
// numbered in Z (Morton), Hilbert or random order.
//
// The tree is built by the library (bcm_tree_build_user), which calls the
// driver's Divider for every node; this class keeps the leaves as Node
// records and answers the reference's queries from the library's block
// table: neighbour tables (makeNeighborInfo), outer-face tests and origins
// in root lengths.  As in the reference there is no 2:1 balancing: a face
// neighbour two levels finer is an error, one two levels coarser is no
// neighbour.
#ifndef BCM_FACADE_BCMOCTREE_H
#define BCM_FACADE_BCMOCTREE_H

#include <cstring>
#include <iostream>
#include <vector>

#include "BCMTools.h"
#include "Divider.h"
#include "NeighborInfo.h"
#include "Node.h"
#include "Partition.h"
#include "Pedigree.h"
#include "RootGrid.h"
#include "Vec3.h"
#include "bcm_hip.h"

using namespace Vec3class;

class BCMOctree {
 public:
  enum Ordering { Z, HILBERT, RANDOM, PEDIGREELIST };

 private:
  RootGrid* rootGrid;
  Divider* divider;
  Ordering ordering;
  std::vector<Node*> leafNodeArray;
  std::vector<bcm_block_info> table;  // the library's block table, by block ID

  static int divide(void* self, uint64_t pedigree) {
    Divider* d = static_cast<BCMOctree*>(self)->divider;
    return int((*d)(Pedigree::fromWord(pedigree)));
  }

 public:
  BCMOctree(RootGrid* rootGrid_, Divider* divider_, Ordering ordering_)
      : rootGrid(rootGrid_), divider(divider_), ordering(ordering_) {
    bcm_tree_params p;
    std::memset(&p, 0, sizeof(p));
    p.root_n[0] = rootGrid->getSizeX();
    p.root_n[1] = rootGrid->getSizeY();
    p.root_n[2] = rootGrid->getSizeZ();
    p.periodic[0] = rootGrid->isPeriodicX();
    p.periodic[1] = rootGrid->isPeriodicY();
    p.periodic[2] = rootGrid->isPeriodicZ();
    p.root_length = 1.0;
    p.divider = BCM_DIV_USER;
    p.min_level = 0;
    p.max_level = int(Pedigree::MaxLevel);
    p.balance = 0;
    p.bc_setter = BCM_BCSET_ALL_DIRICHLET;  // the driver's setter decides (BlockFactory)
    if (ordering == Z) p.ordering = BCM_ORDER_Z;
    else if (ordering == HILBERT) p.ordering = BCM_ORDER_HILBERT;
    else if (ordering == RANDOM) p.ordering = BCM_ORDER_RANDOM;
    else {
      std::cout << "*** error: BCMOctree: ordering PEDIGREELIST needs a pedigree list" << std::endl;
      Exit(EX_FAILURE);
    }
    bcm_tree_t t = nullptr;
    int rc = bcm_tree_build_user(&p, &BCMOctree::divide, this, &t);
    if (rc != BCM_OK) {
      std::cout << "*** error: BCMOctree: bcm_tree_build_user returned " << rc
                << (rc == BCM_ERR_BALANCE ? " (a face neighbour two levels finer)" : "") << std::endl;
      Exit(EX_FAILURE);
    }
    const int nb = bcm_tree_num_blocks(t);
    table.resize(nb > 0 ? nb : 0);
    if (nb > 0) rc = bcm_tree_get_blocks(t, table.data(), nb);
    bcm_tree_destroy(t);
    if (nb < 0 || rc != BCM_OK) {
      std::cout << "*** error: BCMOctree: cannot read the block table" << std::endl;
      Exit(EX_FAILURE);
    }
    leafNodeArray.reserve(table.size());
    for (size_t b = 0; b < table.size(); ++b)
      leafNodeArray.push_back(new Node(Pedigree::fromWord(table[b].pedigree), int(b)));
  }
  ~BCMOctree() {
    for (Node* n : leafNodeArray) delete n;
  }

  void broadcast(MPI::Intracomm& = MPI::COMM_WORLD) {}
  const RootGrid* getRootGrid() const { return rootGrid; }
  int getNumLeafNode() const { return int(leafNodeArray.size()); }
  std::vector<Node*>& getLeafNodeArray() { return leafNodeArray; }
  const std::vector<Node*>& getLeafNodeArray() const { return leafNodeArray; }

  // BCMOctree::checkOnOuterBoundary (Utils/src/BCMOctree.cpp:368-396).
  bool checkOnOuterBoundary(const Node* node, Face face) const {
    const Pedigree& pd = node->getPedigree();
    const unsigned c[3] = {pd.getX(), pd.getY(), pd.getZ()};
    const unsigned edge = (face & 1) ? pd.getUpperBound() - 1 : 0u;
    return c[face / 2] == edge && rootGrid->isOuterBoundary(int(pd.getRootID()), face);
  }

  // BCMOctree::getOrigin (Utils/src/BCMOctree.cpp:408-419): in root lengths.
  Vec3r getOrigin(const Node* node) const {
    const Pedigree& pd = node->getPedigree();
    const int r = int(pd.getRootID());
    const int ub = int(pd.getUpperBound());
    return Vec3r(rootGrid->rootID2indexX(r) + REAL_TYPE(pd.getX()) / ub,
                 rootGrid->rootID2indexY(r) + REAL_TYPE(pd.getY()) / ub,
                 rootGrid->rootID2indexZ(r) + REAL_TYPE(pd.getZ()) / ub);
  }

  // BCMOctree::makeNeighborInfo (Utils/src/BCMOctree.cpp:423-478): a new
  // NeighborInfo[NUM_FACE] for the node (the Block made from it owns it).
  NeighborInfo* makeNeighborInfo(const Node* node, const Partition* partition) const {
    NeighborInfo* ni = new NeighborInfo[NUM_FACE];
    const bcm_block_info& bi = table[node->getBlockID()];
    for (int f = 0; f < NUM_FACE; ++f) {
      const Face face = Face(f);
      ni[f].setOuterBoundary(checkOnOuterBoundary(node, face));
      ni[f].setLevelDifference(bi.nb_level_diff[f]);
      if (bi.nb_level_diff[f] == 1) {
        for (int s = 0; s < NUM_SUBFACE; ++s) {
          const int id = bi.nb_id[f][s];
          ni[f].setID(Subface(s), id);
          ni[f].setRank(Subface(s), id >= 0 ? partition->getRank(id) : MPI::PROC_NULL);
        }
      } else {
        const int id = bi.nb_id[f][0];
        ni[f].setID(id);
        ni[f].setRank(id >= 0 ? partition->getRank(id) : MPI::PROC_NULL);
        if (bi.nb_level_diff[f] == -1) ni[f].setNeighborSubface(Subface(bi.nb_subface[f]));
      }
    }
    return ni;
  }

 private:
  BCMOctree(const BCMOctree&);
  BCMOctree& operator=(const BCMOctree&);
};

#endif
