// Partition (Utils/include/Partition.h:29-108): nItems items split into
// nProcs contiguous ranges, the first nItems % nProcs ranges one longer.
#ifndef BCM_FACADE_PARTITION_H
#define BCM_FACADE_PARTITION_H

#include <algorithm>
#include <iostream>
#include <vector>

#include "mpi.h"

class Partition {
  int nProcs;
  int nItems;
  std::vector<int> end;  // one past the last item of each rank

 public:
  Partition(int nProcs_, int nItems_) : nProcs(nProcs_), nItems(nItems_), end(nProcs_) {
    const int m = nItems / nProcs, r = nItems % nProcs;
    int e = 0;
    for (int i = 0; i < nProcs; ++i) {
      e += m + (i < r ? 1 : 0);
      end[i] = e;
    }
  }
  int getStart(int rank) const { return rank == 0 ? 0 : end[rank - 1]; }
  int getEnd(int rank) const { return end[rank]; }
  int getNum(int rank) const { return getEnd(rank) - getStart(rank); }
  int getRank(int i) const {
    if (i < 0 || i >= nItems) return MPI::PROC_NULL;
    return int(std::upper_bound(end.begin(), end.end(), i) - end.begin());
  }
  void print() const {
    for (int rank = 0; rank < nProcs; ++rank)
      std::cout << rank << ": [" << getStart(rank) << ":" << end[rank] - 1 << "]"
                << " #" << getNum(rank) << std::endl;
  }
};

#endif
