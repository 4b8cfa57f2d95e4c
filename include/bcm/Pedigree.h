// Pedigree (Utils/include/Pedigree.h:36-196): a node's level, its integer
// position inside its root (x, y, z < 2^level) and its root ID packed in
// one 64-bit word as x16 | y16 | z16 | root12 | level4 -- the word the
// library's block table carries (bcm_block_info.pedigree).
#ifndef BCM_FACADE_PEDIGREE_H
#define BCM_FACADE_PEDIGREE_H

#include <stdint.h>

#include <cstring>
#include <iostream>

#include "BCMTools.h"

class Pedigree {
 public:
  const static unsigned MaxLevel = 0xf;
  const static unsigned MaxRootID = 0xfff;
  const static unsigned MaxCoord = 0xffff;

 private:
  uint64_t p;
  static uint64_t pack(unsigned level, unsigned x, unsigned y, unsigned z, unsigned rootID) {
    return (uint64_t(x) << 48) | (uint64_t(y) << 32) | (uint64_t(z) << 16) | (uint64_t(rootID & MaxRootID) << 4) |
           uint64_t(level & MaxLevel);
  }

 public:
  Pedigree(unsigned rootID = 0) : p(pack(0, 0, 0, 0, rootID)) {}
  Pedigree(unsigned level, unsigned x, unsigned y, unsigned z, unsigned rootID = 0)
      : p(pack(level, x, y, z, rootID)) {}
  // Child ijk = i + 2j + 4k of a parent.
  Pedigree(const Pedigree& parent, unsigned ijk)
      : p(pack(parent.getLevel() + 1, 2 * parent.getX() + (ijk & 1), 2 * parent.getY() + ((ijk >> 1) & 1),
               2 * parent.getZ() + ((ijk >> 2) & 1), parent.getRootID())) {}
  unsigned getLevel() const { return unsigned(p & MaxLevel); }
  unsigned getX() const { return unsigned((p >> 48) & MaxCoord); }
  unsigned getY() const { return unsigned((p >> 32) & MaxCoord); }
  unsigned getZ() const { return unsigned((p >> 16) & MaxCoord); }
  unsigned getRootID() const { return unsigned((p >> 4) & MaxRootID); }
  unsigned getUpperBound() const { return 1u << getLevel(); }
  // Bit of the position at `level` (1 <= level <= getLevel()).
  unsigned getX(unsigned level) const { return unsigned((p >> (getLevel() - level + 48)) & 1); }
  unsigned getY(unsigned level) const { return unsigned((p >> (getLevel() - level + 32)) & 1); }
  unsigned getZ(unsigned level) const { return unsigned((p >> (getLevel() - level + 16)) & 1); }
  unsigned getChildId(unsigned level) const { return getX(level) + 2 * getY(level) + 4 * getZ(level); }
  static size_t GetSerializeSize() { return sizeof(uint64_t); }
  void serialize(void* buf) const { std::memcpy(buf, &p, sizeof(p)); }
  void deserialize(const void* buf) { std::memcpy(&p, buf, sizeof(p)); }
  uint64_t word() const { return p; }
  static Pedigree fromWord(uint64_t w) {
    Pedigree q;
    q.p = w;
    return q;
  }
};

inline std::ostream& operator<<(std::ostream& os, const Pedigree& p) {
  return os << "(" << p.getRootID() << ":" << p.getLevel() << "(" << p.getX() << "," << p.getY() << "," << p.getZ()
            << "))";
}

#endif
