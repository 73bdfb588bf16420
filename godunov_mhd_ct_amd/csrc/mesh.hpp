// mesh.hpp -- host-side block octree (a forest of quad/octrees on a brick of trees).
//
// Restates the part of p4est / kalypsso::core the time step depends on
// (SolverGodunovMHD.h:59-84, scheme/GodunovImplemV0.cpp:40-75): leaves in tree-then-Morton
// order, 2:1 balance across faces, edges and corners, periodic or bounded bricks, the
// "outside" blocks beyond non-periodic borders, each the outward image of a border leaf at
// that leaf's level (border_conditions/FillOutside.h),
// and the per-block neighbour tables the GPU kernels read instead of a hash map.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace mhdct
{

struct Leaf
{
  int level;
  int x, y, z; // block coordinates at `level` over the whole brick
};

// One neighbour-table entry: for block C, direction delta in {-1,0,1}^dim and sub-position
// s (which half of C along each tangential axis), the leaf holding the ghost cells there.
//   rel  0 : local = i + off            (same level)
//   rel -1 : local = floor((i+off)/2)   (coarser)
//   rel +1 : local = 2*i + off          (finer, lower fine cell of the 2^dim covering i)
struct NbEntry
{
  int32_t oct;    // -1 if none
  int8_t  rel;
  int8_t  pad0;
  int16_t off[3];
};
static_assert(sizeof(NbEntry) == 12, "NbEntry layout");

// One corrected border edge of the gather list: dst = (sum of nsrc sources from `first`) /
// cnt / ratio.
struct EmfFix
{
  int64_t dst;
  int32_t first; // index into the source list
  int16_t nsrc;
  int16_t cnt;
  double  ratio;
};
static_assert(sizeof(EmfFix) == 24, "EmfFix layout");


// Block key -> index map with open addressing and linear probing: the neighbour-table build
// and the 2:1 balance look keys up millions of times per mesh.  Indices are >= 0, so an index
// of -1 marks an empty bucket (every key value, 0 included, can be stored).
class KeyIndex
{
public:
  void clear(size_t n)
  {
    size_t cap = 16;
    while (cap < 2 * n + 16)
      cap <<= 1;
    keys_.assign(cap, 0);
    vals_.assign(cap, -1);
    mask_ = cap - 1;
  }
  void put(uint64_t k, int v)
  {
    size_t h = mix(k) & mask_;
    while (vals_[h] >= 0 && keys_[h] != k)
      h = (h + 1) & mask_;
    keys_[h] = k;
    vals_[h] = v;
  }
  int get(uint64_t k) const
  {
    if (keys_.empty())
      return -1;
    size_t h = mix(k) & mask_;
    while (vals_[h] >= 0)
    {
      if (keys_[h] == k)
        return vals_[h];
      h = (h + 1) & mask_;
    }
    return -1;
  }

private:
  static size_t mix(uint64_t k)
  {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return size_t(k);
  }
  std::vector<uint64_t> keys_;
  std::vector<int>      vals_;
  size_t                mask_ = 0;
};

class Mesh
{
public:
  int  dim = 3;
  int  b = 16;
  int  nbrick[3] = { 1, 1, 1 };
  int  periodic[3] = { 1, 1, 1 };
  int  bc[6] = { 0, 0, 0, 0, 0, 0 }; // BC_MHD per face when not periodic (2 ZERO_GRADIENT, 3 WALL)
  int  max_level = 20;
  std::vector<Leaf> leaves;          // owned leaves, in tree-then-Morton order
  std::vector<Leaf> outside;         // outside blocks (each at its source leaf's level)
  std::vector<int>  outside_src;     // source leaf of each outside block
  std::vector<int>  outside_dirs;    // 3 per outside block: outward components

  // derived tables (rebuilt by finalize())
  std::vector<NbEntry> nb;           // [leaf][3^dim][2^dim]
  std::vector<uint8_t> conformal;    // [leaf][2*dim]: 1 if the face neighbour is finer
  std::string          err;

  static uint64_t key(int l, int x, int y, int z);

  int  ntotal() const { return int(leaves.size() + outside.size()); }
  long ncells(int l, int d) const { return long(nbrick[d]) * (long(1) << l) * b; }
  long nblocks(int l, int d) const { return long(nbrick[d]) * (long(1) << l); }

  void uniform(int level);
  // flags per leaf: +1 refine, -1 coarsen (only applied when all siblings agree), 0 keep
  bool adapt(const std::vector<int8_t> & flags);
  // replace the leaves by a given set that tiles the brick (no overlap, nothing missing)
  bool set_leaves(const std::vector<Leaf> & given);
  // full (face+edge+corner) 2:1 balance
  void balance();
  void sort_leaves();
  bool finalize(); // outside blocks + tables; false on error (see err)

  int find(int l, int x, int y, int z) const;
  // leaf (owned or outside) holding global cell g (level-L cell units, unwrapped); rel as above
  bool locate(int L, const long g[3], int & oct, int & rel, long loc[3]) const;

  int  nb_index(int dir27, int sub) const { return dir27 * (1 << dim) + sub; }
  int  ndirs() const { return dim == 3 ? 27 : 9; }

private:
  KeyIndex hash_; // owned + outside
  void rebuild_hash();
  bool build_outside();
  bool build_nb();
};

} // namespace mhdct
