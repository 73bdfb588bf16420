// mesh.cpp -- host block octree: adaptation, 2:1 balance, ordering, neighbour tables.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <climits>
#include <cstdint>
#include <thread>
#include <cstdio>

namespace mhdct
{

uint64_t
Mesh::key(int l, int x, int y, int z)
{
  const uint64_t off = 1u << 20;
  return (uint64_t(l) << 60) | ((uint64_t(int64_t(x) + off) & 0xFFFFF) << 40) |
         ((uint64_t(int64_t(y) + off) & 0xFFFFF) << 20) | (uint64_t(int64_t(z) + off) & 0xFFFFF);
}

void
Mesh::rebuild_hash()
{
  hash_.clear(size_t(ntotal()));
  for (size_t i = 0; i < leaves.size(); ++i)
    hash_.put(key(leaves[i].level, leaves[i].x, leaves[i].y, leaves[i].z), int(i));
  for (size_t i = 0; i < outside.size(); ++i)
    hash_.put(key(outside[i].level, outside[i].x, outside[i].y, outside[i].z), int(leaves.size() + i));
}

int
Mesh::find(int l, int x, int y, int z) const
{
  if (l < 0)
    return -1;
  return hash_.get(key(l, x, y, z));
}

void
Mesh::uniform(int level)
{
  leaves.clear();
  const int nz = dim == 3 ? int(nblocks(level, 2)) : 1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < nblocks(level, 1); ++y)
      for (int x = 0; x < nblocks(level, 0); ++x)
        leaves.push_back({ level, x, y, z });
  sort_leaves();
}

static inline long
floordiv(long a, long q)
{
  return a >= 0 ? a / q : -((-a + q - 1) / q);
}

bool
Mesh::adapt(const std::vector<int8_t> & flags)
{
  if (flags.size() != leaves.size())
  {
    err = "adapt: flag count does not match the number of leaves";
    return false;
  }
  rebuild_hash();
  const int         nchild = 1 << dim;
  std::vector<Leaf> out;
  out.reserve(leaves.size() * 2);
  std::vector<char> done(leaves.size(), 0);
  for (size_t i = 0; i < leaves.size(); ++i)
  {
    if (done[i])
      continue;
    const Leaf & L = leaves[i];
    if (flags[i] > 0 && L.level < max_level)
    {
      for (int c = 0; c < nchild; ++c)
        out.push_back({ L.level + 1, 2 * L.x + (c & 1), 2 * L.y + ((c >> 1) & 1),
                        dim == 3 ? 2 * L.z + ((c >> 2) & 1) : 0 });
      done[i] = 1;
      continue;
    }
    if (flags[i] < 0 && L.level > 0)
    {
      const int px = L.x >> 1, py = L.y >> 1, pz = dim == 3 ? (L.z >> 1) : 0;
      bool      all = true;
      int       sib[8];
      for (int c = 0; c < nchild && all; ++c)
      {
        const int o = find(L.level, 2 * px + (c & 1), 2 * py + ((c >> 1) & 1),
                           dim == 3 ? 2 * pz + ((c >> 2) & 1) : 0);
        if (o < 0 || o >= int(leaves.size()) || flags[o] >= 0 || done[o])
          all = false;
        else
          sib[c] = o;
      }
      if (all)
      {
        for (int c = 0; c < nchild; ++c)
          done[sib[c]] = 1;
        out.push_back({ L.level - 1, px, py, pz });
        continue;
      }
    }
    out.push_back(L);
    done[i] = 1;
  }
  leaves.swap(out);
  sort_leaves();
  return true;
}

bool
Mesh::set_leaves(const std::vector<Leaf> & given)
{
  // each leaf inside the brick at its level, no leaf inside another, and the leaves' volumes
  // (in units of the finest level) summing to the brick's: together, an exact tiling
  int lmax = 0;
  for (const Leaf & L : given)
  {
    if (L.level < 0 || L.level > max_level)
    {
      err = "set_leaves: level out of range";
      return false;
    }
    const int c[3] = { L.x, L.y, L.z };
    for (int d = 0; d < 3; ++d)
      if (c[d] < 0 || (d < dim && c[d] >= nblocks(L.level, d)) || (d >= dim && c[d] != 0))
      {
        err = "set_leaves: a leaf lies outside the brick";
        return false;
      }
    lmax = std::max(lmax, L.level);
  }
  std::unordered_map<uint64_t, int> h;
  h.reserve(given.size() * 2 + 16);
  for (size_t i = 0; i < given.size(); ++i)
    if (!h.emplace(key(given[i].level, given[i].x, given[i].y, given[i].z), int(i)).second)
    {
      err = "set_leaves: a leaf is given twice";
      return false;
    }
  long double vol = 0, total = 1;
  for (int d = 0; d < dim; ++d)
    total *= (long double)nbrick[d];
  for (const Leaf & L : given)
  {
    for (int l = L.level - 1; l >= 0; --l)
    {
      const int sh = L.level - l;
      if (h.count(key(l, L.x >> sh, L.y >> sh, dim == 3 ? (L.z >> sh) : 0)))
      {
        err = "set_leaves: a leaf lies inside another";
        return false;
      }
    }
    vol += 1.0L / (long double)(1L << (dim * L.level));
  }
  if (std::fabs(double(vol - total)) > 1e-9 * double(total))
  {
    err = "set_leaves: the leaves do not cover the brick";
    return false;
  }
  leaves = given;
  outside.clear();
  sort_leaves();
  rebuild_hash();
  return true;
}

void
Mesh::balance()
{
  // iterate until no leaf has a neighbour (face/edge/corner) more than one level coarser; each
  // pass scans the leaves in contiguous ranges on host threads, each marking its own flags
  for (;;)
  {
    rebuild_hash();
    const int nleaves = int(leaves.size());
    const int nthr = std::max(1, std::min<int>(int(std::thread::hardware_concurrency()), std::min(16, nleaves / 256)));
    std::vector<std::vector<int8_t>> tflags(static_cast<size_t>(nthr));
    auto scan = [&](int t) {
      std::vector<int8_t> & flags = tflags[size_t(t)];
      flags.assign(size_t(nleaves), 0);
      const int lo = int(long(nleaves) * t / nthr), hi = int(long(nleaves) * (t + 1) / nthr);
      for (int i = lo; i < hi; ++i)
      {
        const Leaf & L = leaves[size_t(i)];
        if (L.level < 2)
          continue;
        for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx)
            {
              if (!dx && !dy && !dz)
                continue;
              long p[3] = { long(L.x) + dx, long(L.y) + dy, long(L.z) + dz };
              bool skip = false;
              for (int d = 0; d < dim; ++d)
              {
                const long n = nblocks(L.level, d);
                if (p[d] < 0 || p[d] >= n)
                {
                  if (periodic[d])
                    p[d] = ((p[d] % n) + n) % n;
                  else
                    skip = true;
                }
              }
              if (skip)
                continue;
              // walk up: find the leaf holding position p; one more than a level coarser
              // gets refined
              for (int l = L.level; l >= 0; --l)
              {
                const int sh = L.level - l;
                const int o = find(l, int(p[0] >> sh), int(p[1] >> sh), int(p[2] >> sh));
                if (o >= 0)
                {
                  if (l < L.level - 1)
                    flags[size_t(o)] = 1;
                  break;
                }
              }
            }
      }
    };
    if (nthr == 1)
      scan(0);
    else
    {
      std::vector<std::thread> pool;
      for (int t = 0; t < nthr; ++t)
        pool.emplace_back(scan, t);
      for (auto & th : pool)
        th.join();
    }
    std::vector<int8_t> flags(size_t(nleaves), 0);
    bool                any = false;
    for (const auto & f : tflags)
      for (int i = 0; i < nleaves; ++i)
        if (f[size_t(i)])
        {
          flags[size_t(i)] = 1;
          any = true;
        }
    if (!any)
      break;
    adapt(flags);
  }
  rebuild_hash();
}

// Tree numbers of a brick of nbrick[0] x nbrick[1] (x nbrick[2]) trees, indexed by the tree's
// position tx + nbrick[0] (ty + nbrick[1] tz): the order in which p4est's brick connectivity
// numbers its trees (p4est_connectivity_new_brick / p8est_connectivity_new_brick of p4est, the
// reference's mesh library, CMakeLists.txt:77, version not pinned; not in the reference tree).
// Linear indices ti = 0, 1, ... of the enclosing power-of-two brick are mapped to positions by
// interleaving bits: the axes are ranked by their bit count log_d = ceil(log2 nbrick[d]) (fewest
// first, ties in axis order); while every remaining axis still has bits, one bit per axis is
// taken in turn (lowest-ranked remaining axes first in axis order), then the axis with fewer
// bits drops out; positions outside the brick are skipped and the rest numbered in order.
std::vector<int>
brick_tree_numbers(int dim, const int nbrick[3])
{
  int logx[3] = { 0, 0, 0 };
  for (int d = 0; d < dim; ++d)
    while ((1 << logx[d]) < nbrick[d])
      ++logx[d];
  int rank[3] = { 0, 1, 2 };
  std::stable_sort(rank, rank + dim, [&](int a, int b) { return logx[a] < logx[b]; });
  int total = 0;
  for (int d = 0; d < dim; ++d)
    total += logx[d];
  const long       n = long(nbrick[0]) * nbrick[1] * (dim == 3 ? nbrick[2] : 1);
  std::vector<int> num(size_t(n), -1);
  int              next = 0;
  for (long ti = 0; ti < (long(1) << total); ++ti)
  {
    long pos[3] = { 0, 0, 0 };
    long t = ti;
    int  last = 0;
    for (int i = 0; i < dim - 1; ++i)
    {
      const int rounds = logx[rank[i]] - last;
      // the axes still taking bits in this phase, in axis order
      int active[3], na = 0;
      for (int d = 0; d < dim; ++d)
        for (int k = i; k < dim; ++k)
          if (rank[k] == d)
            active[na++] = d;
      for (int j = 0; j < rounds; ++j)
        for (int a = 0; a < na; ++a)
        {
          pos[active[a]] |= ((t >> (na * j + a)) & 1) << (last + j);
        }
      t >>= na * rounds;
      last += rounds;
    }
    pos[rank[dim - 1]] += t << last;
    if (pos[0] < nbrick[0] && pos[1] < nbrick[1] && (dim == 2 || pos[2] < nbrick[2]))
      num[size_t(pos[0] + nbrick[0] * (pos[1] + long(nbrick[1]) * pos[2]))] = next++;
  }
  return num;
}

void
Mesh::sort_leaves()
{
  int M = 0;
  for (auto & L : leaves)
    M = std::max(M, L.level);
  const std::vector<int> treeno = brick_tree_numbers(dim, nbrick);
  auto morton = [&](const Leaf & L) {
    const long tx = L.x >> L.level, ty = L.y >> L.level, tz = dim == 3 ? (L.z >> L.level) : 0;
    const uint64_t tree = uint64_t(treeno[size_t(tx + nbrick[0] * (ty + long(nbrick[1]) * tz))]);
    const uint64_t lx = uint64_t(L.x - (tx << L.level)) << (M - L.level);
    const uint64_t ly = uint64_t(L.y - (ty << L.level)) << (M - L.level);
    const uint64_t lz = dim == 3 ? uint64_t(L.z - (tz << L.level)) << (M - L.level) : 0;
    uint64_t       code = 0;
    for (int bit = 0; bit < M; ++bit)
    {
      if (dim == 3)
        code |= (((lx >> bit) & 1) << (3 * bit)) | (((ly >> bit) & 1) << (3 * bit + 1)) |
                (((lz >> bit) & 1) << (3 * bit + 2));
      else
        code |= (((lx >> bit) & 1) << (2 * bit)) | (((ly >> bit) & 1) << (2 * bit + 1));
    }
    return std::make_pair(tree, code);
  };
  std::vector<std::pair<std::pair<uint64_t, uint64_t>, size_t>> keys(leaves.size());
  for (size_t i = 0; i < leaves.size(); ++i)
    keys[i] = { morton(leaves[i]), i };
  std::sort(keys.begin(), keys.end(), [&](auto & a, auto & b2) {
    if (a.first != b2.first)
      return a.first < b2.first;
    return leaves[a.second].level < leaves[b2.second].level;
  });
  std::vector<Leaf> out(leaves.size());
  for (size_t i = 0; i < keys.size(); ++i)
    out[i] = leaves[keys[i].second];
  leaves.swap(out);
}

bool
Mesh::locate(int L, const long gin[3], int & oct, int & rel, long loc[3]) const
{
  // periodic directions wrap; beyond a bounded border the outside blocks answer
  long g[3] = { gin[0], gin[1], dim == 3 ? gin[2] : 0 };
  for (int d = 0; d < dim; ++d)
  {
    const long n = ncells(L, d);
    if ((g[d] < 0 || g[d] >= n) && periodic[d])
      g[d] = ((g[d] % n) + n) % n;
  }
  for (int r = 0; r >= -1; --r)
  {
    // r = 0: same level; r = -1: coarser
    long gg[3] = { g[0], g[1], g[2] };
    if (r == -1)
    {
      for (int d = 0; d < 3; ++d)
        gg[d] = floordiv(g[d], 2);
    }
    const int lv = L + r;
    const int o = find(lv, int(floordiv(gg[0], b)), int(floordiv(gg[1], b)),
                       dim == 3 ? int(floordiv(gg[2], b)) : 0);
    if (o >= 0)
    {
      const Leaf & F = o < int(leaves.size()) ? leaves[o] : outside[o - leaves.size()];
      oct = o;
      rel = r;
      loc[0] = gg[0] - long(F.x) * b;
      loc[1] = gg[1] - long(F.y) * b;
      loc[2] = dim == 3 ? gg[2] - long(F.z) * b : 0;
      return true;
    }
  }
  long gf[3] = { 2 * g[0], 2 * g[1], 2 * g[2] };
  const int o = find(L + 1, int(floordiv(gf[0], b)), int(floordiv(gf[1], b)),
                     dim == 3 ? int(floordiv(gf[2], b)) : 0);
  if (o < 0)
    return false;
  const Leaf & F = o < int(leaves.size()) ? leaves[o] : outside[o - leaves.size()];
  oct = o;
  rel = 1;
  loc[0] = gf[0] - long(F.x) * b;
  loc[1] = gf[1] - long(F.y) * b;
  loc[2] = dim == 3 ? gf[2] - long(F.z) * b : 0;
  return true;
}

bool
Mesh::build_outside()
{
  outside.clear();
  outside_src.clear();
  outside_dirs.clear();
  bool any_bounded = false;
  for (int d = 0; d < dim; ++d)
    if (!periodic[d])
      any_bounded = true;
  if (!any_bounded)
  {
    rebuild_hash();
    return true;
  }
  rebuild_hash();
  // Every leaf on a bounded border gets its outward images, at its own level: one per
  // direction whose nonzero components all leave the brick (across a face, and across the
  // edges and corners where borders meet). Each image's source is the leaf itself, the inside
  // block FillOutside reads (border_conditions/FillOutside.cpp:133-152 looks it up at the
  // outside block's level). The images of the border leaves tile the layer beyond the border,
  // so a position there resolves, like one inside, to a block of the same, the next coarser or
  // the next finer level.
  for (size_t i = 0; i < leaves.size(); ++i)
  {
    const Leaf & L = leaves[i];
    for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
        {
          const int dl[3] = { dx, dy, dz };
          const long p[3] = { long(L.x) + dx, long(L.y) + dy, long(L.z) + dz };
          bool       outward = dx || dy || dz;
          for (int d = 0; d < dim && outward; ++d)
            if (dl[d] != 0 && (periodic[d] || (p[d] >= 0 && p[d] < nblocks(L.level, d))))
              outward = false;
          if (!outward)
            continue;
          outside.push_back({ L.level, int(p[0]), int(p[1]), int(p[2]) });
          outside_src.push_back(int(i));
          outside_dirs.push_back(dl[0]);
          outside_dirs.push_back(dl[1]);
          outside_dirs.push_back(dl[2]);
        }
  }
  rebuild_hash();
  return true;
}

bool
Mesh::build_nb()
{
  const int nd = ndirs(), ns = 1 << dim;
  nb.assign(leaves.size() * size_t(nd) * ns, NbEntry{ -1, 0, 0, { 0, 0, 0 } });
  conformal.assign(leaves.size() * 2 * dim, 0);
  // leaves in contiguous ranges on host threads (each writes its own entries); the first error
  // in leaf order is reported
  const int           nleaves = int(leaves.size());
  const int           nthr = std::max(1, std::min<int>(int(std::thread::hardware_concurrency()), std::min(16, nleaves / 64)));
  std::vector<std::string> errs(nthr);
  auto one_range = [&](int t) -> bool {
    const int lo = int(long(nleaves) * t / nthr), hi = int(long(nleaves) * (t + 1) / nthr);
    std::string & err = errs[t];
  for (size_t o = size_t(lo); o < size_t(hi); ++o)
  {
    const Leaf & C = leaves[o];
    const long   org[3] = { long(C.x) * b, long(C.y) * b, long(C.z) * b };
    for (int di = 0; di < nd; ++di)
    {
      const int dl[3] = { di % 3 - 1, (di / 3) % 3 - 1, dim == 3 ? di / 9 - 1 : 0 };
      if (!dl[0] && !dl[1] && !dl[2])
        continue;
      // A neighbour of this leaf's level or coarser covers the whole face, edge or corner, and
      // its offset is the difference of the block origins (org - org_nb, or org - 2 org_nb),
      // whatever the sub-position: sub-position 0 answers for all.  Only a finer neighbour
      // differs per sub-position.
      bool    share = false;
      NbEntry first{ -1, 0, 0, { 0, 0, 0 } };
      for (int s = 0; s < ns; ++s)
      {
        // representative cell of this (direction, sub-position)
        long i[3] = { 0, 0, 0 };
        bool valid_sub = true;
        for (int d = 0; d < dim; ++d)
        {
          if (dl[d] < 0)
            i[d] = -1;
          else if (dl[d] > 0)
            i[d] = b;
          else
            i[d] = ((s >> d) & 1) ? b / 2 : 0;
          if (dl[d] != 0 && ((s >> d) & 1))
            valid_sub = false;
        }
        if (!valid_sub)
          continue;
        if (share)
        {
          nb[(o * nd + di) * ns + s] = first;
          continue;
        }
        const long g[3] = { org[0] + i[0], org[1] + i[1], org[2] + i[2] };
        int        oct, rel;
        long       loc[3];
        if (!locate(C.level, g, oct, rel, loc))
        {
          char buf[160];
          std::snprintf(buf, sizeof(buf), "neighbour missing for leaf %zu (level %d) direction (%d,%d,%d)",
                        o, C.level, dl[0], dl[1], dl[2]);
          err = buf;
          return false;
        }
        NbEntry e{ oct, int8_t(rel), 0, { 0, 0, 0 } };
        for (int d = 0; d < dim; ++d)
        {
          long off = 0;
          if (rel == 0)
            off = loc[d] - i[d];
          else if (rel == -1)
          {
            // local = floor((i + off)/2): off = 2*loc + (g mod 2) - i, with g unwrapped
            const long gm = ((g[d] % 2) + 2) % 2;
            off = 2 * loc[d] + gm - i[d];
          }
          else
            off = loc[d] - 2 * i[d];
          if (off < -32000 || off > 32000)
          {
            err = "neighbour offset overflow";
            return false;
          }
          e.off[d] = int16_t(off);
        }
        nb[(o * nd + di) * ns + s] = e;
        if (s == 0 && rel <= 0)
        {
          first = e;
          share = true;
        }
      }
    }
    for (int d = 0; d < dim; ++d)
      for (int side = 0; side < 2; ++side)
      {
        int dl[3] = { 0, 0, 0 };
        dl[d] = side ? 1 : -1;
        const int di = (dl[0] + 1) + 3 * (dl[1] + 1) + (dim == 3 ? 9 * (dl[2] + 1) : 0);
        conformal[o * 2 * dim + 2 * d + side] = nb[(o * nd + di) * ns + 0].rel == 1 ? 1 : 0;
      }
  }
    return true;
  };
  std::vector<char> ok(nthr, 1);
  if (nthr == 1)
    ok[0] = one_range(0);
  else
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; ++t)
      pool.emplace_back([&, t] { ok[t] = one_range(t); });
    for (auto & th : pool)
      th.join();
  }
  for (int t = 0; t < nthr; ++t)
    if (!ok[t])
    {
      err = errs[t];
      return false;
    }
  return true;
}

bool
Mesh::finalize()
{
  if (b < 4)
  {
    err = "block size must be >= 2*ghostWidth = 4 (SolverGodunovMHD.cpp:117-137)";
    return false;
  }
  if (b % 2)
  {
    err = "block size must be even";
    return false;
  }
  if (!build_outside())
    return false;
  if (!build_nb())
    return false;
  return true;
}

} // namespace mhdct
