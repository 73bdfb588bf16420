// mesh.cpp -- host block octree: adaptation, 2:1 balance, ordering, neighbour tables.
#include "mesh.hpp"

#include <algorithm>
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
  hash_.clear();
  hash_.reserve(size_t(ntotal()) * 2 + 16);
  for (size_t i = 0; i < leaves.size(); ++i)
    hash_[key(leaves[i].level, leaves[i].x, leaves[i].y, leaves[i].z)] = int(i);
  for (size_t i = 0; i < outside.size(); ++i)
    hash_[key(outside[i].level, outside[i].x, outside[i].y, outside[i].z)] =
      int(leaves.size() + i);
}

int
Mesh::find(int l, int x, int y, int z) const
{
  if (l < 0)
    return -1;
  auto it = hash_.find(key(l, x, y, z));
  return it == hash_.end() ? -1 : it->second;
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

void
Mesh::balance()
{
  // iterate until no leaf has a neighbour (face/edge/corner) more than one level coarser
  for (;;)
  {
    rebuild_hash();
    std::vector<int8_t> flags(leaves.size(), 0);
    bool                any = false;
    for (size_t i = 0; i < leaves.size(); ++i)
    {
      const Leaf & L = leaves[i];
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
            // walk up: find the leaf holding position p
            for (int l = L.level; l >= 0; --l)
            {
              const int sh = L.level - l;
              const int o = find(l, int(p[0] >> sh), int(p[1] >> sh), int(p[2] >> sh));
              if (o >= 0)
              {
                if (l < L.level - 1 && !flags[o])
                {
                  flags[o] = 1;
                  any = true;
                }
                break;
              }
            }
          }
    }
    if (!any)
      break;
    adapt(flags);
  }
  rebuild_hash();
}

void
Mesh::sort_leaves()
{
  int M = 0;
  for (auto & L : leaves)
    M = std::max(M, L.level);
  auto morton = [&](const Leaf & L) {
    const long tx = L.x >> L.level, ty = L.y >> L.level, tz = dim == 3 ? (L.z >> L.level) : 0;
    const uint64_t tree = uint64_t(tx + nbrick[0] * (ty + long(nbrick[1]) * tz));
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
  long g[3] = { gin[0], gin[1], dim == 3 ? gin[2] : 0 };
  bool outside_pos = false;
  for (int d = 0; d < dim; ++d)
  {
    const long n = ncells(L, d);
    if (g[d] < 0 || g[d] >= n)
    {
      if (periodic[d])
        g[d] = ((g[d] % n) + n) % n;
      else
        outside_pos = true;
    }
  }
  for (int r = 0; r >= -1; --r)
  {
    // r = 0: same level; r = -1: coarser
    long gg[3] = { g[0], g[1], g[2] };
    if (r == -1)
    {
      if (outside_pos)
        break;
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
  if (outside_pos)
    return false;
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
  std::unordered_map<uint64_t, int> made;
  for (size_t i = 0; i < leaves.size(); ++i)
  {
    const Leaf & L = leaves[i];
    for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
        {
          const int dl[3] = { dx, dy, dz };
          long      p[3] = { long(L.x) + dx, long(L.y) + dy, long(L.z) + dz };
          int       od[3] = { 0, 0, 0 };
          bool      out = false;
          for (int d = 0; d < dim; ++d)
          {
            const long n = nblocks(L.level, d);
            if (p[d] < 0 || p[d] >= n)
            {
              if (periodic[d])
                p[d] = ((p[d] % n) + n) % n;
              else
              {
                out = true;
                od[d] = dl[d];
              }
            }
          }
          if (!out)
            continue;
          const uint64_t k = key(L.level, int(p[0]), int(p[1]), int(p[2]));
          if (made.count(k))
            continue;
          // source: the same-level leaf across the border (position minus outward part)
          long s[3] = { p[0] - od[0], p[1] - od[1], p[2] - od[2] };
          const int src = find(L.level, int(s[0]), int(s[1]), int(s[2]));
          if (src < 0 || src >= int(leaves.size()))
          {
            err = "non-periodic border: the leaves along a bounded border must share one level "
                  "next to each outside block (source leaf missing)";
            return false;
          }
          made[k] = int(outside.size());
          outside.push_back({ L.level, int(p[0]), int(p[1]), int(p[2]) });
          outside_src.push_back(src);
          outside_dirs.push_back(od[0]);
          outside_dirs.push_back(od[1]);
          outside_dirs.push_back(od[2]);
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
  for (size_t o = 0; o < leaves.size(); ++o)
  {
    const Leaf & C = leaves[o];
    const long   org[3] = { long(C.x) * b, long(C.y) * b, long(C.z) * b };
    for (int di = 0; di < nd; ++di)
    {
      const int dl[3] = { di % 3 - 1, (di / 3) % 3 - 1, dim == 3 ? di / 9 - 1 : 0 };
      if (!dl[0] && !dl[1] && !dl[2])
        continue;
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
}

// Border-edge corrections (scheme/CorrectEmfAtBlockBorderFunctor.cpp:94-360) as an explicit
// gather list.  The four cells around an edge are the quadrants (s1,s2) over the two
// tangential axes (cell = edge index - 1 + s); the block's own quadrant rounds the edge index
// down (clamped to b-1); the other three are the sibling slots, invalid when inside the block.
bool
Mesh::build_emf_fix()
{
  emf_fix.clear();
  emf_src.clear();
  const int  ncomp = dim == 3 ? 3 : 1;
  const long n1 = b + 1;
  const long ne = dim == 3 ? n1 * n1 * n1 : n1 * n1;
  const double surface_ratio = double(1 << (dim - 1));
  auto eidx = [&](long oct, int comp, long i, long j, long k) {
    return (oct * ncomp + comp) * ne + i + n1 * (j + n1 * k);
  };
  const int nowned = int(leaves.size());
  const int nd = ndirs(), ns = 1 << dim;
  // the gather list of one leaf, appended to (fx, sx); `first` indexes sx
  auto one_leaf = [&](int o, std::vector<EmfFix> & fx, std::vector<int64_t> & sx) {
    const Leaf & C = leaves[o];
    const long   org[3] = { long(C.x) * b, long(C.y) * b, long(C.z) * b };
    // level relation of the face neighbour across each face (uniform over the face in a 2:1
    // balanced forest; from the neighbour table built before)
    int face_rel[3][2];
    for (int d = 0; d < dim; ++d)
      for (int side = 0; side < 2; ++side)
      {
        int dl[3] = { 0, 0, 0 };
        dl[d] = side ? 1 : -1;
        const int di = (dl[0] + 1) + 3 * (dl[1] + 1) + (dim == 3 ? 9 * (dl[2] + 1) : 0);
        const NbEntry & E = nb[(size_t(o) * nd + di) * ns + 0];
        face_rel[d][side] = E.oct < 0 ? 99 : E.rel;
      }
    for (int comp = 0; comp < ncomp; ++comp)
    {
      const int a = dim == 3 ? comp : 2;
      const int t1 = a == 0 ? 1 : 0;
      const int t2 = a == 2 ? 1 : 2;
      const int nal = dim == 3 ? b : 1;
      // the border ring of the (e1, e2) square: rows e2 = 0 and e2 = b, then columns e1 = 0
      // and e1 = b between them (4b edges per ea)
      const int nring = 4 * b;
      for (int ea = 0; ea < nal; ++ea)
        for (int ri = 0; ri < nring; ++ri)
          {
            int e1, e2;
            if (ri < 2 * (b + 1))
            {
              e2 = ri < b + 1 ? 0 : b;
              e1 = ri % (b + 1);
            }
            else
            {
              const int rj = ri - 2 * (b + 1); // 0 .. 2(b-1)-1
              e1 = rj < b - 1 ? 0 : b;
              e2 = 1 + rj % (b - 1);
            }
            const bool b1 = (e1 == 0 || e1 == b), b2 = (e2 == 0 || e2 == b);
            // an edge inside one face whose neighbour has this block's level: the two cells
            // across are that neighbour's and the other one is this block's, so no correction
            if (b1 != b2)
            {
              const int fd = b1 ? t1 : t2;
              const int fs = (b1 ? e1 : e2) == b ? 1 : 0;
              if (face_rel[fd][fs] == 0)
                continue;
            }
            long e[3] = { 0, 0, 0 };
            e[t1] = e1;
            e[t2] = e2;
            if (dim == 3)
              e[a] = ea;
            const int own1 = e1 < b ? 1 : 0, own2 = e2 < b ? 1 : 0;
            const int ownq = own1 + 2 * own2;
            struct S
            {
              bool valid = false, out = false;
              int  oct = -1, rel = 0;
              long e[3] = { 0, 0, 0 };
            } sl[4];
            for (int q = 0; q < 4; ++q)
            {
              if (q == ownq)
                continue;
              const int s1 = q & 1, s2 = q >> 1;
              long      c[3] = { e[0], e[1], e[2] };
              c[t1] = e[t1] - 1 + s1;
              c[t2] = e[t2] - 1 + s2;
              bool inside = true;
              for (int d = 0; d < dim; ++d)
                if (c[d] < 0 || c[d] >= b)
                  inside = false;
              if (inside)
                continue;
              // the leaf holding cell c, from the neighbour table (the cell's direction and,
              // along the directions it does not leave, its half of the block)
              int dir = 0, sub = 0, mul = 1;
              for (int d = 0; d < dim; ++d)
              {
                const int dl = c[d] < 0 ? -1 : (c[d] >= b ? 1 : 0);
                dir += (dl + 1) * mul;
                mul *= 3;
                if (dl == 0 && c[d] >= b / 2)
                  sub |= 1 << d;
              }
              const NbEntry & E = nb[(size_t(o) * nd + dir) * ns + sub];
              if (E.oct < 0)
                continue;
              const int oct = E.oct, rel = E.rel;
              sl[q].valid = true;
              sl[q].oct = oct;
              sl[q].rel = rel;
              sl[q].out = oct >= nowned;
              const Leaf & F = oct < nowned ? leaves[oct] : outside[oct - nowned];
              for (int d = 0; d < dim; ++d)
              {
                const int  Ls = C.level + rel;
                const long n = long(nbrick[d]) * b << Ls; // ncells(Ls, d)
                long       G = org[d] + e[d];
                long       Gs = rel == 1 ? 2 * G : (rel == 0 ? G : (G >> 1)); // floor(G / 2)
                long       lc = Gs - long(d == 0 ? F.x : (d == 1 ? F.y : F.z)) * b;
                if (d == a)
                {
                  // into [0, n): lc lies within a few periods of it
                  while (lc < 0)
                    lc += n;
                  while (lc >= n)
                    lc -= n;
                }
                else
                {
                  if (lc < 0)
                    lc += n;
                  if (lc > b)
                    lc -= n;
                }
                sl[q].e[d] = lc;
              }
            }
            bool coarser = false, finer = false, all_same = true;
            for (int q = 0; q < 4; ++q)
            {
              if (q == ownq)
                continue;
              if (sl[q].valid && sl[q].rel < 0)
                coarser = true;
              if (sl[q].valid && sl[q].rel > 0)
                finer = true;
              if (!(sl[q].valid && sl[q].rel == 0))
                all_same = false;
            }
            int64_t srcv[8];
            int     nsrc = 0;
            auto    push = [&](int64_t v) { srcv[nsrc++] = v; };
            int     cnt = 0;
            double               ratio = 1.0;
            bool                 write = false;
            const int64_t own_idx = eidx(o, comp, e[0], e[1], e[2]);
            if (coarser || all_same)
            {
              for (int q = 0; q < 4; ++q)
              {
                if (q == ownq)
                {
                  push(own_idx);
                  ++cnt;
                  continue;
                }
                const bool take = all_same ? (sl[q].valid && !sl[q].out)
                                           : (sl[q].valid && sl[q].rel == 0 && !sl[q].out);
                if (take)
                {
                  push(eidx(sl[q].oct, comp, sl[q].e[0], sl[q].e[1], sl[q].e[2]));
                  ++cnt;
                }
              }
              write = true;
            }
            else if (finer)
            {
              for (int q = 0; q < 4; ++q)
              {
                if (q == ownq)
                  continue;
                if (sl[q].valid && sl[q].rel > 0 && !sl[q].out)
                {
                  push(eidx(sl[q].oct, comp, sl[q].e[0], sl[q].e[1], sl[q].e[2]));
                  if (dim == 3)
                  {
                    long e2v[3] = { sl[q].e[0], sl[q].e[1], sl[q].e[2] };
                    e2v[a] += 1;
                    push(eidx(sl[q].oct, comp, e2v[0], e2v[1], e2v[2]));
                  }
                  ++cnt;
                }
              }
              if (cnt > 0)
              {
                ratio = surface_ratio;
                write = true;
              }
            }
            if (!write)
              continue;
            // a lone own value divided by one is the identity: nothing to correct
            if (nsrc == 1 && srcv[0] == own_idx && cnt == 1)
              continue;
            EmfFix f;
            f.dst = own_idx;
            f.first = int32_t(sx.size());
            f.nsrc = int16_t(nsrc);
            f.cnt = int16_t(cnt);
            f.ratio = ratio;
            fx.push_back(f);
            for (int q = 0; q < nsrc; ++q)
              sx.push_back(srcv[q]);
          }
    }
  };
  // leaves in contiguous ranges on host threads, concatenated in leaf order (the list is the
  // same for any thread count)
  const int nthr = std::max(1, std::min<int>(int(std::thread::hardware_concurrency()), std::min(16, nowned / 64)));
  std::vector<std::vector<EmfFix>>  fxs(nthr);
  std::vector<std::vector<int64_t>> sxs(nthr);
  auto run = [&](int t) {
    const int lo = int(long(nowned) * t / nthr), hi = int(long(nowned) * (t + 1) / nthr);
    // room for the block-edge lines of every leaf; faces beside another level add more
    fxs[t].reserve(size_t(hi - lo) * ncomp * 4 * (b + 1));
    sxs[t].reserve(fxs[t].capacity() * 4);
    for (int o = lo; o < hi; ++o)
      one_leaf(o, fxs[t], sxs[t]);
  };
  if (nthr == 1)
    run(0);
  else
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; ++t)
      pool.emplace_back(run, t);
    for (auto & th : pool)
      th.join();
  }
  size_t nf = 0, nsx = 0;
  for (int t = 0; t < nthr; ++t)
    nf += fxs[t].size(), nsx += sxs[t].size();
  if (nsx > size_t(INT32_MAX))
  {
    err = "border-edge gather list too large";
    return false;
  }
  emf_fix.resize(nf);
  emf_src.resize(nsx);
  std::vector<size_t> f0(nthr + 1, 0), s0(nthr + 1, 0);
  for (int t = 0; t < nthr; ++t)
  {
    f0[t + 1] = f0[t] + fxs[t].size();
    s0[t + 1] = s0[t] + sxs[t].size();
  }
  auto gather = [&](int t) {
    const int32_t base = int32_t(s0[t]);
    EmfFix *      dst = emf_fix.data() + f0[t];
    for (size_t i = 0; i < fxs[t].size(); ++i)
    {
      dst[i] = fxs[t][i];
      dst[i].first += base;
    }
    std::copy(sxs[t].begin(), sxs[t].end(), emf_src.begin() + long(s0[t]));
    std::vector<EmfFix>().swap(fxs[t]);
    std::vector<int64_t>().swap(sxs[t]);
  };
  if (nthr == 1)
    gather(0);
  else
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; ++t)
      pool.emplace_back(gather, t);
    for (auto & th : pool)
      th.join();
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
  if (!build_emf_fix())
    return false;
  return true;
}

} // namespace mhdct
