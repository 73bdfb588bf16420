// border.hpp -- the gather list of CorrectEmfAtBlockBorder
// (scheme/CorrectEmfAtBlockBorderFunctor.cpp:94-360), derived on the device from the neighbour
// table, one work item per edge of a block's border ring (built once per mesh by border.hip).
//
// The border edges of a block are, per EMF component (3D: the edge axis a; 2D: z), for each
// index along a, the ring of the (e1, e2) square over the two tangential axes with e1 or e2 on
// the block border: 4b edges.  They are numbered per component corners first (the block-edge
// lines, which always need work), then each side's b - 1 inner edges side by side, so that a
// wavefront mostly holds edges of one kind: the inner edges of a side beside a leaf of this
// level all leave at once.  The four cells around an edge are the quadrants (s1, s2)
// (cell = edge index - 1 + s); the block's own quadrant rounds the edge index down; the other
// three are found through the neighbour table.  With any coarser sibling the corrected EMF is the mean of the block's own
// value and its same-level, not-outside siblings; with all siblings at this level, the mean of
// all four (outside ones excluded); else with finer siblings, the sum of their edges (two along
// the edge in 3D) / count / 2^(d-1).  Sources are summed in quadrant order, the block's own one
// in its place.  An edge inside one face whose neighbour has this block's level needs none.
#pragma once

#include "common.hpp"

namespace mhdct
{

// work items per block: EMF components x edge positions along the edge x ring length
__host__ __device__ inline int
border_slots_per_block(int dim, int b)
{
  return (dim == 3 ? 3 * b : 1) * 4 * b;
}

struct BorderEdge
{
  int  comp, a, t1, t2;
  long e[3];
};

// v[i] of a 3-vector and v[i] = x, written as selects over the three elements so that a
// runtime index keeps the vector in registers (an indexed private array would live in scratch)
template <class T>
__device__ inline T
get3(const T v[3], int i)
{
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

template <class T>
__device__ inline void
set3(T v[3], int i, T x)
{
  v[0] = i == 0 ? x : v[0];
  v[1] = i == 1 ? x : v[1];
  v[2] = i == 2 ? x : v[2];
}

// src[n++] = x, as a select over the eight slots (src stays in registers)
__device__ inline void
push8(long src[8], int & n, long x)
{
#pragma unroll
  for (int k = 0; k < 8; ++k)
    src[k] = k == n ? x : src[k];
  ++n;
}

__device__ inline BorderEdge
border_edge(int dim, int b, int slot)
{
  const int  nal = dim == 3 ? b : 1;
  const int  per_comp = nal * 4 * b;
  BorderEdge E;
  E.comp = slot / per_comp;
  const int r = slot - E.comp * per_comp;
  E.a = dim == 3 ? E.comp : 2;
  E.t1 = E.a == 0 ? 1 : 0;
  E.t2 = E.a == 2 ? 1 : 2;
  int e1, e2, ea;
  if (r < 4 * nal)
  {
    // corners (e1, e2) in {0, b}^2
    ea = r >> 2;
    e1 = (r & 1) ? b : 0;
    e2 = (r & 2) ? b : 0;
  }
  else
  {
    // inner edges of side 0 (e2 = 0), 1 (e2 = b), 2 (e1 = 0), 3 (e1 = b)
    const int j = r - 4 * nal;
    const int side = j / (nal * (b - 1));
    const int rem = j - side * nal * (b - 1);
    ea = rem / (b - 1);
    const int mid = 1 + rem % (b - 1);
    e1 = side < 2 ? mid : (side == 2 ? 0 : b);
    e2 = side < 2 ? (side == 0 ? 0 : b) : mid;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    E.e[d] = d == E.t1 ? e1 : (d == E.t2 ? e2 : (dim == 3 && d == E.a ? ea : 0));
  return E;
}

// flat index of edge e of component comp of block o in the EMF array ((b+1)^d per component)
__device__ inline long
edge_index(int dim, int b, long o, int comp, const long e[3])
{
  const long n1 = b + 1;
  const long ne = dim == 3 ? n1 * n1 * n1 : n1 * n1;
  const int  ncomp = dim == 3 ? 3 : 1;
  return (o * ncomp + comp) * ne + e[0] + n1 * (e[1] + n1 * e[2]);
}

// The sources of border slot `slot` of owned block o: false when the edge needs no correction,
// else dst (its index in the EMF array), src[0..nsrc) (summed in this order), cnt and ratio
// (corrected EMF = sum / cnt / ratio).
__device__ inline bool
border_sources(const DevMesh & m, int o, int slot, long & dst, long src[8], int & nsrc, int & cnt, double & ratio)
{
  const int        dim = m.dim, b = m.b;
  const BorderEdge E = border_edge(dim, b, slot);
  const long *     e = E.e;
  const int        t1 = E.t1, t2 = E.t2, a = E.a;
  const int        e1 = int(get3(e, t1)), e2 = int(get3(e, t2));
  const bool       b1 = (e1 == 0 || e1 == b), b2 = (e2 == 0 || e2 == b);
  const int        nd = dim == 3 ? 27 : 9, ns = 1 << dim;
  if (b1 != b2)
  {
    // an edge inside one face: nothing to correct beside a leaf of this level
    const int fd = b1 ? t1 : t2;
    const int fs = (b1 ? e1 : e2) == b ? 1 : 0;
    int       dl[3] = { 0, 0, 0 };
    set3(dl, fd, fs ? 1 : -1);
    const int     di = (dl[0] + 1) + 3 * (dl[1] + 1) + (dim == 3 ? 9 * (dl[2] + 1) : 0);
    const NbEntry F = m.nb[(long(o) * nd + di) * ns];
    if (F.oct >= 0 && F.rel == 0)
      return false;
  }
  const int  level = m.level[o];
  const long org[3] = { long(m.coord[3 * o]) * b, long(m.coord[3 * o + 1]) * b, long(m.coord[3 * o + 2]) * b };
  const int  ownq = (e1 < b ? 1 : 0) + 2 * (e2 < b ? 1 : 0);
  bool       valid[4] = { false, false, false, false }, out[4] = { false, false, false, false };
  int        soct[4] = { -1, -1, -1, -1 }, srel[4] = { 0, 0, 0, 0 };
  long       se[4][3];
  for (int q = 0; q < 4; ++q)
  {
    se[q][0] = se[q][1] = se[q][2] = 0;
    if (q == ownq)
      continue;
    long c[3] = { e[0], e[1], e[2] };
    set3(c, t1, get3(e, t1) - 1 + (q & 1));
    set3(c, t2, get3(e, t2) - 1 + (q >> 1));
    bool inside = true;
    for (int d = 0; d < dim; ++d)
      if (c[d] < 0 || c[d] >= b)
        inside = false;
    if (inside)
      continue;
    int dir = 0, sub = 0, mul = 1;
    for (int d = 0; d < dim; ++d)
    {
      const int dl = c[d] < 0 ? -1 : (c[d] >= b ? 1 : 0);
      dir += (dl + 1) * mul;
      mul *= 3;
      if (dl == 0 && c[d] >= b / 2)
        sub |= 1 << d;
    }
    const NbEntry N = m.nb[(long(o) * nd + dir) * ns + sub];
    if (N.oct < 0)
      continue;
    valid[q] = true;
    soct[q] = N.oct;
    srel[q] = N.rel;
    out[q] = N.oct >= m.nowned;
    for (int d = 0; d < dim; ++d)
    {
      const int  Ls = level + N.rel;
      const long n = long(m.nbrick[d]) * b << Ls;
      const long G = org[d] + e[d];
      const long Gs = N.rel == 1 ? 2 * G : (N.rel == 0 ? G : (G >> 1)); // floor(G / 2)
      long       lc = Gs - long(m.coord[3 * N.oct + d]) * b;
      if (d == a)
      {
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
      se[q][d] = lc;
    }
  }
  bool coarser = false, finer = false, all_same = true;
  for (int q = 0; q < 4; ++q)
  {
    if (q == ownq)
      continue;
    if (valid[q] && srel[q] < 0)
      coarser = true;
    if (valid[q] && srel[q] > 0)
      finer = true;
    if (!(valid[q] && srel[q] == 0))
      all_same = false;
  }
  const long own = edge_index(dim, b, o, E.comp, e);
  nsrc = 0;
  cnt = 0;
  ratio = 1.0;
  if (coarser || all_same)
  {
    for (int q = 0; q < 4; ++q)
    {
      if (q == ownq)
      {
        push8(src, nsrc, own);
        ++cnt;
        continue;
      }
      const bool take = all_same ? (valid[q] && !out[q]) : (valid[q] && srel[q] == 0 && !out[q]);
      if (take)
      {
        push8(src, nsrc, edge_index(dim, b, soct[q], E.comp, se[q]));
        ++cnt;
      }
    }
    // a lone own value divided by one is the identity: nothing to correct
    if (nsrc == 1 && cnt == 1)
      return false;
  }
  else if (finer)
  {
    for (int q = 0; q < 4; ++q)
    {
      if (q == ownq || !(valid[q] && srel[q] > 0 && !out[q]))
        continue;
      push8(src, nsrc, edge_index(dim, b, soct[q], E.comp, se[q]));
      if (dim == 3)
      {
        long e2v[3] = { se[q][0], se[q][1], se[q][2] };
        set3(e2v, a, get3(e2v, a) + 1);
        push8(src, nsrc, edge_index(dim, b, soct[q], E.comp, e2v));
      }
      ++cnt;
    }
    if (cnt == 0)
      return false;
    ratio = double(1 << (dim - 1));
  }
  else
    return false;
  dst = own;
  return true;
}

} // namespace mhdct
