// common.hpp -- device helpers shared by the kernel forms: index arithmetic of the reference
// layout, the neighbour-table lookup and the face-B ghost rule.
#pragma once

#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "phys.hpp"

namespace mhdct
{

// ---------------------------------------------------------------------------
// index helpers (layout A0: block slowest, then variable, then cell)
// ---------------------------------------------------------------------------
struct Idx
{
  int dim, b;

  __device__ long cells() const { return dim == 3 ? long(b) * b * b : long(b) * b; }
  __device__ long cell(int i, int j, int k) const { return i + long(b) * (j + long(b) * k); }
  __device__ long gn(int g) const { return b + 2 * g; }
  __device__ long gcells(int g) const
  {
    const long n = b + 2 * g;
    return dim == 3 ? n * n * n : n * n;
  }
  __device__ long gcell(int g, int i, int j, int k) const
  {
    const long n = b + 2 * g;
    return (i + g) + n * ((j + g) + n * (dim == 3 ? (k + g) : 0));
  }
  __device__ long fsize(int c) const
  {
    if (dim == 2)
      return c == 2 ? long(b) * b : long(b + 1) * b;
    return long(b + 1) * b * b;
  }
  __device__ long foff(int c) const { return c == 0 ? 0 : (c == 1 ? fsize(0) : fsize(0) + fsize(1)); }
  __device__ long fper() const { return fsize(0) + fsize(1) + fsize(2); }
  __device__ long face(int c, int i, int j, int k) const
  {
    const long nx = c == 0 ? b + 1 : b;
    const long ny = c == 1 ? b + 1 : b;
    return foff(c) + i + nx * (j + ny * (dim == 3 ? k : 0));
  }
  __device__ long gfsize(int c) const
  {
    const long n = b + 4;
    if (dim == 2)
      return c == 2 ? n * n : (n + 1) * n;
    return (n + 1) * n * n;
  }
  __device__ long gfoff(int c) const { return c == 0 ? 0 : (c == 1 ? gfsize(0) : gfsize(0) + gfsize(1)); }
  __device__ long gfper() const { return gfsize(0) + gfsize(1) + gfsize(2); }
  __device__ long gface(int c, int i, int j, int k) const
  {
    const long n = b + 4;
    const long nx = c == 0 ? n + 1 : n;
    const long ny = c == 1 ? n + 1 : n;
    return gfoff(c) + (i + 2) + nx * ((j + 2) + ny * (dim == 3 ? (k + 2) : 0));
  }
  __device__ long ecells() const
  {
    const long n = b + 1;
    return dim == 3 ? n * n * n : n * n;
  }
  __device__ long edge(int i, int j, int k) const
  {
    const long n = b + 1;
    return i + n * (j + n * k);
  }
};

// The same index arithmetic with the block size fixed at compile time (32-bit, constant
// divisors); one instantiation per block size class.
template <int DIM, int BS>
struct IdxC
{
  static constexpr int dim = DIM;
  static constexpr int b = BS;

  __host__ __device__ static constexpr int cells() { return DIM == 3 ? BS * BS * BS : BS * BS; }
  __host__ __device__ static constexpr int cell(int i, int j, int k) { return i + BS * (j + BS * k); }
  __host__ __device__ static constexpr int gcells(int g)
  {
    return DIM == 3 ? (BS + 2 * g) * (BS + 2 * g) * (BS + 2 * g) : (BS + 2 * g) * (BS + 2 * g);
  }
  __host__ __device__ static constexpr int gcell(int g, int i, int j, int k)
  {
    return (i + g) + (BS + 2 * g) * ((j + g) + (BS + 2 * g) * (DIM == 3 ? (k + g) : 0));
  }
  __host__ __device__ static constexpr int fsize(int c)
  {
    return DIM == 2 ? (c == 2 ? BS * BS : (BS + 1) * BS) : (BS + 1) * BS * BS;
  }
  __host__ __device__ static constexpr int foff(int c) { return c == 0 ? 0 : (c == 1 ? fsize(0) : fsize(0) + fsize(1)); }
  __host__ __device__ static constexpr int fper() { return fsize(0) + fsize(1) + fsize(2); }
  __host__ __device__ static constexpr int face(int c, int i, int j, int k)
  {
    return foff(c) + i + (c == 0 ? BS + 1 : BS) * (j + (c == 1 ? BS + 1 : BS) * (DIM == 3 ? k : 0));
  }
  __host__ __device__ static constexpr int gfsize(int c)
  {
    return DIM == 2 ? (c == 2 ? (BS + 4) * (BS + 4) : (BS + 5) * (BS + 4)) : (BS + 5) * (BS + 4) * (BS + 4);
  }
  __host__ __device__ static constexpr int gfoff(int c)
  {
    return c == 0 ? 0 : (c == 1 ? gfsize(0) : gfsize(0) + gfsize(1));
  }
  __host__ __device__ static constexpr int gfper() { return gfsize(0) + gfsize(1) + gfsize(2); }
  __host__ __device__ static constexpr int gface(int c, int i, int j, int k)
  {
    return gfoff(c) + (i + 2) + (c == 0 ? BS + 5 : BS + 4) * ((j + 2) + (c == 1 ? BS + 5 : BS + 4) * (DIM == 3 ? (k + 2) : 0));
  }
  __host__ __device__ static constexpr int ecells() { return DIM == 3 ? (BS + 1) * (BS + 1) * (BS + 1) : (BS + 1) * (BS + 1); }
  __host__ __device__ static constexpr int edge(int i, int j, int k) { return i + (BS + 1) * (j + (BS + 1) * k); }
};

// Workgroup w is dispatched to XCD w % 8 (round robin), and each XCD has its own L2.  Grids are
// launched as multiples of 8 and renumbered so that each XCD takes one contiguous eighth of the
// work groups: cells whose stencils overlap then run on the same XCD and share its L2.
constexpr int kXcd = 8;

__device__ inline long
xcd_group()
{
  const unsigned w = blockIdx.x;
  return long(w % kXcd) * (gridDim.x / kXcd) + w / kXcd;
}

struct Hit
{
  int oct;
  int rel;
  int l[3];
};

// Leaf holding ghost cell c of block `oct` (c outside [0,b)^dim) from the neighbour table.
// IX is the index type (Idx with run-time sizes, or IdxC with compile-time ones).
template <class IX>
__device__ inline Hit
neighbour(const DevMesh & m, const IX & I, int oct, const int c[3])
{
  const int b = I.b;
  int       dir = 0, sub = 0, mul = 1;
  for (int d = 0; d < I.dim; ++d)
  {
    const int dl = c[d] < 0 ? -1 : (c[d] >= b ? 1 : 0);
    dir += (dl + 1) * mul;
    mul *= 3;
    if (dl == 0 && c[d] >= b / 2)
      sub |= 1 << d;
  }
  const int     ndir = I.dim == 3 ? 27 : 9;
  const int     nsub = 1 << I.dim;
  const NbEntry e = m.nb[(long(oct) * ndir + dir) * nsub + sub];
  Hit           h;
  h.oct = e.oct;
  h.rel = e.rel;
  for (int d = 0; d < 3; ++d)
  {
    const int cd = d < I.dim ? c[d] : 0;
    if (d >= I.dim)
      h.l[d] = 0;
    else if (e.rel == 0)
      h.l[d] = cd + e.off[d];
    else if (e.rel < 0)
      h.l[d] = (cd + e.off[d]) >> 1; // floor division by 2
    else
      h.l[d] = 2 * cd + e.off[d];
  }
  return h;
}

__device__ inline Hit
neighbour(const DevMesh & m, int oct, const int c[3])
{
  return neighbour(m, Idx{ m.dim, m.b }, oct, c);
}

// ---------------------------------------------------------------------------
// Ghost cells beside a coarser block, EXTRAPOLATE_LINEAR_MINMOD
// (linear_extrapolate_hydro_vars_using_limited_slopes, ConvertToPrimitivesVariablesFunctor.cpp:
// 471-635).  The five conserved variables at coarse-level cell c of block oct as the slopes read
// them: the cell itself inside the block; beyond it the neighbour's cell (same level), the mean
// of its 2^dim cells (finer) or its containing cell (coarser); false beyond an outside block,
// which has no neighbour table.
// ---------------------------------------------------------------------------
template <class IX>
__device__ inline bool
stencil_u(const DevMesh & m, const IX & I, const double * __restrict__ U, int oct, const int c[3], double * u)
{
  const int b = I.b;
  bool      inside = true;
  for (int d = 0; d < I.dim; ++d)
    inside = inside && c[d] >= 0 && c[d] < b;
  const long ncell = I.cells();
  if (inside)
  {
    const long ci = I.cell(c[0], c[1], I.dim == 3 ? c[2] : 0);
    for (int v = 0; v < 5; ++v)
      u[v] = U[(long(oct) * 5 + v) * ncell + ci];
    return true;
  }
  if (oct >= m.nowned)
    return false;
  const Hit h = neighbour(m, I, oct, c);
  if (h.oct < 0)
    return false;
  if (h.rel > 0)
  {
    for (int v = 0; v < 5; ++v)
    {
      const double * Un = U + (long(h.oct) * 5 + v) * ncell;
      double         sum = 0;
      for (int dk = 0; dk < (I.dim == 3 ? 2 : 1); ++dk)
        for (int dj = 0; dj < 2; ++dj)
          for (int di = 0; di < 2; ++di)
            sum += Un[I.cell(h.l[0] + di, h.l[1] + dj, h.l[2] + dk)];
      u[v] = sum / double(1 << I.dim);
    }
  }
  else
  {
    const long ci = I.cell(h.l[0], h.l[1], h.l[2]);
    for (int v = 0; v < 5; ++v)
      u[v] = U[(long(h.oct) * 5 + v) * ncell + ci];
  }
  return true;
}

// The conserved variables of ghost cell (i, j, k) from the coarse cell h that contains it:
// u = u_c + 1/4 s_x du/dx + 1/4 s_y du/dy (+ 1/4 s_z du/dz), summed in that order, du/dx the
// minmod slope (slope_type 1) over the coarse cell's neighbours along x, zero when one of them
// lies beyond an outside block; s = -1 / +1 in the low / high half of the coarse cell.
template <class IX>
__device__ inline void
extrapolate_linear_minmod(const DevMesh & m, const IX & I, const double * __restrict__ U, const Hit & h, int i,
                          int j, int k, double * u)
{
  const int dim = I.dim;
  const int cell[3] = { i, j, k };
  const int c0[3] = { h.l[0], h.l[1], h.l[2] };
  double    u0[5];
  stencil_u(m, I, U, h.oct, c0, u0);
  for (int v = 0; v < 5; ++v)
    u[v] = u0[v];
  for (int d = 0; d < dim; ++d)
  {
    int cp[3] = { c0[0], c0[1], c0[2] }, cm[3] = { c0[0], c0[1], c0[2] };
    cp[d] += 1;
    cm[d] -= 1;
    double     up[5], um[5];
    const bool ok = stencil_u(m, I, U, h.oct, cp, up) & stencil_u(m, I, U, h.oct, cm, um);
    const double sgn = (cell[d] & 1) ? 1.0 : -1.0;
    for (int v = 0; v < 5; ++v)
    {
      const double slope = ok ? limited_slope(u0[v], up[v], um[v], 1.0) : 0.0;
      u[v] = u[v] + 0.25 * sgn * slope;
    }
  }
}

__device__ inline double
dx_of(const DevMesh & m, double scaling, int oct)
{
  return 1.0 / double(long(m.b) << m.level[oct]) * scaling;
}

// ---------------------------------------------------------------------------
// Face B of block `o` at face (c, f) of its 2-ghost-layer face grid: the block's own face when
// inside, else the ghost rule (FillBlockGhostFacesFunctor, GodunovImplemV0.cpp:224-235):
// same level -> copy; coarser -> coarse face on a coarse face plane, else the mean of the two
// bracketing coarse faces; finer -> mean of the 2^(dim-1) fine faces.  A face beside the block
// that lies on a leaf boundary takes the leaf at this block's level, else the finer one.
// ---------------------------------------------------------------------------
template <class IX>
__device__ inline double
face_value_ghosted(const DevMesh & m, const IX & I, const double * __restrict__ B, int o, int c, const int fin[3])
{
  const int b = I.b, dim = I.dim;
  int       f[3] = { fin[0], fin[1], fin[2] };
  const bool cellc = (dim == 2 && c == 2);
  bool       inner = true;
  for (int d = 0; d < dim; ++d)
  {
    const int hi = (d == c && !cellc) ? b : b - 1;
    if (f[d] < 0 || f[d] > hi)
      inner = false;
  }
  const double * Bo = B + long(o) * I.fper();
  double         v;
  if (inner)
  {
    v = Bo[I.face(c, f[0], f[1], f[2])];
  }
  else if (cellc)
  {
    const Hit      h = neighbour(m, I, o, f);
    const double * Bn = B + long(h.oct) * I.fper();
    if (h.rel <= 0)
      v = Bn[I.face(2, h.l[0], h.l[1], 0)];
    else
    {
      double s = 0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di)
          s += Bn[I.face(2, h.l[0] + di, h.l[1] + dj, 0)];
      v = s / 4.0;
    }
  }
  else
  {
    // the face separates cells lo (f[c]-1) and hi (f[c]); beyond the block along c take the
    // nearer one, beside it prefer the leaf at this block's level, then the finer (ties: hi)
    int  cc[3] = { f[0], f[1], f[2] };
    bool high = true;
    Hit  h;
    if (f[c] > b)
    {
      cc[c] = f[c] - 1;
      high = false;
      h = neighbour(m, I, o, cc);
    }
    else if (f[c] < 0)
      h = neighbour(m, I, o, cc);
    else
    {
      const Hit hh = neighbour(m, I, o, cc);
      int       lo[3] = { f[0], f[1], f[2] };
      lo[c] -= 1;
      const Hit hl = neighbour(m, I, o, lo);
      // same level first, then finer, then coarser
      const int rl = hl.rel == 0 ? 2 : (hl.rel > 0 ? 1 : 0);
      const int rh = hh.rel == 0 ? 2 : (hh.rel > 0 ? 1 : 0);
      if (rl > rh)
      {
        h = hl;
        cc[c] = lo[c];
        high = false;
      }
      else
        h = hh;
    }
    const double * Bn = B + long(h.oct) * I.fper();
    if (h.rel == 0)
    {
      int q[3] = { h.l[0], h.l[1], h.l[2] };
      if (!high)
        q[c] += 1;
      v = Bn[I.face(c, q[0], q[1], q[2])];
    }
    else if (h.rel < 0)
    {
      // fine face position relative to the coarse block origin, in fine units
      const int    ndir = dim == 3 ? 27 : 9;
      int          dir = 0, sub = 0, mul = 1;
      for (int d = 0; d < dim; ++d)
      {
        const int dl = cc[d] < 0 ? -1 : (cc[d] >= b ? 1 : 0);
        dir += (dl + 1) * mul;
        mul *= 3;
        if (dl == 0 && cc[d] >= b / 2)
          sub |= 1 << d;
      }
      const NbEntry e = m.nb[(long(o) * ndir + dir) * (1 << dim) + sub];
      const int     F = f[c] + e.off[c];
      int           q[3] = { h.l[0], h.l[1], h.l[2] };
      if ((F & 1) == 0)
      {
        q[c] = F >> 1;
        v = Bn[I.face(c, q[0], q[1], q[2])];
      }
      else
      {
        q[c] = (F - 1) >> 1;
        const double lo = Bn[I.face(c, q[0], q[1], q[2])];
        q[c] += 1;
        const double hi = Bn[I.face(c, q[0], q[1], q[2])];
        v = 0.5 * (lo + hi);
      }
    }
    else
    {
      int base[3] = { h.l[0], h.l[1], h.l[2] };
      if (!high)
        base[c] += 2;
      if (dim == 2)
      {
        const int t2 = c == 0 ? 1 : 0;
        double    s = 0;
        for (int a = 0; a < 2; ++a)
        {
          int q[3] = { base[0], base[1], 0 };
          q[t2] += a;
          s += Bn[I.face(c, q[0], q[1], 0)];
        }
        v = s / 2.0;
      }
      else
      {
        const int t1 = c == 0 ? 1 : 0;
        const int t2 = c == 2 ? 1 : 2;
        double    s = 0;
        for (int a2 = 0; a2 < 2; ++a2)
          for (int a1 = 0; a1 < 2; ++a1)
          {
            int q[3] = { base[0], base[1], base[2] };
            q[t1] += a1;
            q[t2] += a2;
            s += Bn[I.face(c, q[0], q[1], q[2])];
          }
        v = s / 4.0;
      }
    }
  }
  return v;
}

} // namespace mhdct
