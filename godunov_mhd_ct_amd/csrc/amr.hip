// amr.hip -- the device side of the AMR cycle (SolverGodunovMHD::do_amr_cycle,
// SolverGodunovMHD.cpp:263-319): the refinement indicator and flags of every owned block
// (compute_refine_flags, :1094-1185), and the remap of U and face B from the leaves before an
// adaptation to the leaves after it (map_userdata_after_adapt, :1191-1269).
//
// The reference delegates both to kalypsso::core (ComputeRefineFlags, UserDataRemapper), which
// is not in its tree.  Restated here:
//  * the indicator is Loehner's second-derivative error estimate (R. Loehner, "An adaptive
//    finite element scheme for transient problems in CFD", CMAME 61, 1987), unsplit (with the
//    mixed derivatives), with the noise filter epsilon_lohner, on one conserved variable of U;
//    a block's indicator is the maximum over its cells;
//  * cells: SIMPLE_COPY prolongation (a fine cell takes its parent cell's value) and
//    restriction by the mean of the 2^d fine cells (summed in pairs, so that restricting a
//    prolongation is exact);
//  * face B: prolongation by injection on the parent's face planes, and inside each parent cell
//    the mean of the opposite fine faces plus a minimum-norm divergence correction (every fine
//    cell divergence free; k_prolong_faces); restriction by the mean of the 2^(d-1) fine faces,
//    summed in pairs (the parent's flux is the fine fluxes' sum).  On a leaf's boundary plane
//    across which an old leaf of the same level lies, that leaf's faces are kept, so shared
//    faces stay single valued.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mhdct.h"
#include "amr.hpp"
#include "common.hpp"
#include "handles.hpp"

#include <mutex>

namespace mhdct
{

namespace
{

constexpr int kThreads = 256;

// Refinement variable v of block o at its cell c inside [0, b)^dim: 0-4 a conserved variable of U;
// 10-17 a derived quantity (ComputeDerivedQuantities, utils/ComputeDerivedQuantities.h:120-228)
// from the cell's primitives: thermal pressure, speed of sound, specific kinetic energy, magnetic
// pressure, local Mach number, vx, vy, vz.  The reference computes the derived array for the
// owned blocks only into a zero-initialised view (:104, 109-111): outside blocks read as zero.
__device__ inline double
field_cell(const DevMesh & m, const Idx & I, const Gas<double> & g, const double * __restrict__ U,
           const double * __restrict__ B, int v, int o, const int c[3])
{
  const long ncell = I.cells();
  const long ci = I.cell(c[0], c[1], I.dim == 3 ? c[2] : 0);
  if (v < 10)
    return U[(long(o) * 5 + v) * ncell + ci];
  if (o >= m.nowned)
    return 0.0;
  double u[8], q[8];
  for (int n = 0; n < 5; ++n)
    u[n] = U[(long(o) * 5 + n) * ncell + ci];
  const double * Bo = B + long(o) * I.fper();
  const int      i = c[0], j = c[1], k = c[2];
  if (I.dim == 2)
  {
    u[kA] = 0.5 * (Bo[I.face(0, i, j, 0)] + Bo[I.face(0, i + 1, j, 0)]);
    u[kB] = 0.5 * (Bo[I.face(1, i, j, 0)] + Bo[I.face(1, i, j + 1, 0)]);
    u[kC] = 0.5 * (Bo[I.face(2, i, j, 0)] + Bo[I.face(2, i, j, 0)]);
  }
  else
  {
    u[kA] = 0.5 * (Bo[I.face(0, i, j, k)] + Bo[I.face(0, i + 1, j, k)]);
    u[kB] = 0.5 * (Bo[I.face(1, i, j, k)] + Bo[I.face(1, i, j + 1, k)]);
    u[kC] = 0.5 * (Bo[I.face(2, i, j, k)] + Bo[I.face(2, i, j, k + 1)]);
  }
  to_primitive(u, q, g);
  switch (v)
  {
  case 10: return q[kP];
  case 11: return sqrt(g.gamma0 * q[kP] / q[kD]);
  case 12:
    return I.dim == 2 ? 0.5 * (q[kU] * q[kU] + q[kV] * q[kV]) : 0.5 * (q[kU] * q[kU] + q[kV] * q[kV] + q[kW] * q[kW]);
  case 13: return 0.5 * (q[kA] * q[kA] + q[kB] * q[kB] + q[kC] * q[kC]);
  case 14:
  {
    double un = q[kU] * q[kU] + q[kV] * q[kV];
    if (I.dim == 3)
      un += q[kW] * q[kW];
    un = sqrt(un);
    const double cs = sqrt(g.gamma0 * q[kP] / q[kD]);
    return un / cs;
  }
  case 15: return q[kU];
  case 16: return q[kV];
  default: return I.dim == 3 ? q[kW] : 0.0;
  }
}

// variable v of block o at cell c (at most one cell outside [0, b)^dim): the ghost cell rule
// of the primitive conversion -- same level: copy; finer: mean of the 2^dim fine cells (x
// fastest, then y, then z); coarser: the containing cell (SIMPLE_COPY)
__device__ inline double
u_ghosted(const DevMesh & m, const Idx & I, const Gas<double> & g, const double * __restrict__ U,
          const double * __restrict__ B, int v, int o, const int c[3])
{
  const int b = I.b;
  bool      inside = true;
  for (int d = 0; d < I.dim; ++d)
    inside = inside && c[d] >= 0 && c[d] < b;
  if (inside)
    return field_cell(m, I, g, U, B, v, o, c);
  const Hit h = neighbour(m, I, o, c);
  if (h.rel > 0)
  {
    double s = 0;
    for (int dk = 0; dk < (I.dim == 3 ? 2 : 1); ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di)
        {
          const int f[3] = { h.l[0] + di, h.l[1] + dj, h.l[2] + dk };
          s += field_cell(m, I, g, U, B, v, h.oct, f);
        }
    return s / double(1 << I.dim);
  }
  return field_cell(m, I, g, U, B, v, h.oct, h.l);
}

// SIMPLE_GRADIENT (kalypsso::core's ComputeRefineFlags, absent from the reference tree; restated):
// the largest normalised jump |q_n - q_c| / max(|q_n|, |q_c|) between the cell and its 2 dim face
// neighbours (zero where both vanish), directions x, y, z, the low neighbour first
template <int dim>
__device__ inline double
simple_gradient(const double * s)
{
  auto at = [&](int dx, int dy, int dz) { return s[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]; };
  const double u0 = at(0, 0, 0);
  double       e = 0.0;
#pragma unroll
  for (int a = 0; a < dim; ++a)
#pragma unroll
    for (int sg = -1; sg <= 1; sg += 2)
    {
      const double un = at(a == 0 ? sg : 0, a == 1 ? sg : 0, a == 2 ? sg : 0);
      const double mx = fmax(fabs(un), fabs(u0));
      const double x = mx > 0.0 ? fabs(un - u0) / mx : 0.0;
      e = fmax(e, x);
    }
  return e;
}

// Loehner's estimate at cell (i, j, k) of block o from the 3^dim stencil s[(dz+1)*9+(dy+1)*3+dx+1]:
//   E^2 = sum_kl N_kl^2 / sum_kl D_kl^2
//   N_kk = (u+ - u0) - (u0 - u-),  D_kk = |u+ - u0| + |u0 - u-| + eps (|u+| + 2|u0| + |u-|)
//   N_kl = ((u++ - u+-) - (u-+ - u--)) / 4
//   D_kl = (|u++ - u+-| + |u-+ - u--|) / 4 + eps (|u++| + |u+-| + |u-+| + |u--|) / 4   (k != l)
// summed over k, l in dimension order (k outer), E = 0 where the denominator vanishes.  The
// dimension is a template parameter: every stencil index is then a constant and s stays in
// registers.
template <int dim>
__device__ inline double
lohner(const double * s, double eps)
{
  auto at = [&](int dx, int dy, int dz) { return s[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]; };
  auto off = [](int d, int sg, int * o3) { o3[d] += sg; };
  double num = 0, den = 0;
#pragma unroll
  for (int k = 0; k < dim; ++k)
#pragma unroll
    for (int l = 0; l < dim; ++l)
    {
      if (k == l)
      {
        int p[3] = { 0, 0, 0 }, q[3] = { 0, 0, 0 };
        off(k, 1, p);
        off(k, -1, q);
        const double up = at(p[0], p[1], p[2]), u0 = at(0, 0, 0), um = at(q[0], q[1], q[2]);
        const double n = (up - u0) - (u0 - um);
        const double d = fabs(up - u0) + fabs(u0 - um) + eps * (fabs(up) + 2.0 * fabs(u0) + fabs(um));
        num += n * n;
        den += d * d;
      }
      else
      {
        int pp[3] = { 0, 0, 0 }, pm[3] = { 0, 0, 0 }, mp[3] = { 0, 0, 0 }, mm[3] = { 0, 0, 0 };
        off(k, 1, pp), off(l, 1, pp);
        off(k, 1, pm), off(l, -1, pm);
        off(k, -1, mp), off(l, 1, mp);
        off(k, -1, mm), off(l, -1, mm);
        const double upp = at(pp[0], pp[1], pp[2]), upm = at(pm[0], pm[1], pm[2]);
        const double ump = at(mp[0], mp[1], mp[2]), umm = at(mm[0], mm[1], mm[2]);
        const double n = ((upp - upm) - (ump - umm)) * 0.25;
        const double d = (fabs(upp - upm) + fabs(ump - umm)) * 0.25 +
                         eps * (fabs(upp) + fabs(upm) + fabs(ump) + fabs(umm)) * 0.25;
        num += n * n;
        den += d * d;
      }
    }
  return den > 0.0 ? sqrt(num / den) : 0.0;
}

// One thread per owned cell; the block maximum (over its cells and the refinement variables)
// through the bits of the non-negative double.  Only the face neighbours are read for
// SIMPLE_GRADIENT (CRIT 1), the whole 3^dim stencil for LOHNER_UNSPLIT (CRIT 0).
template <int DIM, int CRIT>
__global__ void __launch_bounds__(kThreads)
k_refine_indicator(DevMesh m, AmrParams p, Gas<double> g, const double * __restrict__ U,
                   const double * __restrict__ B, unsigned long long * emax)
{
  const Idx  I{ DIM, m.b }; // m.dim == DIM: the launch picks the instantiation
  const long ncell = I.cells();
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = t < long(m.nowned) * ncell;
  double     e = 0.0;
  int        o = 0;
  if (live)
  {
    o = int(t / ncell);
    const long ci = t - long(o) * ncell;
    const int  b = m.b;
    const int  i = int(ci % b), j = int((ci / b) % b), k = DIM == 3 ? int(ci / (long(b) * b)) : 0;
    const int  nv = p.nvars > 0 ? p.nvars : 1;
    for (int n = 0; n < nv; ++n)
    {
      const int v = p.nvars > 0 ? p.vars[n] : p.ivar;
      double    s[27];
#pragma unroll
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx)
          {
            if (CRIT == 1 && (dx != 0) + (dy != 0) + (dz != 0) > 1)
              continue;
            const int c[3] = { i + dx, j + dy, k + dz };
            s[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] = u_ghosted(m, I, g, U, B, v, o, c);
          }
      double x = CRIT == 1 ? simple_gradient<DIM>(s) : lohner<DIM>(s, p.epsilon_lohner);
      if (!(x >= 0.0)) // NaN state: flag it as a maximal indicator
        x = __builtin_huge_val();
      e = fmax(e, x);
    }
  }
  // when every workgroup lies inside one block, reduce over the wavefront first
  if ((ncell % kThreads) == 0)
  {
    for (int w = 32; w > 0; w >>= 1)
      e = fmax(e, __shfl_xor(e, w));
    if ((threadIdx.x & 63) == 0 && live)
      atomicMax(emax + o, static_cast<unsigned long long>(__double_as_longlong(e)));
  }
  else if (live)
    atomicMax(emax + o, static_cast<unsigned long long>(__double_as_longlong(e)));
}

__global__ void __launch_bounds__(kThreads)
k_refine_flags(DevMesh m, AmrParams p, const unsigned long long * __restrict__ emax, int8_t * __restrict__ flags,
               double * __restrict__ ind)
{
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= m.nowned)
    return;
  const double e = __longlong_as_double(static_cast<long long>(emax[o]));
  const int    l = m.level[o];
  int8_t       f = 0;
  if (l < p.level_min)
    f = 1;
  else if (l > p.level_max)
    f = -1;
  else if (e > p.epsilon_refine && l < p.level_max)
    f = 1;
  else if (e < p.epsilon_coarsen && l > p.level_min)
    f = -1;
  flags[o] = f;
  ind[o] = e;
}

// words[0, n) = 0 (the per-block maxima start from zero)
__global__ void __launch_bounds__(kThreads)
k_clear_words(unsigned long long * __restrict__ words, long n)
{
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    words[i] = 0ull;
}

// ---------------------------------------------------------------------------------------------
// remap.  Copied leaves: whole-block copies (16-byte accesses).  Prolonged and restricted
// leaves: one thread per cell (all five variables) and per face; 32-bit index arithmetic.
// ---------------------------------------------------------------------------------------------
// blocks of `per` doubles: dst block pairs[2n+1] <- src block pairs[2n]
__global__ void __launch_bounds__(kThreads)
k_copy_blocks(int per, int npair, const int * __restrict__ pairs, const double * __restrict__ src,
              double * __restrict__ dst, int vec)
{
  for (int n = blockIdx.y; n < npair; n += gridDim.y)
  {
    const double * a = src + long(pairs[2 * n]) * per;
    double *       d = dst + long(pairs[2 * n + 1]) * per;
    if (vec)
    {
      const int h = per / 2;
      for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < h; t += gridDim.x * blockDim.x)
        reinterpret_cast<double2 *>(d)[t] = reinterpret_cast<const double2 *>(a)[t];
    }
    else
      for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < per; t += gridDim.x * blockDim.x)
        d[t] = a[t];
  }
}

__global__ void __launch_bounds__(kThreads)
k_remap_cells(int dim, int b, int nmoved, const int * __restrict__ moved, const RemapEntry * __restrict__ E,
              const double * __restrict__ Uo, double * __restrict__ Un)
{
  const int ncell = dim == 3 ? b * b * b : b * b;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nmoved * ncell)
    return;
  const int        nb = moved[t / ncell];
  const int        ci = t % ncell;
  const int        i = ci % b, j = (ci / b) % b, k = dim == 3 ? ci / (b * b) : 0;
  const RemapEntry e = E[nb];
  auto             cell = [b](int a, int c, int d) { return a + b * (c + b * d); };
  double *         out = Un + long(nb) * 5 * ncell + ci;
  if (e.kind == kRemapProlong)
  {
    const int ox = e.oct & 1, oy = (e.oct >> 1) & 1, oz = (e.oct >> 2) & 1;
    const int pc = cell((ox * b + i) >> 1, (oy * b + j) >> 1, dim == 3 ? (oz * b + k) >> 1 : 0);
    const double * P = Uo + long(e.src[0]) * 5 * ncell + pc;
    for (int v = 0; v < 5; ++v)
      out[long(v) * ncell] = P[long(v) * ncell];
    return;
  }
  // restriction: fine global index 2i + di of the new block's span lies in child octant
  // (2i >= b) at local 2i mod b
  const int cx = (2 * i >= b), cy = (2 * j >= b), cz = dim == 3 ? (2 * k >= b) : 0;
  const int fi = 2 * i - cx * b, fj = 2 * j - cy * b, fk = dim == 3 ? 2 * k - cz * b : 0;
  const double * F = Uo + long(e.src[cx | (cy << 1) | (cz << 2)]) * 5 * ncell;
  for (int v = 0; v < 5; ++v)
  {
    const double * src = F + long(v) * ncell;
    // pairwise sums, x pairs first: the mean of 2^dim equal values is that value exactly, so
    // restricting a prolongation gives back the parent bit for bit
    auto row = [&](int dj, int dk) { return src[cell(fi, fj + dj, fk + dk)] + src[cell(fi + 1, fj + dj, fk + dk)]; };
    double s = row(0, 0) + row(1, 0);
    if (dim == 3)
      s = s + (row(0, 1) + row(1, 1));
    out[long(v) * ncell] = s / double(1 << dim);
  }
}

// face B of restricted leaves, one thread per face
__global__ void __launch_bounds__(kThreads)
k_restrict_faces(int dim, int b, int nrest, const int * __restrict__ restricted, const RemapEntry * __restrict__ E,
                 const double * __restrict__ Bo, double * __restrict__ Bn)
{
  const Idx I{ dim, b };
  const int fper = int(I.fper());
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrest * fper)
    return;
  const int        nb = restricted[t / fper];
  const RemapEntry e = E[nb];
  long             r = t % fper;
  const long       tout = long(nb) * fper + r;
  int c = 0;
  while (c < 2 && r >= I.fsize(c))
    r -= I.fsize(c), ++c;
  const int  nx = c == 0 ? b + 1 : b, ny = c == 1 ? b + 1 : b;
  const int  f[3] = { int(r % nx), int((r / nx) % ny), dim == 3 ? int(r / (long(nx) * ny)) : 0 };
  const bool cellc = dim == 2 && c == 2; // 2D B_z is cell centred
  // a boundary face plane shared with an old leaf of this level: that leaf's faces
  if (!cellc && (f[c] == 0 || f[c] == b))
  {
    const int side = f[c] == b;
    const int o = e.nb[2 * c + side];
    if (o >= 0)
    {
      int g[3] = { f[0], f[1], f[2] };
      g[c] = side ? 0 : b;
      Bn[tout] = Bo[long(o) * fper + I.face(c, g[0], g[1], g[2])];
      return;
    }
  }
  // the children's faces covering this face: along c the plane 2 f_c of the fine span (the high
  // child holds the middle plane and the last one), across c the 2 x 2 (2D: 2) faces, summed
  // in pairs
  int ch[3] = { 0, 0, 0 }, fl[3] = { 0, 0, 0 };
  for (int d = 0; d < dim; ++d)
  {
    ch[d] = 2 * f[d] >= b;
    fl[d] = 2 * f[d] - ch[d] * b;
  }
  const double * F = Bo + long(e.src[ch[0] | (ch[1] << 1) | (ch[2] << 2)]) * fper;
  double         val;
  if (cellc)
  {
    const double s = (F[I.face(2, fl[0], fl[1], 0)] + F[I.face(2, fl[0] + 1, fl[1], 0)]) +
                     (F[I.face(2, fl[0], fl[1] + 1, 0)] + F[I.face(2, fl[0] + 1, fl[1] + 1, 0)]);
    val = s / 4.0;
  }
  else
  {
    // transverse directions t0 < t1 of component c (2D: one); pairs along t0 first
    const int t0 = c == 0 ? 1 : 0;
    const int t1 = c == 2 ? 1 : 2;
    auto      pair = [&](int d1) {
      int q[3] = { fl[0], fl[1], fl[2] };
      if (dim == 3)
        q[t1] += d1;
      const double a = F[I.face(c, q[0], q[1], q[2])];
      q[t0] += 1;
      return a + F[I.face(c, q[0], q[1], q[2])];
    };
    const double s = dim == 3 ? pair(0) + pair(1) : pair(0);
    val = s / double(1 << (dim - 1));
  }
  Bn[tout] = val;
}

// Face B of a prolonged leaf, one thread per parent cell inside it: the 2^dim fine cells of that
// parent cell get
//  * on the parent cell's face planes: the parent face (injection), except on the leaf's own
//    boundary planes across which an old leaf of this level lies, which keep that leaf's faces;
//  * inside the parent cell: the mean of the two opposite outer fine faces of the same row,
//    plus the minimum-norm correction that makes every fine cell divergence free.  With fine
//    cells as the nodes of the cube graph Q_dim and the interior faces as its edges, the
//    correction of the face between cells n0 (low) and n1 (high) is phi[n0] - phi[n1], where
//    L phi = -div and L is the graph Laplacian; its Walsh functions w_s(n) = (-1)^|s & n| are
//    eigenvectors of eigenvalue 2 |s|, so phi = -sum_{s != 0} (w_s . div) w_s / (2^dim 2 |s|).
// Each thread writes the low and interior planes of its parent cell, and the high planes on the
// leaf's high boundary, so every face is written once.  2D B_z is cell centred: copied.
template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_prolong_faces(int b, int nprol, const int * __restrict__ prolonged, const RemapEntry * __restrict__ E,
                const double * __restrict__ Bo, double * __restrict__ Bn)
{
  // the dimension is a template parameter: the face arrays below are then indexed by constants
  // and stay in registers
  constexpr int dim = DIM;
  const Idx  I{ dim, b };
  const long fper = I.fper();
  const int  h = b / 2; // parent cells per leaf edge
  const long per = dim == 3 ? long(h) * h * h : long(h) * h;
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(nprol) * per)
    return;
  const int        nbk = prolonged[t / per];
  const long       r = t % per;
  const RemapEntry e = E[nbk];
  const int        pl[3] = { int(r % h), int((r / h) % h), dim == 3 ? int(r / (long(h) * h)) : 0 }; // local
  const int        o3[3] = { e.oct & 1, (e.oct >> 1) & 1, (e.oct >> 2) & 1 };
  int              pc[3]; // parent cell in the parent's grid
  for (int d = 0; d < 3; ++d)
    pc[d] = d < dim ? o3[d] * h + pl[d] : 0;
  const double * P = Bo + long(e.src[0]) * fper;
  double *       out = Bn + long(nbk) * fper;
  const int      nc = 1 << dim;
  // outer fine faces: ofc[c][side][m], m the transverse fine index (t0 bit 0, t1 bit 1)
  double ofc[3][2][4];
  for (int c = 0; c < dim; ++c)
  {
    const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
    for (int side = 0; side < 2; ++side)
    {
      const bool edge = pl[c] == (side ? h - 1 : 0);
      const int  o = edge ? e.nb[2 * c + side] : -1;
      for (int m = 0; m < (1 << (dim - 1)); ++m)
      {
        double v;
        if (o >= 0)
        {
          int g[3] = { 0, 0, 0 };
          g[c] = side ? 0 : b;
          g[t0] = 2 * pl[t0] + (m & 1);
          if (dim == 3)
            g[t1] = 2 * pl[t1] + ((m >> 1) & 1);
          v = Bo[long(o) * fper + I.face(c, g[0], g[1], g[2])];
        }
        else
        {
          int q[3] = { pc[0], pc[1], pc[2] };
          q[c] += side;
          v = P[I.face(c, q[0], q[1], q[2])];
        }
        ofc[c][side][m] = v;
      }
    }
  }
  // interior guesses: ifc[c][m] between the two fine cells of row m along c
  double ifc[3][4];
  for (int c = 0; c < dim; ++c)
    for (int m = 0; m < (1 << (dim - 1)); ++m)
      ifc[c][m] = 0.5 * (ofc[c][0][m] + ofc[c][1][m]);
  // transverse row index m of fine cell n (bits x, y, z) for component c
  auto row = [&](int c, int n) {
    const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
    return ((n >> t0) & 1) | (dim == 3 ? (((n >> t1) & 1) << 1) : 0);
  };
  double div[8];
  for (int n = 0; n < nc; ++n)
  {
    double sdiv = 0;
    for (int c = 0; c < dim; ++c)
    {
      const int    m = row(c, n);
      const bool   hi = (n >> c) & 1;
      const double lo_f = hi ? ifc[c][m] : ofc[c][0][m];
      const double hi_f = hi ? ofc[c][1][m] : ifc[c][m];
      sdiv = sdiv + (hi_f - lo_f);
    }
    div[n] = sdiv;
  }
  double phi[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  for (int sw = 1; sw < nc; ++sw)
  {
    double w = 0;
    for (int n = 0; n < nc; ++n)
      w = (__popc(sw & n) & 1) ? w - div[n] : w + div[n];
    const double coef = w / double(nc * 2 * __popc(sw));
    for (int n = 0; n < nc; ++n)
      phi[n] = (__popc(sw & n) & 1) ? phi[n] + coef : phi[n] - coef;
  }
  for (int c = 0; c < dim; ++c)
    for (int m = 0; m < (1 << (dim - 1)); ++m)
    {
      // the two cells of row m along c
      const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
      const int n0 = ((m & 1) << t0) | (dim == 3 ? (((m >> 1) & 1) << t1) : 0);
      const int n1 = n0 | (1 << c);
      ifc[c][m] = ifc[c][m] + (phi[n0] - phi[n1]);
    }
  // write: local fine face coordinates of this parent cell's fine cells start at 2 pl
  for (int c = 0; c < dim; ++c)
  {
    const int t0 = c == 0 ? 1 : 0, t1 = c == 2 ? 1 : 2;
    for (int m = 0; m < (1 << (dim - 1)); ++m)
    {
      int g[3] = { 2 * pl[0], 2 * pl[1], 2 * pl[2] };
      g[t0] += m & 1;
      if (dim == 3)
        g[t1] += (m >> 1) & 1;
      out[I.face(c, g[0], g[1], g[2])] = ofc[c][0][m];
      g[c] += 1;
      out[I.face(c, g[0], g[1], g[2])] = ifc[c][m];
      if (pl[c] == h - 1)
      {
        g[c] += 1;
        out[I.face(c, g[0], g[1], g[2])] = ofc[c][1][m];
      }
    }
  }
  if (dim == 2)
  {
    const double bz = P[I.face(2, pc[0], pc[1], 0)];
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di)
        out[I.face(2, 2 * pl[0] + di, 2 * pl[1] + dj, 0)] = bz;
  }
}

unsigned
grid_of(long n)
{
  return unsigned((n + kThreads - 1) / kThreads);
}

} // namespace

#define AMR_HIP(errs, x)                                                                           \
  do                                                                                               \
  {                                                                                                \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess)                                                                          \
    {                                                                                              \
      errs = std::string(#x) + ": " + hipGetErrorString(e_);                                       \
      return -1;                                                                                   \
    }                                                                                              \
  } while (0)

// The indicator and flags of every owned block into the device buffers emax (n words), ind
// (n doubles) and fl (n bytes), then read back into the host arrays; emax is cleared first, by
// hipMemsetAsync (clear_by_kernel false) or by k_clear_words.
int
refine_flags_into(Context & c, const double * U, const double * B, const AmrParams & p, int8_t * flags_host,
                  double * ind_host, hipStream_t s, unsigned long long * emax, double * ind, int8_t * fl,
                  bool clear_by_kernel)
{
  const DevMesh & m = c.dm;
  const long      n = m.nowned;
  if (clear_by_kernel)
    hipLaunchKernelGGL(k_clear_words, grid_of(n), kThreads, 0, s, emax, n);
  else
    AMR_HIP(c.err, hipMemsetAsync(emax, 0, size_t(n) * sizeof(unsigned long long), s));
  const Shapes      S{ c.dim, c.b };
  const Gas<double> g = gas_of(c.st);
  const unsigned    grid = grid_of(n * S.cells());
  if (c.dim == 3 && p.criterion == 1)
    hipLaunchKernelGGL((k_refine_indicator<3, 1>), grid, kThreads, 0, s, m, p, g, U, B, emax);
  else if (c.dim == 3)
    hipLaunchKernelGGL((k_refine_indicator<3, 0>), grid, kThreads, 0, s, m, p, g, U, B, emax);
  else if (p.criterion == 1)
    hipLaunchKernelGGL((k_refine_indicator<2, 1>), grid, kThreads, 0, s, m, p, g, U, B, emax);
  else
    hipLaunchKernelGGL((k_refine_indicator<2, 0>), grid, kThreads, 0, s, m, p, g, U, B, emax);
  hipLaunchKernelGGL(k_refine_flags, grid_of(n), kThreads, 0, s, m, p, emax, fl, ind);
  AMR_HIP(c.err, hipGetLastError());
  AMR_HIP(c.err, hipMemcpyAsync(flags_host, fl, size_t(n), hipMemcpyDeviceToHost, s));
  if (ind_host)
    AMR_HIP(c.err, hipMemcpyAsync(ind_host, ind, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, s));
  AMR_HIP(c.err, hipStreamSynchronize(s));
  return 0;
}

int
launch_refine_flags(Context & c, const double * U, const double * B, const AmrParams & p, int8_t * flags_host,
                    double * ind_host, hipStream_t s)
{
  const long n = c.dm.nowned;
  if (n == 0)
    return 0;
  const size_t need = size_t(n) * (2 * sizeof(unsigned long long) + 1) + 256;
  if (c.amr_bytes < need)
  {
    if (c.amr_buf)
      (void)hipFree(c.amr_buf);
    c.amr_buf = nullptr;
    c.amr_bytes = 0;
    AMR_HIP(c.err, hipMalloc(&c.amr_buf, need));
    c.amr_bytes = need;
  }
  auto *   emax = static_cast<unsigned long long *>(c.amr_buf);
  double * ind = reinterpret_cast<double *>(emax + n);
  auto *   fl = reinterpret_cast<int8_t *>(ind + n);
  return refine_flags_into(c, U, B, p, flags_host, ind_host, s, emax, ind, fl, false);
}

bool
remap_plan(const Mesh & old_mesh, const Mesh & new_mesh, std::vector<RemapEntry> & plan, RemapLists & lists,
           int counts[3], std::string & err)
{
  std::vector<int> copies, moved, restricted, prolonged;
  if (old_mesh.dim != new_mesh.dim || old_mesh.b != new_mesh.b)
  {
    err = "remap: the meshes differ in dimension or block size";
    return false;
  }
  for (int d = 0; d < 3; ++d)
    if (old_mesh.nbrick[d] != new_mesh.nbrick[d])
    {
      err = "remap: the meshes differ in brick size";
      return false;
    }
  const int dim = old_mesh.dim;
  std::unordered_map<uint64_t, int> old_index;
  old_index.reserve(old_mesh.leaves.size() * 2);
  for (size_t i = 0; i < old_mesh.leaves.size(); ++i)
  {
    const Leaf & L = old_mesh.leaves[i];
    old_index[Mesh::key(L.level, L.x, L.y, L.z)] = int(i);
  }
  auto find = [&](int l, int x, int y, int z) -> int {
    if (l < 0)
      return -1;
    const auto it = old_index.find(Mesh::key(l, x, y, z));
    return it == old_index.end() ? -1 : it->second;
  };
  for (int d = 0; d < 3; ++d)
    if (old_mesh.periodic[d] != new_mesh.periodic[d])
    {
      err = "remap: the meshes differ in periodicity";
      return false;
    }
  // the old leaf of level l across face plane (c, side) of block (l, x, y, z), or -1
  auto across = [&](int l, int x, int y, int z, int c, int side) -> int {
    long q[3] = { x, y, z };
    q[c] += side ? 1 : -1;
    const long n = old_mesh.nblocks(l, c);
    if (q[c] < 0 || q[c] >= n)
    {
      if (!old_mesh.periodic[c])
        return -1;
      q[c] = (q[c] + n) % n;
    }
    return find(l, int(q[0]), int(q[1]), int(q[2]));
  };
  counts[0] = counts[1] = counts[2] = 0;
  plan.assign(new_mesh.leaves.size(), RemapEntry{});
  for (size_t n = 0; n < new_mesh.leaves.size(); ++n)
  {
    const Leaf & L = new_mesh.leaves[n];
    RemapEntry & e = plan[n];
    std::fill(e.src, e.src + 8, -1);
    const int same = find(L.level, L.x, L.y, L.z);
    if (same >= 0)
    {
      e.kind = kRemapCopy;
      e.src[0] = same;
      copies.push_back(same);
      copies.push_back(int(n));
      ++counts[0];
      continue;
    }
    const int parent = find(L.level - 1, L.x >> 1, L.y >> 1, dim == 3 ? (L.z >> 1) : 0);
    if (parent >= 0)
    {
      e.kind = kRemapProlong;
      e.src[0] = parent;
      e.oct = (L.x & 1) | ((L.y & 1) << 1) | (dim == 3 ? ((L.z & 1) << 2) : 0);
      for (int c = 0; c < dim; ++c)
        for (int side = 0; side < 2; ++side)
          e.nb[2 * c + side] = across(L.level, L.x, L.y, L.z, c, side);
      prolonged.push_back(int(n));
      moved.push_back(int(n));
      ++counts[1];
      continue;
    }
    bool all = true;
    for (int ch = 0; ch < (1 << dim) && all; ++ch)
    {
      const int s = find(L.level + 1, 2 * L.x + (ch & 1), 2 * L.y + ((ch >> 1) & 1),
                         dim == 3 ? 2 * L.z + ((ch >> 2) & 1) : 0);
      all = s >= 0;
      e.src[ch] = s;
    }
    if (!all)
    {
      err = "remap: a new leaf is neither an old leaf, nor a child of one, nor the parent of old leaves "
            "(more than one level of change in one cycle)";
      return false;
    }
    e.kind = kRemapRestrict;
    for (int c = 0; c < dim; ++c)
      for (int side = 0; side < 2; ++side)
        e.nb[2 * c + side] = across(L.level, L.x, L.y, L.z, c, side);
    restricted.push_back(int(n));
    moved.push_back(int(n));
    ++counts[2];
  }
  lists.ncopy = int(copies.size() / 2);
  lists.nmoved = int(moved.size());
  lists.nrest = int(restricted.size());
  lists.nprol = int(prolonged.size());
  lists.ints = copies;
  lists.ints.insert(lists.ints.end(), moved.begin(), moved.end());
  lists.ints.insert(lists.ints.end(), restricted.begin(), restricted.end());
  lists.ints.insert(lists.ints.end(), prolonged.begin(), prolonged.end());
  return true;
}

int
launch_remap(int dim, int b, const RemapEntry * plan_dev, const RemapLists & L, const int * lists_dev, const double * Uo,
             const double * Bo, double * Un, double * Bn, hipStream_t s, std::string & err)
{
  const int ncell = dim == 3 ? b * b * b : b * b;
  const int fper = dim == 3 ? 3 * (b + 1) * b * b : 2 * (b + 1) * b + b * b;
  if (long(L.nmoved) * fper > long(INT32_MAX) || long(L.nmoved) * ncell > long(INT32_MAX))
  {
    err = "remap: too many prolonged or restricted leaves for one launch";
    return -1;
  }
  if (L.ncopy > 0)
  {
    // whole blocks; 16-byte accesses when every block start is 16-byte aligned
    const auto aligned = [](const void * p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const int  vec = aligned(Uo) && aligned(Un) && aligned(Bo) && aligned(Bn) ? 1 : 0; // per is even
    const dim3 gu(std::min<unsigned>(grid_of(vec ? 5 * ncell / 2 : 5 * ncell), 64), std::min(L.ncopy, 65535));
    const dim3 gb(std::min<unsigned>(grid_of(vec ? fper / 2 : fper), 64), std::min(L.ncopy, 65535));
    hipLaunchKernelGGL(k_copy_blocks, gu, kThreads, 0, s, 5 * ncell, L.ncopy, L.copies(lists_dev), Uo, Un, vec);
    hipLaunchKernelGGL(k_copy_blocks, gb, kThreads, 0, s, fper, L.ncopy, L.copies(lists_dev), Bo, Bn, vec);
  }
  if (L.nmoved > 0)
    hipLaunchKernelGGL(k_remap_cells, grid_of(long(L.nmoved) * ncell), kThreads, 0, s, dim, b, L.nmoved,
                       L.moved(lists_dev), plan_dev, Uo, Un);
  if (L.nrest > 0)
    hipLaunchKernelGGL(k_restrict_faces, grid_of(long(L.nrest) * fper), kThreads, 0, s, dim, b, L.nrest,
                       L.restricted(lists_dev), plan_dev, Bo, Bn);
  if (L.nprol > 0)
  {
    const long per = dim == 3 ? long(b / 2) * (b / 2) * (b / 2) : long(b / 2) * (b / 2);
    if (dim == 3)
      hipLaunchKernelGGL(k_prolong_faces<3>, grid_of(long(L.nprol) * per), kThreads, 0, s, b, L.nprol,
                         L.prolonged(lists_dev), plan_dev, Bo, Bn);
    else
      hipLaunchKernelGGL(k_prolong_faces<2>, grid_of(long(L.nprol) * per), kThreads, 0, s, b, L.nprol,
                         L.prolonged(lists_dev), plan_dev, Bo, Bn);
  }
  AMR_HIP(err, hipGetLastError());
  return 0;
}

// The checks of the refinement entry points, and their parameters in device form
int
check_amr_params(Context & c, const double * U, const double * B, const mhdct_amr_params * p, AmrParams & out)
{
  if (!c.has_mesh)
  {
    c.err = "no mesh set";
    return -1;
  }
  if (c.dm.nowned > 0 && !U)
  {
    c.err = "null state pointer";
    return -1;
  }
  const int nv = p->nvars;
  bool      ok = p->level_min >= 0 && p->level_max >= p->level_min && p->epsilon_lohner >= 0.0 && nv >= 0 &&
          nv <= 4 && (p->criterion == MHDCT_REFINE_LOHNER_UNSPLIT || p->criterion == MHDCT_REFINE_SIMPLE_GRADIENT);
  bool derived = false;
  for (int n = 0; ok && n < (nv > 0 ? nv : 1); ++n)
  {
    const int v = nv > 0 ? p->vars[n] : p->ivar;
    ok = (v >= 0 && v <= 4) || (nv > 0 && v >= MHDCT_RVAR_THERMAL_PRESSURE && v <= MHDCT_RVAR_VZ);
    derived = derived || v >= MHDCT_RVAR_THERMAL_PRESSURE;
  }
  if (!ok)
  {
    c.err = "refine flags: invalid parameters (0 <= level_min <= level_max, epsilon_lohner >= 0, criterion "
            "LOHNER_UNSPLIT or SIMPLE_GRADIENT, 0 to 4 variables 0-4 or MHDCT_RVAR_*)";
    return -1;
  }
  if (derived && c.dm.nowned > 0 && !B)
  {
    c.err = "refine flags: a derived quantity needs the face field B";
    return -1;
  }
  out.level_min = p->level_min;
  out.level_max = p->level_max;
  out.epsilon_refine = p->epsilon_refine;
  out.epsilon_coarsen = p->epsilon_coarsen;
  out.epsilon_lohner = p->epsilon_lohner;
  out.ivar = p->ivar;
  out.criterion = p->criterion;
  out.nvars = nv;
  for (int n = 0; n < 4; ++n)
    out.vars[n] = n < nv ? p->vars[n] : 0;
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Refinement flags without allocation: the per-block maxima, indicators and flags live in a
// file-scope device array ("scratch") sized for MHDCT_NOALLOC_MAX_BLOCKS owned blocks, one copy
// per GPU for the life of the process.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name is only a host-side stand-in) and kept per device.
// ---------------------------------------------------------------------------------------------
namespace
{

constexpr long kScratchWords = 2L * MHDCT_NOALLOC_MAX_BLOCKS + MHDCT_NOALLOC_MAX_BLOCKS / 8 + 1;

__device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mutex;
void *        g_scratch_ptr[kMaxDevices] = {};

// the device address of this device's copy of scratch (the current device), or nullptr
void *
scratch_address(int device, std::string & err)
{
  if (device < 0 || device >= kMaxDevices)
  {
    err = "refine flags (no allocation): device index out of range";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device])
  {
    void *           ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
    {
      err = std::string("refine flags (no allocation): hipGetSymbolAddress: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

} // namespace

int
launch_refine_flags_noalloc(Context & c, const double * U, const double * B, const AmrParams & p,
                            int8_t * flags_host, double * ind_host, hipStream_t s)
{
  const long n = c.dm.nowned;
  if (n > MHDCT_NOALLOC_MAX_BLOCKS)
  {
    c.err = "refine flags (no allocation): " + std::to_string(n) + " owned blocks, more than the " +
            std::to_string(long(MHDCT_NOALLOC_MAX_BLOCKS)) + " its scratch holds";
    return -8;
  }
  if (n == 0)
    return 0;
  int dev = -1;
  AMR_HIP(c.err, hipGetDevice(&dev));
  if (dev != c.device)
  {
    c.err = "refine flags (no allocation): the context's device is not the current device";
    return -1;
  }
  void * base = scratch_address(dev, c.err);
  if (!base)
    return -1;
  auto *   emax = static_cast<unsigned long long *>(base);
  double * ind = reinterpret_cast<double *>(emax + n);
  auto *   fl = reinterpret_cast<int8_t *>(ind + n);
  return refine_flags_into(c, U, B, p, flags_host, ind_host, s, emax, ind, fl, true);
}

} // namespace mhdct

extern "C"
{
  int
  mhdct_refine_flags_noalloc(mhdct_ctx * h, const double * U, const double * B, const mhdct_amr_params * p,
                             int8_t * flags, double * indicator, void * stream)
  {
    if (!h || !p || !flags)
      return -1;
    mhdct::Context & c = h->c;
    mhdct::AmrParams ap;
    if (mhdct::check_amr_params(c, U, B, p, ap))
      return -1;
    return mhdct::launch_refine_flags_noalloc(c, U, B, ap, flags, indicator, static_cast<hipStream_t>(stream));
  }
}
