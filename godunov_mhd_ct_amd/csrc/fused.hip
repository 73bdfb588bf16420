// fused.hip -- the fused form of the MHD-CT time step.
//
// The 16 device operations of GodunovImplemV0::do_time_step (scheme/GodunovImplemV0.cpp:101-207)
// are regrouped so that each intermediate crosses HBM at most once:
//
//   k_prim         face-B ghost fill + primitive conversion (GodunovImplemV0.cpp:128-139)
//   k_efield       edge electric field v x B at t_n (:141-149)
//   k_cell         limited slopes, Hancock predictor, face-B source, transverse
//                  face-B slopes: one compact record per cell (:141-156, and the slopes the EMF
//                  reconstruction reads, ComputeEmfAndStoreFunctor.h:363-419)
//   k_border_flux  HLLD on the block-border faces that a coarser neighbour sums
//   k_flux_tile    HLLD on the faces of an 8^dim tile, one direction at a time through LDS, and
//                  the conservative update U2 = ((U + dX) + dY) + dZ in FP64 registers (:161-177)
//   k_edge         2D-HLLD EMF, one thread per edge, every edge computed once (:197)
//   k_emf_fix      CorrectEmfAtBlockBorder from the precomputed gather list (:628-638)
//   k_emf_scatter  corrected values back into the EMF array
//   k_bupdate      face-B update by the discrete curl (:641)
//
// T is the arithmetic type of the bulk (slopes, predictor, HLLD, EMF): double for the FP64
// form, float for the mixed-precision form; primitive recovery from total energy, the
// conservative accumulation and the face-B accumulation are FP64 in both.
//
// Every kernel is instantiated per (dimension, block size): the block geometry is a set of
// compile-time constants, so index arithmetic is 32-bit with constant divisors.  A launch walks
// a work list of one size class that may span several meshes ("slots"): each slot's range of
// work items is padded to whole 512-item groups, so the slot of a wavefront is uniform.
// A launch has one workgroup per work group (grid-stride beyond a cap), numbered XCD by XCD.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "fused.hpp"
#include "phys_fast.hpp"

namespace mhdct
{

namespace
{

constexpr int kTS = 8; // tile edge of the flux kernel

// Calls CALL with the compile-time block size BS: BSEL when the launch is specialised to one
// block size, else the one matching the run-time value bs (8, 16 or 32; a slot property,
// uniform across the wavefront).
#define MHDCT_BS_DISPATCH(bs, CALL)                                                                                    \
  do                                                                                                                   \
  {                                                                                                                    \
    if constexpr (BSEL != 0)                                                                                           \
    {                                                                                                                  \
      constexpr int BS = BSEL;                                                                                         \
      CALL;                                                                                                            \
    }                                                                                                                  \
    else if ((bs) == 8)                                                                                                \
    {                                                                                                                  \
      constexpr int BS = 8;                                                                                            \
      CALL;                                                                                                            \
    }                                                                                                                  \
    else if ((bs) == 16)                                                                                               \
    {                                                                                                                  \
      constexpr int BS = 16;                                                                                           \
      CALL;                                                                                                            \
    }                                                                                                                  \
    else                                                                                                               \
    {                                                                                                                  \
      constexpr int BS = 32;                                                                                           \
      CALL;                                                                                                            \
    }                                                                                                                  \
  } while (0)

// compile-time geometry of one block size class
template <int DIM, int BS>
struct Geo
{
  using I = IdxC<DIM, BS>;
  static constexpr int N2 = BS + 4;  // cells per axis with 2 ghost layers
  static constexpr int N1 = BS + 2;  // with 1 ghost layer
  static constexpr int NE = BS + 1;  // edge-grid points per axis
  static constexpr int G2 = DIM == 3 ? N2 * N2 * N2 : N2 * N2;
  static constexpr int G1 = DIM == 3 ? N1 * N1 * N1 : N1 * N1;
  static constexpr int NC = I::cells();
  static constexpr int NFS = DIM == 3 ? 6 : 2;           // transverse face-B slopes per cell
  static constexpr int NT = DIM == 3 ? BS * BS : BS;      // faces on one block side
  static constexpr int NCE = DIM == 3 ? 3 : 1;            // EMF components
  static constexpr int NEC = I::ecells();                 // edge-grid points per component
  static constexpr int EPD = DIM == 3 ? BS * NE * NE : NE * NE; // edges per direction
  // k_edge (3D): the plane wavefronts of the x-, y- and z-edges, one (i, j) position per lane
  // (NPLX, NPLY, NE x NE); EWB: k_edge wavefronts per block (2D: one edge per lane)
  static constexpr int NPLY = (BS + 1) * BS;
  static constexpr int NWP = (NPLY + 63) / 64;
  static constexpr int NPLX = BS * (BS + 1);
  static constexpr int NWX = (NPLX + 63) / 64;
  static constexpr int NWZ = (NE * NE + 63) / 64;
  static constexpr int EWB = fused_edge_waves(DIM, BS);
  static_assert(DIM == 2 || EWB == NWX + NWP + NWZ, "k_edge wavefronts per block");
  static constexpr int NTP = (NT + 63) / 64 * 64;          // k_border_flux items per side
  static constexpr int GSH = G2 - NC;                      // ghost shell of the 2-ghost tile
  static constexpr int PRIM = NC + (GSH + 63) / 64 * 64;  // k_prim items: interior, then shell
  static constexpr int NP = BS + 3;                       // edge points per axis of the E grid
  static constexpr int NPE = DIM == 3 ? NP * NP * NP : NP * NP;
  static constexpr int EFP = DIM == 3 ? (NP * NP + 63) / 64 * 64 : (NPE + 63) / 64 * 64; // k_efield items
  static constexpr int NB1 = BS / kTS;                    // flux tiles per axis
  static constexpr int TPB = DIM == 3 ? NB1 * NB1 * NB1 : NB1 * NB1;
  static_assert(BS % kTS == 0, "block size must be a multiple of the flux tile");
};

__device__ __forceinline__ int
slot_of(const long * off, int nslots, long t)
{
  int s = 0;
  while (s + 1 < nslots && t >= off[s + 1])
    ++s;
  return s;
}

// Raw buffer resource over `bytes` bytes at `base` (bounds-checked vector loads with a 32-bit
// offset; word 3 selects the 32-bit data format of gfx9-class buffer descriptors).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t
rsrc(const void * base, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T
bload(__amdgpu_buffer_rsrc_t r, int idx)
{
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4, 0, 0));
}

// Non-temporal store (cache policy "nt"): the scratch records are read back by later kernels,
// long after they would have left the L2; streaming them past it raised the write rate from
// 4.3 to 4.8 TB/s in a 41-stream store test on MI355X.
constexpr int kCacheNT = 2;

template <typename T>
__device__ __forceinline__ void
bstore(__amdgpu_buffer_rsrc_t r, int idx, T x)
{
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r,
                                          idx * 8, 0, kCacheNT);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, idx * 4, 0, kCacheNT);
}

// value x of the lane whose byte address (4 x lane) is addr
template <typename T>
__device__ __forceinline__ T
lane_take(T x, int addr)
{
  if constexpr (sizeof(T) == 8)
  {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u))));
    const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u >> 32))));
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  }
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, x)));
}

template <typename T>
__device__ __forceinline__ void
store_nt(T * p, T x)
{
  __builtin_nontemporal_store(x, p);
}

// View of one block of one slot.  The block index must be uniform across the
// wavefront (every work list pads its per-block range to whole wavefronts): the six arrays are
// then addressed through buffer descriptors held in scalar registers.
template <typename T, int DIM, int BS>
struct View
{
  using G = Geo<DIM, BS>;
  const FusedSlot<T> &   S;
  __amdgpu_buffer_rsrc_t r_qp, r_bg, r_q2, r_sl, r_sf, r_fs;
  __device__ View(const FusedSlot<T> & s, int o)
    : S(s)
    , r_qp(rsrc(s.Qp + long(o) * 5 * G::G2, 5 * G::G2 * int(sizeof(T))))
    , r_bg(rsrc(s.Bg + long(o) * G::I::gfper(), G::I::gfper() * int(sizeof(T))))
    , r_q2(rsrc(s.Q2 + long(o) * 8 * G::G1, 8 * G::G1 * int(sizeof(T))))
    , r_sl(rsrc(s.S + long(o) * DIM * 8 * G::G1, DIM * 8 * G::G1 * int(sizeof(T))))
    , r_sf(rsrc(s.sF + long(o) * DIM * G::G1, DIM * G::G1 * int(sizeof(T))))
    , r_fs(rsrc(s.FS + long(o) * G::NFS * G::G1, G::NFS * G::G1 * int(sizeof(T))))
  {
  }
  __device__ static int c2(int i, int j, int k)
  {
    return (i + 2) + G::N2 * ((j + 2) + (DIM == 3 ? G::N2 * (k + 2) : 0));
  }
  __device__ static int c1(int i, int j, int k)
  {
    return (i + 1) + G::N1 * ((j + 1) + (DIM == 3 ? G::N1 * (k + 1) : 0));
  }
  __device__ T qp(int v, int i, int j, int k) const { return bload<T>(r_qp, v * G::G2 + c2(i, j, k)); }
  __device__ T bf(int c, int i, int j, int k) const { return bload<T>(r_bg, G::I::gface(c, i, j, k)); }
  __device__ T q2(int v, int c) const { return bload<T>(r_q2, v * G::G1 + c); }
  __device__ T sl(int d, int v, int c) const { return bload<T>(r_sl, (d * 8 + v) * G::G1 + c); }
  __device__ T sf(int d, int c) const { return bload<T>(r_sf, d * G::G1 + c); }
  __device__ T fs(int n, int c) const { return bload<T>(r_fs, n * G::G1 + c); }
};

// Where the cell and edge stages read the primitive state, the ghosted face field and the edge
// electric field: the HBM scratch of k_prim / k_efield (GlobalSrc) or a front tile in LDS
// (LdsSrc, k_front).  Both give the same values, so the records are the same bit for bit.
template <typename T, int DIM, int BS>
struct GlobalSrc
{
  using G = Geo<DIM, BS>;
  const View<T, DIM, BS> & V;
  __amdgpu_buffer_rsrc_t   r_ed;
  __device__ GlobalSrc(const View<T, DIM, BS> & v, const FusedSlot<T> & S, int o)
    : V(v)
    , r_ed(rsrc(S.Ed + long(o) * G::NCE * G::NPE, G::NCE * G::NPE * int(sizeof(T))))
  {
  }
  __device__ T qp(int v, int a, int c, int e) const { return V.qp(v, a, c, e); }
  __device__ T bf(int fc, int a, int c, int e) const { return V.bf(fc, a, c, e); }
  // component comp (0 x, 1 y, 2 z; 2D: 0 = z) of E at edge point (a, c, e), a in [-1, b+1]
  __device__ T ed(int comp, int a, int c, int e) const
  {
    return bload<T>(r_ed, comp * G::NPE + (a + 1) + G::NP * ((c + 1) + (DIM == 3 ? G::NP * (e + 1) : 0)));
  }
};

// Cell (i,j,k) of the g-th ghost cell of the 2-ghost tile, g < (b+4)^dim - b^dim.  3D: two
// full planes below the block, two above, then for each interior plane two full rows below,
// two above and the four side cells of each interior row; 2D: the same for rows and cells.
template <int DIM, int BS>
__host__ __device__ inline void
shell_cell(int g, int & i, int & j, int & k)
{
  constexpr int n = BS + 4;
  constexpr int ring = 4 * n + 4 * BS; // ghost cells of one interior plane (3D) / the 2D shell
  int           rem = g;
  k = 0;
  if (DIM == 3)
  {
    constexpr int P = n * n;
    if (rem < 4 * P)
    {
      k = rem < 2 * P ? rem / P - 2 : BS + (rem - 2 * P) / P;
      rem %= P;
      j = rem / n - 2;
      i = rem % n - 2;
      return;
    }
    rem -= 4 * P;
    k = rem / ring;
    rem %= ring;
  }
  if (rem < 2 * n)
  {
    j = rem / n - 2;
    i = rem % n - 2;
  }
  else if (rem < 4 * n)
  {
    j = BS + (rem - 2 * n) / n;
    i = (rem - 2 * n) % n - 2;
  }
  else
  {
    rem -= 4 * n;
    j = rem / 4;
    const int q = rem % 4;
    i = q < 2 ? q - 2 : BS + q - 2;
  }
}

template <int BS>
__device__ __forceinline__ double
dxd(const DevMesh & m, double scaling, int o)
{
  return 1.0 / double(long(BS) << m.level[o]) * scaling;
}

// ---------------------------------------------------------------------------------------------
// k_prim: U, B -> primitive (rho, p, u, v, w) on the 2-ghost tile and the ghosted face field.
// The cell owns its low faces (and its high faces on the last ghost layer).
// ---------------------------------------------------------------------------------------------
// The conserved state u[8] (5 of U, cell-centred B) of cell (i, j, k) of the 2-ghost tile of block
// o: the cell itself inside the block (inside_state); a ghost cell from the neighbour (ghost_state:
// same level: copy; finer: mean of the 2^dim cells and this block's ghosted faces; coarser:
// SIMPLE_COPY of the coarse cell, or EXTRAPOLATE_LINEAR_MINMOD and this block's ghosted faces).
// gf(c, a0, a1, a2) is the ghosted face value face_value_ghosted gives.
template <typename T, int DIM, int BS>
__device__ __forceinline__ void
inside_state(const FusedSlot<T> & S, int o, int i, int j, int k, double * u, double * lo)
{
  using IX = IdxC<DIM, BS>;
  constexpr int  ncell = IX::cells();
  const int      cidx = IX::cell(i, j, k);
  const double * Bo = S.B + long(o) * IX::fper();
  for (int v = 0; v < 5; ++v)
    u[v] = S.U[(long(o) * 5 + v) * ncell + cidx];
  // the low faces (also the cell's ghosted face values: an interior cell's faces are its own)
  lo[0] = Bo[IX::face(0, i, j, DIM == 3 ? k : 0)];
  lo[1] = Bo[IX::face(1, i, j, DIM == 3 ? k : 0)];
  lo[2] = Bo[IX::face(2, i, j, DIM == 3 ? k : 0)];
  if (DIM == 2)
  {
    u[kA] = 0.5 * (lo[0] + Bo[IX::face(0, i + 1, j, 0)]);
    u[kB] = 0.5 * (lo[1] + Bo[IX::face(1, i, j + 1, 0)]);
    u[kC] = 0.5 * (lo[2] + lo[2]);
  }
  else
  {
    u[kA] = 0.5 * (lo[0] + Bo[IX::face(0, i + 1, j, k)]);
    u[kB] = 0.5 * (lo[1] + Bo[IX::face(1, i, j + 1, k)]);
    u[kC] = 0.5 * (lo[2] + Bo[IX::face(2, i, j, k + 1)]);
  }
}

template <typename T, int DIM, int BS, class GF>
__device__ __forceinline__ void
ghost_state(const FusedSlot<T> & S, int o, int i, int j, int k, double * u, const GF & gf)
{
  using IX = IdxC<DIM, BS>;
  constexpr IX    I{};
  constexpr int   ncell = IX::cells();
  const DevMesh & m = S.m;
  const double * __restrict__ U = S.U;
  const double * __restrict__ B = S.B;
  const int  c[3] = { i, j, k };
  const Hit  h = neighbour(m, I, o, c);
  const bool lin = h.rel < 0 && m.prolong == 1;
  if (h.rel > 0 || lin)
  {
    if (lin)
      extrapolate_linear_minmod(m, I, U, h, i, j, k, u);
    else
      for (int v = 0; v < 5; ++v)
      {
        const double * Un = U + (long(h.oct) * 5 + v) * ncell;
        double         sum = 0;
        for (int dk = 0; dk < (DIM == 3 ? 2 : 1); ++dk)
          for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di)
              sum += Un[IX::cell(h.l[0] + di, h.l[1] + dj, h.l[2] + dk)];
        u[v] = sum / double(1 << DIM);
      }
    // the cell-centred field from this block's ghosted faces (finer neighbour, and the
    // linear prolongation: ConvertToPrimitivesVariablesFunctor.cpp:536-541, 623-628)
    if (DIM == 2)
    {
      u[kA] = 0.5 * (gf(0, i, j, 0) + gf(0, i + 1, j, 0));
      u[kB] = 0.5 * (gf(1, i, j, 0) + gf(1, i, j + 1, 0));
      u[kC] = gf(2, i, j, 0);
    }
    else
    {
      u[kA] = 0.5 * (gf(0, i, j, k) + gf(0, i + 1, j, k));
      u[kB] = 0.5 * (gf(1, i, j, k) + gf(1, i, j + 1, k));
      u[kC] = 0.5 * (gf(2, i, j, k) + gf(2, i, j, k + 1));
    }
  }
  else
  {
    const int cidx = IX::cell(h.l[0], h.l[1], h.l[2]);
    for (int v = 0; v < 5; ++v)
      u[v] = U[(long(h.oct) * 5 + v) * ncell + cidx];
    const double * Bn = B + long(h.oct) * IX::fper();
    const int      a = h.l[0], bb = h.l[1], cc = h.l[2];
    if (DIM == 2)
    {
      u[kA] = 0.5 * (Bn[IX::face(0, a, bb, 0)] + Bn[IX::face(0, a + 1, bb, 0)]);
      u[kB] = 0.5 * (Bn[IX::face(1, a, bb, 0)] + Bn[IX::face(1, a, bb + 1, 0)]);
      u[kC] = 0.5 * (Bn[IX::face(2, a, bb, 0)] + Bn[IX::face(2, a, bb, 0)]);
    }
    else
    {
      u[kA] = 0.5 * (Bn[IX::face(0, a, bb, cc)] + Bn[IX::face(0, a + 1, bb, cc)]);
      u[kB] = 0.5 * (Bn[IX::face(1, a, bb, cc)] + Bn[IX::face(1, a, bb + 1, cc)]);
      u[kC] = 0.5 * (Bn[IX::face(2, a, bb, cc)] + Bn[IX::face(2, a, bb, cc + 1)]);
    }
  }
}

// ComputeDtMHD's rate of one cell (ComputeDtMHDFunctor.cpp:86-151) from its primitive state q:
// the sum over directions of (c_f + |v|)/dx (with gravity, cell_rate's bound)
template <int DIM, int BS, typename T>
__device__ __forceinline__ double
cell_inv_dt(const double * q, const FusedSlot<T> & S, int o)
{
  const Gas<double> & g = S.gd;
  const double        b2 = q[kA] * q[kA] + q[kB] * q[kB] + q[kC] * q[kC];
  const double        c2 = g.gamma0 * q[kP] / q[kD];
  const double        d2 = 0.5 * (b2 / q[kD] + c2);
  const double        dx = dxd<BS>(S.m, S.scaling, o);
  double              v[3];
  v[0] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kA] * q[kA] / q[kD])) + fabs(q[kU]);
  v[1] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kB] * q[kB] / q[kD])) + fabs(q[kV]);
  if (DIM == 3)
    v[2] = sqrt(d2 + sqrt(d2 * d2 - c2 * q[kC] * q[kC] / q[kD])) + fabs(q[kW]);
  return cell_rate<DIM>(v, dx, g);
}

// one item of k_prim; returns the cell's inverse time step for an owned cell of a slot that
// reduces dt here (S.dt_red), else 0
template <typename T, int DIM, int BS>
__device__ __forceinline__ double
prim_item(const FusedSlot<T> & S, long r0)
{

  using G = Geo<DIM, BS>;
  using IX = IdxC<DIM, BS>;
  constexpr IX I{};
    const DevMesh &      m = S.m;
    // interior cells first, then the ghost shell: every wavefront is one or the other
    const int o = int(r0) / G::PRIM;
    int       g = int(r0) - o * G::PRIM;
    constexpr int n = G::N2, b = BS;
    int           i, j, k = 0;
    const bool    inside = g < G::NC;
    if (inside)
    {
      i = g % BS;
      j = (g / BS) % BS;
      k = DIM == 3 ? g / (BS * BS) : 0;
    }
    else
    {
      g -= G::NC;
      if (g >= G::GSH)
        return 0.0;
      shell_cell<DIM, BS>(g, i, j, k);
    }
    const int ci = (i + 2) + n * ((j + 2) + (DIM == 3 ? n * (k + 2) : 0));
    const double * __restrict__ B = S.B;
    auto       gfv = [&](int c, int a0, int a1, int a2) {
      const int f[3] = { a0, a1, a2 };
      return face_value_ghosted(m, I, B, o, c, f);
    };
    double    u[8], lo[3] = { 0.0, 0.0, 0.0 };
    if (inside)
      inside_state<T, DIM, BS>(S, o, i, j, k, u, lo);
    else
      ghost_state<T, DIM, BS>(S, o, i, j, k, u, gfv);
    double       q[8];
    const double eint = to_primitive(u, q, S.gd);
    if (S.abort_eint && eint < 0)
      atomicOr(S.red + 1, 1ull);
    T * Qo = S.Qp + long(o) * 5 * G::G2;
    for (int v = 0; v < 5; ++v)
      store_nt(Qo + v * G::G2 + ci, T(q[v]));
    // face field: low faces of this cell, and the high faces on the last ghost layer
    T * Go = S.Bg + long(o) * IX::gfper();
    if (DIM == 2)
    {
      store_nt(Go + IX::gface(0, i, j, 0), T(gfv(0, i, j, 0)));
      store_nt(Go + IX::gface(1, i, j, 0), T(gfv(1, i, j, 0)));
      store_nt(Go + IX::gface(2, i, j, 0), T(gfv(2, i, j, 0)));
      if (i == b + 1)
        store_nt(Go + IX::gface(0, b + 2, j, 0), T(gfv(0, b + 2, j, 0)));
      if (j == b + 1)
        store_nt(Go + IX::gface(1, i, b + 2, 0), T(gfv(1, i, b + 2, 0)));
    }
    else if (inside)
    {
      // an interior cell: its low faces as read, no high faces on the last ghost layer
      store_nt(Go + IX::gface(0, i, j, k), T(lo[0]));
      store_nt(Go + IX::gface(1, i, j, k), T(lo[1]));
      store_nt(Go + IX::gface(2, i, j, k), T(lo[2]));
    }
    else
    {
      store_nt(Go + IX::gface(0, i, j, k), T(gfv(0, i, j, k)));
      store_nt(Go + IX::gface(1, i, j, k), T(gfv(1, i, j, k)));
      store_nt(Go + IX::gface(2, i, j, k), T(gfv(2, i, j, k)));
      if (i == b + 1)
        store_nt(Go + IX::gface(0, b + 2, j, k), T(gfv(0, b + 2, j, k)));
      if (j == b + 1)
        store_nt(Go + IX::gface(1, i, b + 2, k), T(gfv(1, i, b + 2, k)));
      if (k == b + 1)
        store_nt(Go + IX::gface(2, i, j, b + 2), T(gfv(2, i, j, b + 2)));
    }
    return (inside && S.dt_red) ? cell_inv_dt<DIM, BS>(q, S, o) : 0.0;
}

// k_prim, and for a slot with S.dt_red the dt reduction of ComputeDtMHD over its owned cells
// (the interior items read exactly the U and B the reduction needs): per 256 items a workgroup
// maximum, folded into the slot's reduction word by one atomic only when it exceeds the value
// there (which only grows), so that few of the workgroups write it; a NaN rate raises flag bit
// 2.  A workgroup's 256 items lie in one slot (slot ranges are padded to 512 items).  (One
// atomic check per wavefront instead, without the barriers, was slower: 2.40 against 2.35 ms on
// the mixed-block mesh.)
template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_prim(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  __shared__ double wmax[4];
  for (long g0 = xcd_group() * blockDim.x; g0 < total; g0 += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, g0));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = g0 + threadIdx.x - off[s];
    double               rate = 0.0;
    if (r0 < S.cnt[0])
      MHDCT_BS_DISPATCH(S.m.b, (rate = prim_item<T, DIM, BS>(S, r0)));
    if (!S.dt_red) // uniform across the workgroup
      continue;
    const bool nan = rate != rate;
    double     c = nan ? 0.0 : rate;
    for (int o = 32; o > 0; o >>= 1)
      c = fmax(c, __shfl_xor(c, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
      wmax[wave] = c;
    if (__any(nan) && lane == 0)
      atomicOr(S.red + 1, 2ull);
    __syncthreads();
    if (threadIdx.x == 0)
    {
      const double mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
      const unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
      if (bits > __hip_atomic_load(S.red, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(S.red, bits);
    }
    __syncthreads();
  }
}

// value v (0-4 hydro, 5-7 low and 8-10 high faces of Bx, By, Bz) of cell (x, y, z) from src
template <typename T, int DIM, class Src>
__device__ __forceinline__ T
cell_q_src(const Src & src, int v, int x, int y, int z)
{
  if (v < 5)
    return src.qp(v, x, y, z);
  const int fc = (v - 5) % 3;
  const int hi = v >= 8;
  if (DIM == 2 && fc == 2)
    return src.bf(2, x, y, 0);
  return src.bf(fc, x + (hi && fc == 0), y + (hi && fc == 1), z + (hi && fc == 2));
}

// ---------------------------------------------------------------------------------------------
// k_efield: edge electric field v x B at t_n (ComputeElectricFieldFunctor.cpp:97-235) on the
// edge points [-1, b+1]^dim of each block, from the primitive state of the 2-ghost tile.
// ---------------------------------------------------------------------------------------------
// E = v x B at edge point (a, c, e) from the primitive state and the face field of src: out[0..2]
// = E_x, E_y, E_z (3D), out[0] = E_z (2D)
template <typename T, int DIM, class Src>
__device__ __forceinline__ void
efield_point(const Src & src, int a, int c, int e, T * out)
{
  auto Q = [&](int v, int x, int y, int z) -> T {
    if (v < 5)
      return src.qp(v, x, y, z);
    const int fc = (v - 5) % 3;
    const int hi = v >= 8;
    if (DIM == 2 && fc == 2)
      return src.bf(2, x, y, 0);
    return src.bf(fc, x + (hi && fc == 0), y + (hi && fc == 1), z + (hi && fc == 2));
  };
  if (DIM == 3)
  {
    const T vv = T(0.25) * (Q(3, a, c - 1, e - 1) + Q(3, a, c - 1, e) + Q(3, a, c, e - 1) + Q(3, a, c, e));
    const T ww = T(0.25) * (Q(4, a, c - 1, e - 1) + Q(4, a, c - 1, e) + Q(4, a, c, e - 1) + Q(4, a, c, e));
    const T BB = T(0.5) * (Q(6, a, c, e - 1) + Q(6, a, c, e));
    const T CC = T(0.5) * (Q(7, a, c - 1, e) + Q(7, a, c, e));
    out[0] = vv * CC - ww * BB;
  }
  if (DIM == 3)
  {
    const T uu = T(0.25) * (Q(2, a - 1, c, e - 1) + Q(2, a - 1, c, e) + Q(2, a, c, e - 1) + Q(2, a, c, e));
    const T ww = T(0.25) * (Q(4, a - 1, c, e - 1) + Q(4, a - 1, c, e) + Q(4, a, c, e - 1) + Q(4, a, c, e));
    const T AA = T(0.5) * (Q(5, a, c, e - 1) + Q(5, a, c, e));
    const T CC = T(0.5) * (Q(7, a - 1, c, e) + Q(7, a, c, e));
    out[1] = ww * AA - uu * CC;
  }
  {
    const T uu = T(0.25) * (Q(2, a - 1, c - 1, e) + Q(2, a - 1, c, e) + Q(2, a, c - 1, e) + Q(2, a, c, e));
    const T vv = T(0.25) * (Q(3, a - 1, c - 1, e) + Q(3, a - 1, c, e) + Q(3, a, c - 1, e) + Q(3, a, c, e));
    const T AA = T(0.5) * (Q(5, a, c - 1, e) + Q(5, a, c, e));
    const T BB = T(0.5) * (Q(6, a - 1, c, e) + Q(6, a, c, e));
    out[DIM == 3 ? 2 : 0] = uu * BB - vv * AA;
  }
}

// E at the edge points of an (a, c) column, marching along z (3D): each plane e reads the 18
// values of plane e that the three components take (velocities at the four cells (a-1 .. a,
// c-1 .. c), the faces) and carries the nine that E_x and E_y take again at e + 1 as plane e - 1.
// The same sums in the same order as efield_point, bit for bit.
template <typename T, int BS>
__device__ __forceinline__ void
efield_column(const FusedSlot<T> & S, int o, int a, int c)
{
  using G = Geo<3, BS>;
  const View<T, 3, BS>      V(S, o);
  const GlobalSrc<T, 3, BS> src(V, S, o);
  auto                      Q = [&](int v, int x, int y, int z) -> T { return cell_q_src<T, 3>(src, v, x, y, z); };
  // plane e - 1: v(a, c-1), v(a, c), w(a, c-1), w(a, c), w(a-1, c), u(a-1, c), u(a, c), By(a, c),
  // Bx(a, c)
  int e = -1;
  T   pv0 = Q(3, a, c - 1, e - 1), pv1 = Q(3, a, c, e - 1), pw0 = Q(4, a, c - 1, e - 1), pw1 = Q(4, a, c, e - 1);
  T   pw2 = Q(4, a - 1, c, e - 1), pu0 = Q(2, a - 1, c, e - 1), pu1 = Q(2, a, c, e - 1);
  T   pB = Q(6, a, c, e - 1), pA = Q(5, a, c, e - 1);
  T * Ed = S.Ed + long(o) * G::NCE * G::NPE + (a + 1) + G::NP * (c + 1);
  for (; e <= BS + 1; ++e)
  {
    const T v0 = Q(3, a, c - 1, e), v1 = Q(3, a, c, e), w0 = Q(4, a, c - 1, e), w1 = Q(4, a, c, e);
    const T w2 = Q(4, a - 1, c, e), u0 = Q(2, a - 1, c, e), u1 = Q(2, a, c, e);
    const T Bc = Q(6, a, c, e), Ac = Q(5, a, c, e);
    T       out[3];
    {
      const T vv = T(0.25) * (pv0 + v0 + pv1 + v1);
      const T ww = T(0.25) * (pw0 + w0 + pw1 + w1);
      const T BB = T(0.5) * (pB + Bc);
      const T CC = T(0.5) * (Q(7, a, c - 1, e) + Q(7, a, c, e));
      out[0] = vv * CC - ww * BB;
    }
    {
      const T uu = T(0.25) * (pu0 + u0 + pu1 + u1);
      const T ww = T(0.25) * (pw2 + w2 + pw1 + w1);
      const T AA = T(0.5) * (pA + Ac);
      const T CC = T(0.5) * (Q(7, a - 1, c, e) + Q(7, a, c, e));
      out[1] = ww * AA - uu * CC;
    }
    {
      const T uu = T(0.25) * (Q(2, a - 1, c - 1, e) + u0 + Q(2, a, c - 1, e) + u1);
      const T vv = T(0.25) * (Q(3, a - 1, c - 1, e) + Q(3, a - 1, c, e) + v0 + v1);
      const T AA = T(0.5) * (Q(5, a, c - 1, e) + Ac);
      const T BB = T(0.5) * (Q(6, a - 1, c, e) + Bc);
      out[2] = uu * BB - vv * AA;
    }
    const int pe = G::NP * G::NP * (e + 1);
    for (int comp = 0; comp < 3; ++comp)
      store_nt(Ed + comp * G::NPE + pe, out[comp]);
    pv0 = v0;
    pv1 = v1;
    pw0 = w0;
    pw1 = w1;
    pw2 = w2;
    pu0 = u0;
    pu1 = u1;
    pB = Bc;
    pA = Ac;
  }
}

template <typename T, int DIM, int BS>
__device__ __forceinline__ void
efield_item(const FusedSlot<T> & S, long r0)
{
  using G = Geo<DIM, BS>;
  const int o = __builtin_amdgcn_readfirstlane(int(r0) / G::EFP);
  const int r = int(r0) - o * G::EFP;
  if constexpr (DIM == 3)
  {
    // one (a, c) column of edge points per lane
    if (r >= G::NP * G::NP)
      return;
    efield_column<T, BS>(S, o, r % G::NP - 1, r / G::NP - 1);
    return;
  }
  if (r >= G::NPE)
    return;
  constexpr int          np = G::NP;
  const int              a = r % np - 1, c = (r / np) % np - 1, e = DIM == 3 ? r / (np * np) - 1 : 0;
  const View<T, DIM, BS> V(S, o);
  const GlobalSrc<T, DIM, BS> src(V, S, o);
  T out[3];
  efield_point<T, DIM>(src, a, c, e, out);
  T * Ed = S.Ed + long(o) * G::NCE * G::NPE + r;
  for (int comp = 0; comp < G::NCE; ++comp)
    store_nt(Ed + comp * G::NPE, out[comp]);
}

template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256)
k_efield(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[6])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (efield_item<T, DIM, BS>(S, r0)));
  }
}


// ---------------------------------------------------------------------------------------------
// k_cell: per cell of the 1-ghost tile: slopes, predictor, face source, face-B slopes.  HK:
// HANCOCK (half-step predictor), else RK2_SSP (Q2 = the cell-centred state): a kernel template
// parameter, so that the HANCOCK kernel carries no trace of the other integrator.  VISC: the
// predictor's viscous term is compiled in (FP64 kernels of one block size: viscosity runs per
// context, never in a batch of mixed block sizes, and the FP32 form refuses it).
// ---------------------------------------------------------------------------------------------
// The stencil of one cell, each value read once: qc the 11 values of the cell (5 hydro, the low
// and the high face of each component), qn[d][0] / qn[d][1] at the neighbours +e_d / -e_d the 5
// hydro values and the low and high faces of the components across d (the only ones the slopes
// read), e the edge E values the face source reads.
template <typename T>
struct CellStencil
{
  T qc[11], qn[3][2][11];
  // Ex, Ey, Ez at (i, j, k); Ez at j + 1, Ey at k + 1, Ez at i + 1, Ex at k + 1, Ey at i + 1,
  // Ex at j + 1 (2D: e[2], e[3], e[5] only)
  T e[9];
};

// value v (0-4 hydro, 5-7 low and 8-10 high faces of Bx, By, Bz) of cell (a, c, e)
template <typename T, int DIM, class Src>
__device__ __forceinline__ T
cell_q(const Src & src, int v, int a, int c, int e)
{
  if (v < 5)
    return src.qp(v, a, c, e);
  const int fc = (v - 5) % 3;
  const int hi = v >= 8;
  if (DIM == 2 && fc == 2)
    return src.bf(2, a, c, 0);
  return src.bf(fc, a + (hi && fc == 0), c + (hi && fc == 1), e + (hi && fc == 2));
}

// the values of neighbour n = cell (i, j, k) + sg e_d that the slopes read
template <typename T, int DIM, class Src>
__device__ __forceinline__ void
cell_neighbour(const Src & src, int d, int sg, int i, int j, int k, T * qn)
{
  const int a = i + sg * (d == 0), c = j + sg * (d == 1), e = k + sg * (d == 2);
  for (int v = 0; v < 5; ++v)
    qn[v] = cell_q<T, DIM>(src, v, a, c, e);
  for (int mc = 0; mc < 3; ++mc)
    if (mc != d)
    {
      qn[5 + mc] = cell_q<T, DIM>(src, 5 + mc, a, c, e);
      qn[8 + mc] = cell_q<T, DIM>(src, 8 + mc, a, c, e);
    }
}

template <typename T, int DIM, class Src>
__device__ __forceinline__ void
cell_stencil(const Src & src, int i, int j, int k, CellStencil<T> & C)
{
  for (int v = 0; v < 11; ++v)
    C.qc[v] = cell_q<T, DIM>(src, v, i, j, k);
  for (int d = 0; d < DIM; ++d)
  {
    cell_neighbour<T, DIM>(src, d, 1, i, j, k, C.qn[d][0]);
    cell_neighbour<T, DIM>(src, d, -1, i, j, k, C.qn[d][1]);
  }
  if (DIM == 2)
  {
    C.e[2] = src.ed(0, i, j, 0);
    C.e[3] = src.ed(0, i, j + 1, 0);
    C.e[5] = src.ed(0, i + 1, j, 0);
  }
  else
  {
    C.e[0] = src.ed(0, i, j, k);
    C.e[1] = src.ed(1, i, j, k);
    C.e[2] = src.ed(2, i, j, k);
    C.e[3] = src.ed(2, i, j + 1, k);
    C.e[4] = src.ed(1, i, j, k + 1);
    C.e[5] = src.ed(2, i + 1, j, k);
    C.e[6] = src.ed(0, i, j, k + 1);
    C.e[7] = src.ed(1, i + 1, j, k);
    C.e[8] = src.ed(0, i, j + 1, k);
  }
}

// The stencil of cell (i, j, k + 1) from that of (i, j, k) (3D): the cell becomes the -z
// neighbour, the +z neighbour the cell (its low Bz face is the old high one), E at k + 1 of the
// old cell E at the new one; the rest is read.
template <typename T, class Src>
__device__ __forceinline__ void
cell_stencil_up(const Src & src, int i, int j, int k, CellStencil<T> & C)
{
  constexpr int keep[9] = { 0, 1, 2, 3, 4, 5, 6, 8, 9 };
  for (int n = 0; n < 9; ++n)
  {
    C.qn[2][1][keep[n]] = C.qc[keep[n]];
    C.qc[keep[n]] = C.qn[2][0][keep[n]];
  }
  C.qc[7] = C.qc[10];
  const int kn = k + 1;
  C.qc[10] = cell_q<T, 3>(src, 10, i, j, kn);
  cell_neighbour<T, 3>(src, 2, 1, i, j, kn, C.qn[2][0]);
  for (int d = 0; d < 2; ++d)
  {
    cell_neighbour<T, 3>(src, d, 1, i, j, kn, C.qn[d][0]);
    cell_neighbour<T, 3>(src, d, -1, i, j, kn, C.qn[d][1]);
  }
  C.e[0] = C.e[6];
  C.e[1] = C.e[4];
  C.e[2] = src.ed(2, i, j, kn);
  C.e[3] = src.ed(2, i, j + 1, kn);
  C.e[4] = src.ed(1, i, j, kn + 1);
  C.e[5] = src.ed(2, i + 1, j, kn);
  C.e[6] = src.ed(0, i, j, kn + 1);
  C.e[7] = src.ed(1, i + 1, j, kn);
  C.e[8] = src.ed(0, i, j + 1, kn);
}

// slopes, predictor, face source and face-B slopes of cell (i, j, k) from its stencil, stored as
// record c1
template <typename T, int DIM, int BS, bool HK, bool VISC>
__device__ __forceinline__ void
cell_finish(const FusedSlot<T> & S, const View<T, DIM, BS> & V, const CellStencil<T> & CS, int o, int c1)
{
  using G = Geo<DIM, BS>;
    const DevMesh &      m = S.m;
    const T   st = S.gt.slope_type;
    const T(&qc)[11] = CS.qc;
    const T(&qn)[3][2][11] = CS.qn;
    // Q(v, cell + o e_d) for o in {0, +1, -1}, from the stencil
    auto QS = [&](int v, int d, int off) -> T { return off == 0 ? qc[v] : qn[d][off > 0 ? 0 : 1][v]; };
    T         sl[3][8];
    constexpr int nh = DIM == 3 ? 5 : 4;
    for (int d = 0; d < DIM; ++d)
    {
      for (int v = 0; v < 8; ++v)
        sl[d][v] = T(0);
      for (int v = 0; v < nh; ++v)
        sl[d][v] = limited_slope(QS(v, d, 0), QS(v, d, 1), QS(v, d, -1), st);
      for (int mc = 0; mc < 3; ++mc)
      {
        const int vL = 5 + mc, vR = 8 + mc;
        if (mc == d)
          sl[d][5 + mc] = QS(vR, d, 0) - QS(vL, d, 0);
        else
        {
          const T dc = (QS(vL, d, 0) + QS(vR, d, 0)) * T(0.5);
          const T dp = (QS(vL, d, 1) + QS(vR, d, 1)) * T(0.5);
          const T dm = (QS(vL, d, -1) + QS(vR, d, -1)) * T(0.5);
          sl[d][5 + mc] = limited_slope(dc, dp, dm, st);
        }
      }
      // every slope but the normal-B one along its own direction, which no consumer reads (the
      // flux replaces the normal B by the face value; the edge reconstructions use the normal
      // component only as the quirk of x-edges in 3D, dBy/dy): 3D skips dBx/dx and dBz/dz, 2D
      // dBx/dx and dBy/dy
    }
    // Hancock predictor (ComputeUpdatedPrimitiveVariablesFunctor.cpp:284-546)
    const double dxo = dxd<BS>(m, S.scaling, o);
    const T      dtdx = T(S.dt / dxo), dtdy = dtdx, dtdz = dtdx;
    const T      gamma = S.gt.gamma0;
    const T r = qc[0], p = qc[1];
    T       u = qc[2], v = qc[3], w = qc[4];
    const T A = (qc[5] + qc[8]) * T(0.5);
    const T B = (qc[6] + qc[9]) * T(0.5);
    const T C = (qc[7] + qc[10]) * T(0.5);
    T q2o[8];
    if (!HK)
    {
      // RK2_SSP: Q2 is the cell-centred state (ComputeUpdatedPrimitiveVariablesFunctor.cpp:396-407,
      // 531-542)
      q2o[0] = r;
      q2o[2] = u;
      q2o[3] = v;
      q2o[4] = w;
      q2o[1] = p;
      q2o[5] = A;
      q2o[6] = B;
      q2o[7] = C;
    }
    else
    {
      const T drx = sl[0][0], dpx = sl[0][1], dux = sl[0][2], dvx = sl[0][3], dwx = sl[0][4], dAx = sl[0][5],
              dBx = sl[0][6], dCx = sl[0][7];
      const T dry = sl[1][0], dpy = sl[1][1], duy = sl[1][2], dvy = sl[1][3], dwy = sl[1][4], dAy = sl[1][5],
              dBy = sl[1][6], dCy = sl[1][7];
      T sr0, su0, sv0, sw0, sp0, sA0, sB0, sC0;
      if (DIM == 2)
      {
        sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy;
        su0 = (-u * dux - dpx / r - B * dBx / r - C * dCx / r) * dtdx + (-v * duy + B * dAy / r) * dtdy;
        sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - dpy / r - A * dAy / r - C * dCy / r) * dtdy;
        sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy;
        sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy;
        sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy;
        sB0 = (-u * dBx - B * dux + v * dAx + A * dvx) * dtdx;
        sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (-v * dCy - C * dvy + w * dBy + B * dwy) * dtdy;
      }
      else
      {
        const T drz = sl[2][0], dpz = sl[2][1], duz = sl[2][2], dvz = sl[2][3], dwz = sl[2][4], dAz = sl[2][5],
                dBz = sl[2][6], dCz = sl[2][7];
        sr0 = (-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy + (-w * drz - dwz * r) * dtdz;
        su0 = (-u * dux - (dpx + B * dBx + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy +
              (-w * duz + C * dAz / r) * dtdz;
        sv0 = (-u * dvx + A * dBx / r) * dtdx + (-v * dvy - (dpy + A * dAy + C * dCy) / r) * dtdy +
              (-w * dvz + C * dBz / r) * dtdz;
        sw0 = (-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy +
              (-w * dwz - (dpz + A * dAz + B * dBz) / r) * dtdz;
        sp0 = (-u * dpx - dux * gamma * p) * dtdx + (-v * dpy - dvy * gamma * p) * dtdy + (-w * dpz - dwz * gamma * p) * dtdz;
        sA0 = (u * dBy + B * duy - v * dAy - A * dvy) * dtdy + (u * dCz + C * duz - w * dAz - A * dwz) * dtdz;
        sB0 = (v * dAx + A * dvx - u * dBx - B * dux) * dtdx + (v * dCz + C * dvz - w * dBz - B * dwz) * dtdz;
        sC0 = (w * dAx + A * dwx - u * dCx - C * dux) * dtdx + (w * dBy + B * dwy - v * dCy - C * dvy) * dtdy;
      }
      // gravity, half time step, after the source terms
      // (ComputeUpdatedPrimitiveVariablesFunctor.cpp:372-377, 505-511)
      if (S.gt.grav)
      {
        const T dtT = T(S.dt);
        u += S.gt.g[0] * T(0.5) * dtT;
        v += S.gt.g[1] * T(0.5) * dtT;
        if (DIM == 3)
          w += S.gt.g[2] * T(0.5) * dtT;
      }
      // viscous predictor, after the gravity kick (ComputeUpdatedPrimitiveVariablesFunctor.cpp:
      // 178-272, called at :380-383, 514-517): per velocity component, per axis,
      // 1/2 nu (second difference) / dx * dt / dx with nu = mu / rho
      if (VISC && S.visc_mu >= 0.0)
      {
        const T nu = T(S.visc_mu) / r;
        T *     vel[3] = { &u, &v, &w };
        for (int c = 0; c < DIM; ++c)
          for (int a = 0; a < DIM; ++a)
          {
            const T d2 = QS(2 + c, a, 1) + QS(2 + c, a, -1) - T(2) * QS(2 + c, a, 0);
            *vel[c] += T(0.5) * nu * d2 / T(dxo) * dtdx;
          }
      }
      q2o[0] = r + T(0.5) * sr0;
      q2o[2] = u + T(0.5) * su0;
      q2o[3] = v + T(0.5) * sv0;
      q2o[4] = w + T(0.5) * sw0;
      q2o[1] = p + T(0.5) * sp0;
      q2o[5] = A + T(0.5) * sA0;
      q2o[6] = B + T(0.5) * sB0;
      q2o[7] = C + T(0.5) * sC0;
    }
    // edge electric field v x B at t_n (k_efield, or the front tile)
    const T(&E)[9] = CS.e;
    T sfo[3];
    if (DIM == 2)
    {
      const T e00 = E[2];
      sfo[0] = (E[3] - e00) * T(0.5) * dtdy;
      sfo[1] = -(E[5] - e00) * T(0.5) * dtdx;
    }
    else
    {
      const T ex = E[0], ey = E[1], ez = E[2];
      sfo[0] = (E[3] - ez) * T(0.5) * dtdy - (E[4] - ey) * T(0.5) * dtdz;
      sfo[1] = -(E[5] - ez) * T(0.5) * dtdx + (E[6] - ex) * T(0.5) * dtdz;
      sfo[2] = (E[7] - ey) * T(0.5) * dtdx - (E[8] - ex) * T(0.5) * dtdy;
    }
    // transverse slopes of the face field at this cell's low faces (dBx/dy, dBx/dz, dBy/dx,
    // dBy/dz, dBz/dx, dBz/dy; 2D: dBx/dy, dBy/dx)
    T fso[G::NFS];
    if (DIM == 2)
    {
      fso[0] = limited_slope(QS(5, 1, 0), QS(5, 1, 1), QS(5, 1, -1), st);
      fso[1] = limited_slope(QS(6, 0, 0), QS(6, 0, 1), QS(6, 0, -1), st);
    }
    else
    {
      fso[0] = limited_slope(QS(5, 1, 0), QS(5, 1, 1), QS(5, 1, -1), st);
      fso[1] = limited_slope(QS(5, 2, 0), QS(5, 2, 1), QS(5, 2, -1), st);
      fso[2] = limited_slope(QS(6, 0, 0), QS(6, 0, 1), QS(6, 0, -1), st);
      fso[3] = limited_slope(QS(6, 2, 0), QS(6, 2, 1), QS(6, 2, -1), st);
      fso[4] = limited_slope(QS(7, 0, 0), QS(7, 0, 1), QS(7, 0, -1), st);
      fso[5] = limited_slope(QS(7, 1, 0), QS(7, 1, 1), QS(7, 1, -1), st);
    }
    // every result is stored after the last load: buffer stores are not reordered across
    // buffer loads, so stores issued early would split the loads into dependent phases
    for (int d = 0; d < DIM; ++d)
      for (int v = 0; v < 8; ++v)
        if (!(v == kA + d && (DIM == 2 || d != 1)))
          bstore<T>(V.r_sl, (d * 8 + v) * G::G1 + c1, sl[d][v]);
    for (int v = 0; v < 8; ++v)
      bstore<T>(V.r_q2, v * G::G1 + c1, q2o[v]);
    for (int d = 0; d < DIM; ++d)
      bstore<T>(V.r_sf, d * G::G1 + c1, sfo[d]);
    for (int n2 = 0; n2 < G::NFS; ++n2)
      bstore<T>(V.r_fs, n2 * G::G1 + c1, fso[n2]);
}

template <typename T, int DIM, int BS, bool HK, bool VISC, class Src>
__device__ __forceinline__ void
cell_core(const FusedSlot<T> & S, const View<T, DIM, BS> & V, const Src & src, int o, int i, int j, int k, int c1)
{
  CellStencil<T> C;
  cell_stencil<T, DIM>(src, i, j, k, C);
  cell_finish<T, DIM, BS, HK, VISC>(S, V, C, o, c1);
}

// k_cell's work per block: NCHK runs of CL cells along z per (i, j) column of the 1-ghost tile
// (CMARCH), else one cell per item; P1 items per block
template <typename T, int DIM, int BS>
struct CellG
{
  static constexpr int  N1 = BS + 2;
  static constexpr int  G1 = DIM == 3 ? N1 * N1 * N1 : N1 * N1;
  static constexpr int  NCHK = fused_cell_runs(DIM, BS, sizeof(T) == sizeof(double));
  static constexpr bool CMARCH = NCHK > 0;
  static constexpr int  CL = CMARCH ? N1 / NCHK : 1;
  static_assert(!CMARCH || CL * NCHK == N1, "k_cell runs tile the column");
  static constexpr int  P1 = CMARCH ? (N1 * N1 * NCHK + 63) / 64 * 64 : (G1 + 63) / 64 * 64;
};

// One item per cell of the 1-ghost tile, or (CellG::CMARCH, FP64 3D) per lane marching a run of
// CellG::CL cells of a column along z: the stencil of each next cell from the last one's
// (cell_stencil_up), 53 reads per cell after the first instead of 74, at 2 wavefronts per SIMD
// (the stencil of the next cell and the results of this one are live together).
template <typename T, int DIM, int BS, bool HK, bool VISC>
__device__ __forceinline__ void
cell_item(const FusedSlot<T> & S, long r0)
{
  using G = Geo<DIM, BS>;
  using CG = CellG<T, DIM, BS>;
  constexpr int n = G::N1;
  const int     o = __builtin_amdgcn_readfirstlane(int(r0) / CG::P1);
  const int     ci = int(r0) - o * CG::P1;
  const View<T, DIM, BS>      V(S, o);
  const GlobalSrc<T, DIM, BS> src(V, S, o);
  if constexpr (DIM == 3 && CG::CMARCH)
  {
    // lane position (i, j, run) over N1 x N1 x NCHK: a column of CL cells from k0
    if (ci >= n * n * CG::NCHK)
      return;
    const int i = ci % n - 1, j = (ci / n) % n - 1, k0 = (ci / (n * n)) * CG::CL - 1;
    CellStencil<T> C;
    cell_stencil<T, DIM>(src, i, j, k0, C);
    for (int t = 0; t < CG::CL; ++t)
    {
      const int k = k0 + t;
      cell_finish<T, DIM, BS, HK, VISC>(S, V, C, o, View<T, DIM, BS>::c1(i, j, k));
      if (t + 1 < CG::CL)
        cell_stencil_up<T>(src, i, j, k, C);
    }
  }
  else
  {
    if (ci >= G::G1)
      return;
    const int i = ci % n - 1, j = (ci / n) % n - 1, k = DIM == 3 ? ci / (n * n) - 1 : 0;
    cell_core<T, DIM, BS, HK, VISC>(S, V, src, o, i, j, k, ci);
  }
}

// VISC: the viscous predictor term compiled in (only the per-mesh FP64 form runs viscosity; its
// registers would otherwise be held by every cell kernel)
template <typename T, int DIM, int BSEL, bool HK, bool VISC>
__global__ void __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu((DIM == 3 && sizeof(T) == sizeof(double) && (BSEL == 0 || fused_cell_runs(DIM, BSEL, true) > 0)) ? 2 : 3)))
k_cell(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[1])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (cell_item<T, DIM, BS, HK, (VISC && BSEL != 0 && sizeof(T) == sizeof(double))>(S, r0)));
  }
}


// ---------------------------------------------------------------------------------------------
// k_front: k_prim, k_efield and k_cell for one sub-tile of a block in one workgroup, with the
// primitive state, the ghosted face field and the edge electric field in LDS.  A sub-tile is
// 8 x 8 x TZ cells (3D: TZ = 4, or 2 in 8^3 blocks, so that two workgroups fit the LDS of a CU;
// 2D: 8 x 8).  Along each axis the sub-tile writes the records of the cells [lo, hi] (its cells;
// the first sub-tile of a block extends to -1 and the last to b, so the ghost-1 layer is covered
// once), which read the primitive state and the face field of the cells [lo-1, hi+1] and E at
// the edge points [lo, hi+1].  The face field of the cells [clo, chi] (its cells; the first
// extends to -2 and the last to b+1) also goes to Bg, which the flux and edge stages read.  The arithmetic is
// that of the three kernels, so the records are the same bit for bit; the primitive state and
// E of the tile's halo are computed by each tile that reads them.  The viscous predictor term
// reads Qp from HBM and keeps the three-kernel path.
// ---------------------------------------------------------------------------------------------
template <int DIM, int BS>
struct FrontG
{
  static constexpr int NB1 = BS / kTS;
  static constexpr int TZ = fused_front_depth(DIM, BS);             // sub-tile depth
  static constexpr int NSZ = DIM == 3 ? BS / TZ : 1;               // sub-tiles along z per block
  static constexpr int NSUB = NB1 * NB1 * NSZ;                     // sub-tiles per block
  static constexpr int R = NB1 == 1 ? kTS + 2 : kTS + 1;          // most record cells along x, y
  static constexpr int RZ = DIM == 3 ? TZ + 1 : 1;                 // ... along z
  static constexpr int P = R + 2;                                  // primitive cells along x, y
  static constexpr int PZ = DIM == 3 ? RZ + 2 : 1;                 // ... along z
  static constexpr int PE = R + 1;                                 // edge points along x, y
  static constexpr int PEZ = DIM == 3 ? RZ + 1 : 1;                // ... along z
  static constexpr int NQ = P * P * PZ;                            // primitive cells
  // faces of each component: one more along its own axis (2D: the cell-centred z field)
  static constexpr int NF0 = (P + 1) * P * PZ;
  static constexpr int NF1 = P * (P + 1) * PZ;
  static constexpr int NF2 = P * P * (DIM == 3 ? PZ + 1 : 1);
  static constexpr int OF1 = NF0, OF2 = NF0 + NF1, NFT = NF0 + NF1 + NF2;
  static constexpr int NPT = PE * PE * PEZ;                        // edge points
  static constexpr int NCE = DIM == 3 ? 3 : 1;
  static_assert(DIM == 2 || BS % TZ == 0, "sub-tile depth");
};

// LDS of one k_front workgroup in doubles: the face field in FP64, then the primitive state and E
// in the bulk type T
template <typename T, int DIM, int BS>
constexpr int
front_lds_doubles()
{
  using F = FrontG<DIM, BS>;
  return F::NFT + int(((5 * F::NQ + F::NCE * F::NPT) * sizeof(T) + 7) / 8);
}

// the largest of the block sizes a kernel serves (BSEL 0: 8, 16 and 32)
template <typename T, int DIM, int BSEL>
constexpr int
front_lds_max()
{
  if constexpr (BSEL != 0)
    return front_lds_doubles<T, DIM, BSEL>();
  else
  {
    constexpr int a = front_lds_doubles<T, DIM, 8>(), b = front_lds_doubles<T, DIM, 16>(),
                  c = front_lds_doubles<T, DIM, 32>();
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
  }
}

template <int DIM>
constexpr int
front_threads()
{
  return DIM == 3 ? 512 : 256;
}

// the LDS tile of k_front: origin (x0, y0, z0) = (lo-1) per axis; each grid is laid out with the
// extents of the largest tile of the block size, so that a grid slot is a compile-time function
// of its position
template <typename T, int DIM, int BS>
struct LdsSrc
{
  using F = FrontG<DIM, BS>;
  double * lf; // [NFT]      face field: component c over [lo-1, hi+2] along c, [lo-1, hi+1] across
  T *      lq; // [5][NQ]    primitive state of the cells [lo-1, hi+1]
  T *      le; // [NCE][NPT] E at the edge points [lo, hi+1]
  int      x0, y0, z0;
  __device__ int cq(int a, int c, int e) const
  {
    return (a - x0) + F::P * ((c - y0) + (DIM == 3 ? F::P * (e - z0) : 0));
  }
  __device__ int cf(int fc, int a, int c, int e) const
  {
    const int e0 = F::P + (fc == 0), e1 = F::P + (fc == 1);
    const int base = fc == 0 ? 0 : (fc == 1 ? F::OF1 : F::OF2);
    return base + (a - x0) + e0 * ((c - y0) + (DIM == 3 ? e1 * (e - z0) : 0));
  }
  __device__ int ce(int comp, int a, int c, int e) const
  {
    return comp * F::NPT + (a - x0 - 1) + F::PE * ((c - y0 - 1) + (DIM == 3 ? F::PE * (e - z0 - 1) : 0));
  }
  __device__ T qp(int v, int a, int c, int e) const { return lq[v * F::NQ + cq(a, c, e)]; }
  __device__ T bf(int fc, int a, int c, int e) const { return T(lf[cf(fc, a, c, e)]); }
  __device__ T ed(int comp, int a, int c, int e) const { return le[ce(comp, a, c, e)]; }
};

template <typename T, int DIM, int BS, bool HK>
__device__ __forceinline__ void
front_item(const FusedSlot<T> & S, long r0, double * lds)
{
  using G = Geo<DIM, BS>;
  using F = FrontG<DIM, BS>;
  using IX = IdxC<DIM, BS>;
  constexpr IX    I{};
  constexpr int   NTHR = front_threads<DIM>();
  constexpr int   nb1 = G::NB1;
  constexpr int   P = F::P;
  const DevMesh & m = S.m;
  const int       o = int(r0) / F::NSUB;
  const int       tl = int(r0) - o * F::NSUB;
  // segment along each axis: 8-cell segments along x, y; TZ-cell segments along z
  const int       seg[3] = { tl % nb1, (tl / nb1) % nb1, DIM == 3 ? tl / (nb1 * nb1) : 0 };
  int             lo[3], hi[3], clo[3], chi[3];
  for (int d = 0; d < 3; ++d)
  {
    const int s = seg[d];
    const int w = d == 2 ? F::TZ : kTS, ns = d == 2 ? F::NSZ : nb1;
    lo[d] = s == 0 ? -1 : w * s;
    hi[d] = s == ns - 1 ? BS : w * s + w - 1;
    clo[d] = s == 0 ? -2 : w * s;
    chi[d] = s == ns - 1 ? BS + 1 : w * s + w - 1;
  }
  if (DIM == 2)
    lo[2] = hi[2] = clo[2] = chi[2] = 0;
  T *                lq = reinterpret_cast<T *>(lds + F::NFT);
  LdsSrc<T, DIM, BS> src{ lds, lq, lq + 5 * F::NQ, lo[0] - 1, lo[1] - 1, DIM == 3 ? lo[2] - 1 : 0 };
  const double * __restrict__ Bo = S.B + long(o) * IX::fper();
  const double * __restrict__ Uo = S.U + long(o) * 5 * IX::cells();
  T * Go = S.Bg + long(o) * IX::gfper();
  // phase 1a: the face field (component c over [lo-1, hi+2] along c, [lo-1, hi+1] across; the 2D
  // cell-centred z field over [lo-1, hi+1]), FP64, and the faces of the cells [clo, chi] to Bg.
  // Slot t of the face grid is face t; the faces inside the block are loaded first, all at once
  // (unconditional loads, clamped to face 0 elsewhere), the ghost faces after.
  {
    constexpr int NF = F::NFT;
    constexpr int FMAX = (NF + NTHR - 1) / NTHR;
    struct Face
    {
      int  c, a0, a1, a2;
      bool valid, inner, own;
    };
    auto at = [&](int t) {
      Face       f;
      f.c = t < F::OF1 ? 0 : (t < F::OF2 ? 1 : 2);
      const int  r = t - (f.c == 0 ? 0 : (f.c == 1 ? F::OF1 : F::OF2));
      const int  e0 = P + (f.c == 0), e1 = P + (f.c == 1);
      const int  rel[3] = { r % e0, (r / e0) % e1, r / (e0 * e1) };
      const bool cellc = DIM == 2 && f.c == 2;
      int        a[3] = { 0, 0, 0 };
      f.valid = t < NF;
      f.inner = true;
      f.own = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d)
      {
        const int along = d == f.c && !cellc;
        a[d] = lo[d] - 1 + rel[d];
        f.valid = f.valid && a[d] <= hi[d] + 1 + along;
        f.inner = f.inner && a[d] >= 0 && a[d] <= BS - 1 + along;
        f.own = f.own && a[d] >= clo[d] && a[d] <= chi[d] + (along && chi[d] == BS + 1);
      }
      f.a0 = a[0];
      f.a1 = a[1];
      f.a2 = a[2];
      f.inner = f.inner && f.valid;
      return f;
    };
    double x[FMAX];
#pragma unroll
    for (int r = 0; r < FMAX; ++r)
    {
      const Face f = at(threadIdx.x + r * NTHR);
      x[r] = Bo[f.inner ? IX::face(f.c, f.a0, f.a1, f.a2) : 0];
    }
#pragma unroll
    for (int r = 0; r < FMAX; ++r)
    {
      const int  t = threadIdx.x + r * NTHR;
      const Face f = at(t);
      if (f.valid && !f.inner)
      {
        // the component as a constant, so that the lookup's arrays indexed by it stay in registers
        const int a[3] = { f.a0, f.a1, f.a2 };
        if (f.c == 0)
          x[r] = face_value_ghosted(m, I, S.B, o, 0, a);
        else if (f.c == 1)
          x[r] = face_value_ghosted(m, I, S.B, o, 1, a);
        else
          x[r] = face_value_ghosted(m, I, S.B, o, 2, a);
      }
      if (f.valid)
      {
        lds[t] = x[r];
        if (f.own)
          store_nt(Go + IX::gface(f.c, f.a0, f.a1, f.a2), T(x[r]));
      }
    }
  }
  __syncthreads();
  // phase 1b: the primitive state of the cells [lo-1, hi+1] (slot t of the grid is cell t): the
  // cells inside the block first (U loaded all at once, the cell-centred field from the face
  // grid), then the ghost cells
  {
    constexpr int CMAX = (F::NQ + NTHR - 1) / NTHR;
    struct Cell
    {
      int  a0, a1, a2;
      bool valid, inside;
    };
    auto at = [&](int t) {
      Cell      cl;
      const int rel[3] = { t % P, (t / P) % P, DIM == 3 ? t / (P * P) : 0 };
      int       a[3] = { 0, 0, 0 };
      cl.valid = t < F::NQ;
      cl.inside = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d)
      {
        a[d] = lo[d] - 1 + rel[d];
        cl.valid = cl.valid && a[d] <= hi[d] + 1;
        cl.inside = cl.inside && a[d] >= 0 && a[d] < BS;
      }
      cl.a0 = a[0];
      cl.a1 = a[1];
      cl.a2 = a[2];
      cl.inside = cl.inside && cl.valid;
      return cl;
    };
    auto gf = [&](int c, int a0, int a1, int a2) { return lds[src.cf(c, a0, a1, a2)]; };
    auto finish = [&](int t, const double * u) {
      double       q[8];
      const double eint = to_primitive(u, q, S.gd);
      if (S.abort_eint && eint < 0)
        atomicOr(S.red + 1, 1ull);
      for (int v = 0; v < 5; ++v)
        lq[v * F::NQ + t] = T(q[v]);
    };
    double u[CMAX][5];
#pragma unroll
    for (int r = 0; r < CMAX; ++r)
    {
      const Cell cl = at(threadIdx.x + r * NTHR);
      const int  cidx = cl.inside ? IX::cell(cl.a0, cl.a1, cl.a2) : 0;
#pragma unroll
      for (int v = 0; v < 5; ++v)
        u[r][v] = Uo[v * IX::cells() + cidx];
    }
#pragma unroll
    for (int r = 0; r < CMAX; ++r)
    {
      const int  t = threadIdx.x + r * NTHR;
      const Cell cl = at(t);
      if (cl.inside)
      {
        const int i = cl.a0, j = cl.a1, k = cl.a2;
        double    w[8] = { u[r][0], u[r][1], u[r][2], u[r][3], u[r][4] };
        if (DIM == 2)
        {
          w[kA] = 0.5 * (gf(0, i, j, 0) + gf(0, i + 1, j, 0));
          w[kB] = 0.5 * (gf(1, i, j, 0) + gf(1, i, j + 1, 0));
          w[kC] = 0.5 * (gf(2, i, j, 0) + gf(2, i, j, 0));
        }
        else
        {
          w[kA] = 0.5 * (gf(0, i, j, k) + gf(0, i + 1, j, k));
          w[kB] = 0.5 * (gf(1, i, j, k) + gf(1, i, j + 1, k));
          w[kC] = 0.5 * (gf(2, i, j, k) + gf(2, i, j, k + 1));
        }
        finish(t, w);
      }
    }
#pragma unroll 1
    for (int r = 0; r < CMAX; ++r)
    {
      const int  t = threadIdx.x + r * NTHR;
      const Cell cl = at(t);
      if (cl.valid && !cl.inside)
      {
        double w[8];
        ghost_state<T, DIM, BS>(S, o, cl.a0, cl.a1, cl.a2, w, gf);
        finish(t, w);
      }
    }
  }
  __syncthreads();
  // phase 2: E at the edge points [lo, hi+1]
  {
    const int ex = hi[0] - lo[0] + 2, ey = hi[1] - lo[1] + 2, ez = DIM == 3 ? hi[2] - lo[2] + 2 : 1;
#pragma unroll 1
    for (int t = threadIdx.x; t < ex * ey * ez; t += NTHR)
    {
      const int a = lo[0] + t % ex, c = lo[1] + (t / ex) % ey, e = DIM == 3 ? lo[2] + t / (ex * ey) : 0;
      T         out[3];
      efield_point<T, DIM>(src, a, c, e, out);
      for (int comp = 0; comp < F::NCE; ++comp)
        src.le[src.ce(comp, a, c, e)] = out[comp];
    }
  }
  __syncthreads();
  // phase 3: the records of the cells [lo, hi]
  {
    const View<T, DIM, BS> V(S, o);
    const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = DIM == 3 ? hi[2] - lo[2] + 1 : 1;
    static_assert(F::R * F::R * F::RZ <= NTHR, "one record cell per thread");
    const int t = threadIdx.x;
    if (t < ex * ey * ez)
    {
      const int i = lo[0] + t % ex, j = lo[1] + (t / ex) % ey, k = DIM == 3 ? lo[2] + t / (ex * ey) : 0;
      cell_core<T, DIM, BS, HK, false>(S, V, src, o, i, j, k, View<T, DIM, BS>::c1(i, j, k));
    }
  }
}

// k_front: one workgroup per 8^dim tile of the flux-tile work list.  No grid-stride loop: with
// the tile bounds known at compile time (one tile per block) the compiler would keep every
// thread's tile indexing live across it, beyond the register budget of 1024-thread workgroups.
template <typename T, int DIM, int BSEL, bool HK>
__global__ void __launch_bounds__((front_threads<DIM>())) __attribute__((amdgpu_waves_per_eu(4)))
k_front(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  __shared__ double lds[front_lds_max<T, DIM, BSEL>()];
  const long        g = xcd_group();
  if (g >= total)
    return;
  const int            s = slot_of(off, nslots, g);
  const FusedSlot<T> & S = slots[s];
  const long           r0 = g - off[s];
  if (r0 >= S.cnt[kStageFront])
    return;
  MHDCT_BS_DISPATCH(S.m.b, (front_item<T, DIM, BS, HK>(S, r0, lds)));
}


// HLLD fluxes (times dt/dx) through the faces along DIR (ComputeHydroFluxesAndStoreFunctor.cpp:
// 218-503), in the reference variable order.
//
// The reconstructed state of cell c (index in the 1-ghost record grid) on its face on side SG
// along DIR (+1: the high face, a left state; -1: the low face, a right state), floored; the
// normal B component is left to the face value (face_hlld).
template <typename T, int DIM, int BS, int DIR, int SG>
__device__ __forceinline__ void
face_side(const View<T, DIM, BS> & V, int c, T * q)
{
  for (int v = 0; v < 8; ++v)
    if (v != kA + DIR)
      q[v] = V.q2(v, c) + T(0.5 * SG) * V.sl(DIR, v, c);
  const Gas<T> & g = V.S.gt;
  q[kD] = fmax(g.smallr, q[kD]);
  q[kP] = fmax(g.smallp * q[kD], q[kP]);
}

// HLLD flux along DIR between the states qL and qR with the face's normal field bn, times dt/dx
// (qL and qR are modified)
template <typename T, int DIR>
__device__ __forceinline__ void
face_hlld(T * qL, T * qR, T bn, T dtdS, const Gas<T> & g, T * f)
{
  qR[kA + DIR] = bn;
  qL[kA + DIR] = bn;
  if (DIR != 0)
  {
    constexpr int vt = kU + DIR, bt = kA + DIR;
    T             x;
    x = qL[kU]; qL[kU] = qL[vt]; qL[vt] = x;
    x = qR[kU]; qR[kU] = qR[vt]; qR[vt] = x;
    x = qL[kA]; qL[kA] = qL[bt]; qL[bt] = x;
    x = qR[kA]; qR[kA] = qR[bt]; qR[bt] = x;
  }
  hlld_fast(qL, qR, f, g);
  if (DIR != 0)
  {
    constexpr int vt = kU + DIR, bt = kA + DIR;
    T             x;
    x = f[kU]; f[kU] = f[vt]; f[vt] = x;
    x = f[kA]; f[kA] = f[bt]; f[bt] = x;
  }
  for (int v = 0; v < 8; ++v)
    f[v] = f[v] * dtdS;
}

// HLLD flux through the low DIR face of cell (i, j, k), times dt/dx
template <typename T, int DIM, int BS, int DIR>
__device__ __forceinline__ void
face_flux(const View<T, DIM, BS> & V, int i, int j, int k, T dtdS, T * f)
{
  using VW = View<T, DIM, BS>;
  const int cR = VW::c1(i, j, k);
  const int cL = VW::c1(i - (DIR == 0), j - (DIR == 1), k - (DIR == 2));
  T         qL[8], qR[8];
  face_side<T, DIM, BS, DIR, 1>(V, cL, qL);
  face_side<T, DIM, BS, DIR, -1>(V, cR, qR);
  const T bn = V.bf(DIM == 2 && DIR == 2 ? 2 : DIR, i, j, k) + V.sf(DIR, cR);
  face_hlld<T, DIR>(qL, qR, bn, dtdS, V.S.gt, f);
}

// the six stored flux components: rho, E, mx, my, mz, Bz (2D, cell-centred Bz)
__device__ constexpr int kFluxVar[6] = { kD, kP, kU, kV, kW, kC };

template <int DIM, int BS>
__device__ __forceinline__ long
bf_index(int o, int dir, int side, int v, int tpos)
{
  constexpr int nt = Geo<DIM, BS>::NT;
  return long((((o * DIM + dir) * 2 + side) * 6 + v)) * nt + tpos;
}

// ---------------------------------------------------------------------------------------------
// k_border_flux: HLLD on every block-border face (both sides, every direction)
// ---------------------------------------------------------------------------------------------
template <typename T, int DIM, int BS, int DIR>
__device__ __forceinline__ void
border_face(const FusedSlot<T> & S, const View<T, DIM, BS> & V, int o, int side, int tp)
{
  constexpr int t0 = DIR == 0 ? 1 : 0;
  constexpr int t1 = DIR == 2 ? 1 : 2;
  int           c[3] = { 0, 0, 0 };
  c[t0] = tp % BS;
  if (DIM == 3)
    c[t1] = tp / BS;
  c[DIR] = side ? BS : 0;
  T       f[8];
  const T dtdS = T(S.dt / dxd<BS>(S.m, S.scaling, o));
  face_flux<T, DIM, BS, DIR>(V, c[0], c[1], c[2], dtdS, f);
  for (int v = 0; v < (DIM == 2 ? 6 : 5); ++v)
    S.BF[bf_index<DIM, BS>(o, DIR, side, v, tp)] = f[kFluxVar[v]];
}

template <typename T, int DIM, int BS>
__device__ __forceinline__ void
border_flux_item(const FusedSlot<T> & S, long r0)
{

  using G = Geo<DIM, BS>;
    constexpr int        per = G::NTP;
    const int q = __builtin_amdgcn_readfirstlane(int(r0) / per);
    const int tp = int(r0) - q * per;
    if (tp >= G::NT)
      return;
    const int              code = S.bside[q];
    const int              o = code / (2 * DIM);
    const int              dir = (code >> 1) % DIM;
    const int              side = code & 1;
    const View<T, DIM, BS> V(S, o);
    if (dir == 0)
      border_face<T, DIM, BS, 0>(S, V, o, side, tp);
    else if (dir == 1)
      border_face<T, DIM, BS, 1>(S, V, o, side, tp);
    else if (DIM == 3)
      border_face<T, DIM, BS, 2>(S, V, o, side, tp);
}

template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256)
k_border_flux(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[2])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (border_flux_item<T, DIM, BS>(S, r0)));
  }
}


// ---------------------------------------------------------------------------------------------
// k_flux_tile: one workgroup per 8^dim tile; HLLD on the tile faces, one direction at a time
// through LDS; U2 = ((U + dX) + dY) + dZ accumulated in FP64 registers.  The workgroup has one
// thread per face of a direction (3D: 9*8*8 = 576), the first 8^dim of them also own a cell.
// ---------------------------------------------------------------------------------------------
// Flux tile of kTS x kTS x TZ cells (3D; 2D: kTS x kTS): the faces of each direction, one per
// thread, go to an LDS plane per direction; the cells of the tile are owned by the first threads.
template <int DIM, int TZ>
struct TileG
{
  static constexpr int TX = kTS, TY = kTS, TD = DIM == 3 ? TZ : 1;
  static constexpr int NCT = TX * TY * TD;
  static constexpr int NF0 = (TX + 1) * TY * TD;
  static constexpr int NF1 = TX * (TY + 1) * TD;
  static constexpr int NF2 = DIM == 3 ? TX * TY * (TD + 1) : 0;
  static constexpr int NFMAX = NF0 > NF1 ? (NF0 > NF2 ? NF0 : NF2) : (NF1 > NF2 ? NF1 : NF2);
  static constexpr int NTHR = DIM == 3 ? (NFMAX + 63) / 64 * 64 : 128;
  static constexpr int SPLIT = DIM == 3 ? kTS / TZ : 1; // tiles per 8^dim work item
};

template <int DIM, int TZ>
constexpr int
tile_threads()
{
  return TileG<DIM, TZ>::NTHR;
}

// The flux through the block-border face at cell c (c[DIR] = 0 or b, `side` 0 or 1) whose
// neighbour is finer: the sum of the fine faces covering it (their border fluxes), divided by
// 2^dim (as tile_faces has always formed it).
template <typename T, int DIM, int BS, int DIR>
__device__ __forceinline__ void
fine_face_flux(const FusedSlot<T> & S, int o, const int * c, int side, T * out)
{
  constexpr int NV = DIM == 2 ? 6 : 5;
  constexpr int t0 = DIR == 0 ? 1 : 0;
  constexpr int t1 = DIR == 2 ? 1 : 2;
  int           cc[3] = { c[0], c[1], c[2] };
  cc[DIR] = side == 0 ? -1 : BS;
  const Hit    h = neighbour(S.m, IdxC<DIM, BS>{}, o, cc);
  const double vr = double(1 << DIM);
  for (int vi = 0; vi < NV; ++vi)
  {
    double sum = 0;
    if (DIM == 2)
    {
      for (int a = 0; a < 2; ++a)
        sum += S.BF[bf_index<DIM, BS>(h.oct, DIR, 1 - side, vi, h.l[t0] + a)];
    }
    else
    {
      for (int q1 = 0; q1 < 2; ++q1)
        for (int q0 = 0; q0 < 2; ++q0)
          sum += S.BF[bf_index<DIM, BS>(h.oct, DIR, 1 - side, vi, (h.l[t0] + q0) + BS * (h.l[t1] + q1))];
    }
    out[vi] = T(sum / vr);
  }
}

template <typename T, int DIM, int BS, int TZ, int DIR>
__device__ __forceinline__ void
tile_faces(const FusedSlot<T> & S, const View<T, DIM, BS> & V, int o, const int org[3], T dtdS,
           T (*lf)[TileG<DIM, TZ>::NFMAX])
{
  using TG = TileG<DIM, TZ>;
  constexpr int NF = DIR == 0 ? TG::NF0 : (DIR == 1 ? TG::NF1 : TG::NF2);
  constexpr int NV = DIM == 2 ? 6 : 5;
  constexpr int e0 = DIR == 0 ? TG::TX + 1 : TG::TX;
  constexpr int e1 = DIR == 1 ? TG::TY + 1 : TG::TY;
  const DevMesh & m = S.m;
  for (int fl = threadIdx.x; fl < NF; fl += TG::NTHR)
  {
    const int a0 = fl % e0, a1 = (fl / e0) % e1, a2 = DIM == 3 ? fl / (e0 * e1) : 0;
    int       c[3] = { org[0] + a0, org[1] + a1, org[2] + a2 };
    T          out[NV];
    const bool border = c[DIR] == 0 || c[DIR] == BS;
    const int  side = c[DIR] == BS;
    if (border && m.conformal[long(o) * 2 * DIM + 2 * DIR + side] != 0)
      fine_face_flux<T, DIM, BS, DIR>(S, o, c, side, out);
    else
    {
      T f[8];
      face_flux<T, DIM, BS, DIR>(V, c[0], c[1], c[2], dtdS, f);
      for (int vi = 0; vi < NV; ++vi)
        out[vi] = f[kFluxVar[vi]];
    }
    for (int vi = 0; vi < NV; ++vi)
      lf[vi][fl] = out[vi];
  }
}

// one tile: r0 is the 8^dim work item within the slot, part selects the TZ-deep slab of it
template <typename T, int DIM, int BS, int TZ>
__device__ __forceinline__ void
flux_tile_item(const FusedSlot<T> & S, long r0, int part, T (*lf)[DIM == 3 ? 5 : 6][TileG<DIM, TZ>::NFMAX])
{
  using G = Geo<DIM, BS>;
  using IX = IdxC<DIM, BS>;
  using TG = TileG<DIM, TZ>;
  constexpr int NV = DIM == 2 ? 6 : 5;
  const DevMesh & m = S.m;
  const int              o = int(r0) / G::TPB;
  const View<T, DIM, BS> V(S, o);
  const int              tl = int(r0) - o * G::TPB;
  constexpr int          nb1 = G::NB1;
  const int org[3] = { (tl % nb1) * kTS, ((tl / nb1) % nb1) * kTS,
                       DIM == 3 ? (tl / (nb1 * nb1)) * kTS + part * TZ : 0 };
  const int     tid = threadIdx.x;
  const bool    owner = tid < TG::NCT;
  const int     li = tid % TG::TX, lj = (tid / TG::TX) % TG::TY, lk = DIM == 3 ? tid / (TG::TX * TG::TY) : 0;
  const int     ci = org[0] + li, cj = org[1] + lj, ck = org[2] + lk;
  constexpr int ncell = IX::cells();
  const int     cidx = IX::cell(ci, cj, ck);
  double        acc[NV];
  if (owner)
  {
    for (int v = 0; v < 5; ++v)
      acc[v] = S.U[(long(o) * 5 + v) * ncell + cidx];
    if (DIM == 2)
      acc[5] = S.B[long(o) * IX::fper() + IX::face(2, ci, cj, 0)];
  }
  const T dtdS = T(S.dt / dxd<BS>(m, S.scaling, o));
  // the faces of every direction first (one LDS plane per direction), then one barrier
  tile_faces<T, DIM, BS, TZ, 0>(S, V, o, org, dtdS, lf[0]);
  tile_faces<T, DIM, BS, TZ, 1>(S, V, o, org, dtdS, lf[1]);
  if (DIM == 3)
    tile_faces<T, DIM, BS, TZ, 2>(S, V, o, org, dtdS, lf[DIM - 1]);
  __syncthreads();
  if (owner)
  {
    for (int dir = 0; dir < DIM; ++dir)
    {
      int ext[3] = { TG::TX, TG::TY, TG::TD };
      ext[dir] += 1;
      int       a[3] = { li, lj, lk };
      const int fL = a[0] + ext[0] * (a[1] + ext[1] * a[2]);
      a[dir] += 1;
      const int fR = a[0] + ext[0] * (a[1] + ext[1] * a[2]);
      for (int vi = 0; vi < NV; ++vi)
      {
        // (0 + F_left) - F_right, as ReadFluxesAndConservativeUpdateFunctor accumulates it
        const double d = double(lf[dir][vi][fL]) - double(lf[dir][vi][fR]);
        acc[vi] = acc[vi] + d;
      }
    }
  }
  __syncthreads();
  if (owner)
  {
    for (int v = 0; v < 5; ++v)
      S.U2[(long(o) * 5 + v) * ncell + cidx] = acc[v];
    if (DIM == 2)
      S.B2[long(o) * IX::fper() + IX::face(2, ci, cj, 0)] = acc[5];
  }
}

// k_flux_tile: one workgroup per tile; `total` counts 8^dim work items, each split into
// TileG::SPLIT tiles of depth TZ
template <typename T, int DIM, int BSEL, int TZ>
__global__ void __launch_bounds__((tile_threads<DIM, TZ>()))
k_flux_tile(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  using TG = TileG<DIM, TZ>;
  constexpr int NV = DIM == 2 ? 6 : 5;
  __shared__ T  lf[DIM][NV][TG::NFMAX];
  for (long g = xcd_group(); g < total * TG::SPLIT; g += gridDim.x)
  {
    const long           t = g / TG::SPLIT;
    const int            part = int(g - t * TG::SPLIT);
    const int            s = slot_of(off, nslots, t);
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[3])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (flux_tile_item<T, DIM, BS, TZ>(S, r0, part, lf)));
  }
}

// ---------------------------------------------------------------------------------------------
// k_flux_march (3D): the flux stage of an 8^3 tile with each cell's record read once per
// direction instead of twice.  A workgroup of 4 wavefronts takes a tile:
//   * wavefront 0 (z-faces) and 1 (y-faces): a lane per (x, other) column of the tile marches
//     along the direction through the 10 cells from -1 to 8, forming from each cell's record its
//     right state (the face below it) and its left state (the face above it, carried to the next
//     cell); face a's flux with the carried one of face a - 1 gives cell a - 1's difference;
//   * wavefronts 2 and 3 (x-faces): rows of the 10 cells from -1 to 8 along x, 6 rows per pass;
//     each lane forms both states of its cell, takes the left state of its face from the lane
//     before it and the flux of the face above from the lane after it (cross-lane reads).
// The three differences of each cell go to LDS (FP64, 60 KB), and after one barrier every cell
// is updated as k_flux_tile updates it: ((U + d_x) + d_y) + d_z.  Block-border faces with a finer
// neighbour take the sum of its border fluxes (fine_face_flux), as in k_flux_tile.  Same fluxes,
// the same U2 bit for bit.
// ---------------------------------------------------------------------------------------------
template <typename T, int BS, int DIR>
__device__ __forceinline__ void
flux_march_dir(const FusedSlot<T> & S, const View<T, 3, BS> & V, int o, const int * org, int lane, T dtdS,
               double (*ld)[512])
{
  static_assert(DIR == 1 || DIR == 2, "flux_march_dir: y or z faces");
  using VW = View<T, 3, BS>;
  constexpr int OT = DIR == 1 ? 2 : 1; // the transverse direction besides x
  const int     li = lane & 7, le = lane >> 3;
  const bool    fine_lo = S.m.conformal[long(o) * 6 + 2 * DIR] != 0;
  const bool    fine_hi = S.m.conformal[long(o) * 6 + 2 * DIR + 1] != 0;
  int           c[3];
  c[0] = org[0] + li;
  c[OT] = org[OT] + le;
  T qLp[8], Fp[5];
  for (int t = 0; t < kTS + 2; ++t)
  {
    const int a = t - 1; // the tile-relative cell along DIR, -1 .. 8
    c[DIR] = org[DIR] + a;
    const int cc = VW::c1(c[0], c[1], c[2]);
    T         qR[8], qLn[8];
    face_side<T, 3, BS, DIR, -1>(V, cc, qR);
    if (t <= kTS)
      face_side<T, 3, BS, DIR, 1>(V, cc, qLn);
    if (t >= 1)
    {
      // the face below cell a (absolute c[DIR], uniform across the wavefront)
      T          out[5];
      const bool lo = c[DIR] == 0, hi = c[DIR] == BS;
      if ((lo && fine_lo) || (hi && fine_hi))
        fine_face_flux<T, 3, BS, DIR>(S, o, c, hi ? 1 : 0, out);
      else
      {
        T       f[8];
        const T bn = V.bf(DIR, c[0], c[1], c[2]) + V.sf(DIR, cc);
        face_hlld<T, DIR>(qLp, qR, bn, dtdS, S.gt, f);
        for (int vi = 0; vi < 5; ++vi)
          out[vi] = f[kFluxVar[vi]];
      }
      if (t >= 2)
      {
        // cell a - 1: (0 + F_low) - F_high, as ReadFluxesAndConservativeUpdateFunctor forms it
        const int cl = DIR == 1 ? li + kTS * ((a - 1) + kTS * le) : li + kTS * (le + kTS * (a - 1));
        for (int vi = 0; vi < 5; ++vi)
          ld[vi][cl] = double(Fp[vi]) - double(out[vi]);
      }
      for (int vi = 0; vi < 5; ++vi)
        Fp[vi] = out[vi];
    }
    if (t <= kTS)
      for (int v = 0; v < 8; ++v)
        qLp[v] = qLn[v];
  }
}

template <typename T, int BS>
__device__ __forceinline__ void
flux_rows_x(const FusedSlot<T> & S, const View<T, 3, BS> & V, int o, const int * org, int xw, int lane, T dtdS,
            double (*ld)[512])
{
  using VW = View<T, 3, BS>;
  constexpr int RPP = 6;                        // rows of 10 lanes per pass
  constexpr int NPASS = (kTS * kTS + RPP - 1) / RPP; // 11 passes over the 64 rows
  constexpr int P0 = (NPASS + 1) / 2;           // passes of the first x wavefront
  const bool    fine_lo = S.m.conformal[long(o) * 6] != 0;
  const bool    fine_hi = S.m.conformal[long(o) * 6 + 1] != 0;
  const int     rr = lane / (kTS + 2), p = lane - rr * (kTS + 2);
  const int     a = p - 1; // the tile-relative cell along x, -1 .. 8
  const int     pass0 = xw == 0 ? 0 : P0, pass1 = xw == 0 ? P0 : NPASS;
  const int     back = (lane > 0 ? lane - 1 : 0) * 4, next = (lane < 63 ? lane + 1 : 63) * 4;
  for (int pass = pass0; pass < pass1; ++pass)
  {
    const int  row = pass * RPP + rr;
    const bool valid = rr < RPP && row < kTS * kTS;
    const int  lj = row % kTS, lk = row / kTS;
    int        c[3] = { org[0] + a, org[1] + lj, org[2] + lk };
    T          qR[8], qLn[8];
    for (int v = 0; v < 8; ++v)
    {
      qR[v] = T(0);
      qLn[v] = T(0);
    }
    const int cc = VW::c1(c[0], c[1], c[2]);
    if (valid)
    {
      face_side<T, 3, BS, 0, -1>(V, cc, qR);
      face_side<T, 3, BS, 0, 1>(V, cc, qLn);
    }
    // the left state of the face below this cell: the cell before it, one lane back
    T qL[8];
    for (int v = 0; v < 8; ++v)
      qL[v] = lane_take(qLn[v], back);
    T          out[5] = { T(0), T(0), T(0), T(0), T(0) };
    const bool own = valid && p >= 1;
    if (own)
    {
      const bool lo = c[0] == 0, hi = c[0] == BS;
      if ((lo && fine_lo) || (hi && fine_hi))
        fine_face_flux<T, 3, BS, 0>(S, o, c, hi ? 1 : 0, out);
      else
      {
        T       f[8];
        const T bn = V.bf(0, c[0], c[1], c[2]) + V.sf(0, cc);
        face_hlld<T, 0>(qL, qR, bn, dtdS, S.gt, f);
        for (int vi = 0; vi < 5; ++vi)
          out[vi] = f[kFluxVar[vi]];
      }
    }
    // cell a: (0 + F_low) - F_high, the high face's flux one lane on
    T up[5];
    for (int vi = 0; vi < 5; ++vi)
      up[vi] = lane_take(out[vi], next);
    if (own && a < kTS)
    {
      const int cl = a + kTS * (lj + kTS * lk);
      for (int vi = 0; vi < 5; ++vi)
        ld[vi][cl] = double(out[vi]) - double(up[vi]);
    }
  }
}

template <typename T, int BS>
__device__ __forceinline__ void
flux_march_item(const FusedSlot<T> & S, long r0, double (*ld)[5][512])
{
  using G = Geo<3, BS>;
  using IX = IdxC<3, BS>;
  const int              o = int(r0) / G::TPB;
  const View<T, 3, BS>   V(S, o);
  const int              tl = int(r0) - o * G::TPB;
  constexpr int          nb1 = G::NB1;
  const int              org[3] = { (tl % nb1) * kTS, ((tl / nb1) % nb1) * kTS, (tl / (nb1 * nb1)) * kTS };
  const T                dtdS = T(S.dt / dxd<BS>(S.m, S.scaling, o));
  const int              w = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int              lane = int(threadIdx.x) & 63;
  if (w == 0)
    flux_march_dir<T, BS, 2>(S, V, o, org, lane, dtdS, ld[2]);
  else if (w == 1)
    flux_march_dir<T, BS, 1>(S, V, o, org, lane, dtdS, ld[1]);
  else
    flux_rows_x<T, BS>(S, V, o, org, w - 2, lane, dtdS, ld[0]);
  __syncthreads();
  constexpr int ncell = IX::cells();
  for (int cl = int(threadIdx.x); cl < kTS * kTS * kTS; cl += 256)
  {
    const int cidx = IX::cell(org[0] + cl % kTS, org[1] + (cl / kTS) % kTS, org[2] + cl / (kTS * kTS));
    for (int v = 0; v < 5; ++v)
    {
      double acc = S.U[(long(o) * 5 + v) * ncell + cidx];
      acc = acc + ld[0][v][cl];
      acc = acc + ld[1][v][cl];
      acc = acc + ld[2][v][cl];
      S.U2[(long(o) * 5 + v) * ncell + cidx] = acc;
    }
  }
  __syncthreads(); // the next tile of this workgroup reuses the LDS
}

template <typename T, int BSEL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_flux_march(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  __shared__ double ld[3][5][kTS * kTS * kTS];
  for (long g = xcd_group(); g < total; g += gridDim.x)
  {
    const int            s = slot_of(off, nslots, g);
    const FusedSlot<T> & S = slots[s];
    const long           r0 = g - off[s];
    if (r0 >= S.cnt[3])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (flux_march_item<T, BS>(S, r0, ld)));
  }
}


// corner state of an edge (ComputeEmfAndStoreFunctor.cpp:105-367); KIND 0 LB, 1 RT, 2 RB, 3 LT,
// (D0, D1) the two directions across the edge; HK: HANCOCK (the face-B source enters the face
// values, ComputeEmfAndStoreFunctor.cpp:189-193, 297-301), else RK2_SSP
template <typename T, int DIM, int BS, int KIND, int D0, int D1, bool HK>
__device__ __forceinline__ void
edge_state(const View<T, DIM, BS> & V, int is, int js, int ks, T * q)
{
  using VW = View<T, DIM, BS>;
  const int c0 = VW::c1(is, js, ks);
  for (int v = 0; v < 8; ++v)
    q[v] = V.q2(v, c0);
  constexpr int s0 = (KIND == 0 || KIND == 3) ? -1 : 1;
  constexpr int s1 = (KIND == 0 || KIND == 2) ? -1 : 1;
  constexpr int b0 = (KIND == 0 || KIND == 2) ? -1 : 1;
  constexpr int b1 = (KIND == 0 || KIND == 3) ? -1 : 1;
  int p0[3] = { is, js, ks }, p1[3] = { is, js, ks };
  if (KIND == 1 || KIND == 2)
    p0[D0] += 1;
  if (KIND == 1 || KIND == 3)
    p1[D1] += 1;
  const int cp0 = VW::c1(p0[0], p0[1], p0[2]);
  const int cp1 = VW::c1(p1[0], p1[1], p1[2]);
  T         B0 = V.bf(D0, p0[0], p0[1], p0[2]);
  T         B1 = V.bf(D1, p1[0], p1[1], p1[2]);
  if (HK)
  {
    B0 += V.sf(D0, cp0);
    B1 += V.sf(D1, cp1);
  }
  // face-slope record index of (component, direction)
  constexpr int fs01 = DIM == 2 ? D0 : (D0 == 0 ? D1 - 1 : (D0 == 1 ? (D1 == 0 ? 2 : 3) : (D1 == 0 ? 4 : 5)));
  constexpr int fs10 = DIM == 2 ? D1 : (D1 == 0 ? D0 - 1 : (D1 == 1 ? (D0 == 0 ? 2 : 3) : (D0 == 0 ? 4 : 5)));
  const T dB0d1 = V.fs(fs01, cp0);
  const T dB1d0 = V.fs(fs10, cp1);
  const T f0 = T(s0), f1 = T(s1);
  for (int v = 0; v < 5; ++v)
    q[v] += T(0.5) * (f0 * V.sl(D0, v, c0) + f1 * V.sl(D1, v, c0));
  auto tr = [&](int v) { return T(0.5) * (f0 * V.sl(D0, v, c0) + f1 * V.sl(D1, v, c0)); };
  if (DIM == 2)
  {
    q[kA] = B0 + T(0.5) * (T(b0) * dB0d1);
    q[kB] = B1 + T(0.5) * (T(b1) * dB1d0);
    q[kC] += tr(kC);
  }
  else
  {
    if (D0 == 0)
      q[kA] = B0 + T(0.5) * (T(b0) * dB0d1);
    else
      q[kA] += tr(kA);
    if (D0 == 1)
      q[kB] = B0 + T(0.5) * (T(b0) * dB0d1);
    else
      q[kB] += tr(kB);
    if (D1 == 1)
      q[kB] = B1 + T(0.5) * (T(b1) * dB1d0);
    else
      q[kB] += tr(kB);
    if (D1 == 2)
      q[kC] = B1 + T(0.5) * (T(b1) * dB1d0);
    else
      q[kC] += tr(kC);
  }
  q[kD] = fmax(V.S.gt.smallr, q[kD]);
  q[kP] = fmax(V.S.gt.smallp * q[kD], q[kP]);
}

// the EMF along axis AX at edge r (of EPD) of block o
template <typename T, int DIM, int BS, int AX, bool HK>
__device__ __forceinline__ void
edge_at(const FusedSlot<T> & S, const View<T, DIM, BS> & V, int o, int r, T dtdx)
{
  using G = Geo<DIM, BS>;
  constexpr int e0 = (DIM == 3 && AX == 0) ? BS : G::NE;
  constexpr int e1 = (DIM == 3 && AX == 1) ? BS : G::NE;
  const int     i = r % e0, j = (r / e0) % e1, k = DIM == 3 ? r / (e0 * e1) : 0;
  T             qe[4][8];
  if (AX == 2)
  {
    edge_state<T, DIM, BS, 0, 0, 1, HK>(V, i, j, k, qe[kLB]);
    edge_state<T, DIM, BS, 1, 0, 1, HK>(V, i - 1, j - 1, k, qe[kRT]);
    edge_state<T, DIM, BS, 2, 0, 1, HK>(V, i - 1, j, k, qe[kRB]);
    edge_state<T, DIM, BS, 3, 0, 1, HK>(V, i, j - 1, k, qe[kLT]);
  }
  else if (AX == 1)
  {
    edge_state<T, DIM, BS, 0, 0, 2, HK>(V, i, j, k, qe[kLB]);
    edge_state<T, DIM, BS, 1, 0, 2, HK>(V, i - 1, j, k - 1, qe[kRT]);
    edge_state<T, DIM, BS, 2, 0, 2, HK>(V, i - 1, j, k, qe[kLT]);
    edge_state<T, DIM, BS, 3, 0, 2, HK>(V, i, j, k - 1, qe[kRB]);
  }
  else
  {
    edge_state<T, DIM, BS, 0, 1, 2, HK>(V, i, j, k, qe[kLB]);
    edge_state<T, DIM, BS, 1, 1, 2, HK>(V, i, j - 1, k - 1, qe[kRT]);
    edge_state<T, DIM, BS, 2, 1, 2, HK>(V, i, j - 1, k, qe[kRB]);
    edge_state<T, DIM, BS, 3, 1, 2, HK>(V, i, j, k - 1, qe[kLT]);
  }
  const T       E = edge_emf_fast<AX>(qe, S.gt) * dtdx;
  constexpr int comp = DIM == 3 ? AX : 0;
  S.emf[(long(o) * G::NCE + comp) * G::NEC + IdxC<DIM, BS>::edge(i, j, k)] = E;
}

// Corner states by parts (ComputeEmfAndStoreFunctor.cpp:105-367).  The corner of KIND (0 LB,
// 1 RT, 2 RB, 3 LT) reconstructed from cell c0 across the directions (D0, D1) is
//   q = Q2(c0) + 1/2 (s0 dQ/dD0 + s1 dQ/dD1)      for rho, p, the velocity and the B component
//                                                  along the edge (the "cell part"),
//   B_D0 = Bf_D0(p0) + 1/2 b0 dB_D0/dD1 (p0),  B_D1 = Bf_D1(p1) + 1/2 b1 dB_D1/dD0 (p1)
// (the "face part"), Bf the face field plus, for HANCOCK, the face-B source term, p0 = c0 + e_D0
// for KIND 1, 2 and p1 = c0 + e_D1 for KIND 1, 3: edge_state's arithmetic, split.  As edge_state
// does, the y-edges add the dBy slope terms twice, and the x-edges add the cell's dBy slope terms
// to their face value of By.
//
// The cell part of KIND at cell c0 across (D0, D1) = (x, y) (z-edges), (x, z) (y-edges) or (y, z)
// (x-edges), edge_part_of: h[0..4] rho, p, u, v, w (floored as in the reference), h[5] the B
// component that stays with the cell (z-edges: Bz; y-edges: By; x-edges: Bx), h[6] (x-edges) the
// dBy slope term of the By face value.
//
// the corner state q[8] from its cell part h and its face values (bf0, fs0) = (Bf_D0, dB_D0/dD1)
// at p0 and (bf1, fs1) = (Bf_D1, dB_D1/dD0) at p1
template <typename T, int KIND, int D0, int D1>
__device__ __forceinline__ void
edge_corner(const T * h, T bf0, T fs0, T bf1, T fs1, T * q)
{
  constexpr int b0 = (KIND == 0 || KIND == 2) ? -1 : 1;
  constexpr int b1 = (KIND == 0 || KIND == 3) ? -1 : 1;
  for (int v = 0; v < 5; ++v)
    q[v] = h[v];
  const T B0 = bf0 + T(0.5) * (T(b0) * fs0);
  const T B1 = bf1 + T(0.5) * (T(b1) * fs1);
  if (D0 == 0 && D1 == 1)
  {
    q[kA] = B0;
    q[kB] = B1;
    q[kC] = h[5];
  }
  else if (D0 == 0)
  {
    q[kA] = B0;
    q[kB] = h[5];
    q[kC] = B1;
  }
  else
  {
    q[kA] = h[5];
    q[kB] = B0 + h[6];
    q[kC] = B1;
  }
}

// Bf_c (plus the face-B source for HANCOCK) and the transverse slope of face component c along
// d at the low face of cell (a, b, e)
template <typename T, int DIM, int BS, int C, int D, bool HK>
__device__ __forceinline__ void
edge_face(const View<T, DIM, BS> & V, int a, int b, int e, T & bf, T & fs)
{
  using VW = View<T, DIM, BS>;
  constexpr int fsi = DIM == 2 ? C : (C == 0 ? D - 1 : (C == 1 ? (D == 0 ? 2 : 3) : (D == 0 ? 4 : 5)));
  const int     c1 = VW::c1(a, b, e);
  bf = V.bf(C, a, b, e);
  if (HK)
    bf += V.sf(C, c1);
  fs = V.fs(fsi, c1);
}


// What a marching lane reads of its cell c for the edges across (D0, D1): the values the cell
// parts take (rho, p, u, v, w and the cell's B component, their slopes along D0 and D1, and for
// the x-edges the slopes of By), Bf_D1 and Bf_D0 at c with their transverse slopes (edge_face;
// Bf_D0 only on edge lanes).
template <typename T, int D0, int D1>
struct EdgeRaw
{
  static constexpr int NQ = 6;                   // q2: rho, p, u, v, w, the cell's B component
  static constexpr int NS = D0 == 1 ? 7 : 6;     // slopes: the same and, x-edges, By
  static constexpr int BC = (D0 == 0 && D1 == 1) ? kC : (D0 == 0 ? kB : kA);
  T                    q[NQ], s0[NS], s1[NS], b1, f1, b0, f0;
};

template <typename T, int DIM, int BS, int D0, int D1, bool HK, bool B1 = true>
__device__ __forceinline__ void
edge_raw_load(const View<T, DIM, BS> & V, const int * c, bool edge, EdgeRaw<T, D0, D1> & R)
{
  using ER = EdgeRaw<T, D0, D1>;
  const int cA = View<T, DIM, BS>::c1(c[0], c[1], c[2]);
  for (int v = 0; v < 5; ++v)
  {
    R.q[v] = V.q2(v, cA);
    R.s0[v] = V.sl(D0, v, cA);
    R.s1[v] = V.sl(D1, v, cA);
  }
  R.q[5] = V.q2(ER::BC, cA);
  R.s0[5] = V.sl(D0, ER::BC, cA);
  R.s1[5] = V.sl(D1, ER::BC, cA);
  if constexpr (ER::NS == 7)
  {
    R.s0[6] = V.sl(D0, kB, cA);
    R.s1[6] = V.sl(D1, kB, cA);
  }
  R.b1 = T(0);
  R.f1 = T(0);
  if (B1)
    edge_face<T, DIM, BS, D1, D0, HK>(V, c[0], c[1], c[2], R.b1, R.f1);
  R.b0 = T(0);
  R.f0 = T(0);
  if (edge)
    edge_face<T, DIM, BS, D0, D1, HK>(V, c[0], c[1], c[2], R.b0, R.f0);
}

// the cell part of KIND from the values read (edge_at's corner arithmetic, the same operations in
// the same order)
template <typename T, int KIND, int D0, int D1>
__device__ __forceinline__ void
edge_part_of(const EdgeRaw<T, D0, D1> & R, const Gas<T> & gt, T * h)
{
  constexpr int s0 = (KIND == 0 || KIND == 3) ? -1 : 1;
  constexpr int s1 = (KIND == 0 || KIND == 2) ? -1 : 1;
  const T       f0 = T(s0), f1 = T(s1);
  auto          tr = [&](int n) { return T(0.5) * (f0 * R.s0[n] + f1 * R.s1[n]); };
  for (int v = 0; v < 5; ++v)
    h[v] = R.q[v] + T(0.5) * (f0 * R.s0[v] + f1 * R.s1[v]);
  if (D0 == 0 && D1 == 1)
    h[5] = R.q[5] + tr(5);
  else if (D0 == 0)
  {
    const T b = R.q[5] + tr(5);
    h[5] = b + tr(5);
  }
  else
  {
    h[5] = R.q[5] + tr(5);
    h[6] = tr(6);
  }
  h[kD] = fmax(gt.smallr, h[kD]);
  h[kP] = fmax(gt.smallp * h[kD], h[kP]);
}

// The x- (AX 0) or y-edges (AX 1) of an (i, j) position, marching along z: each iteration reads
// the records of A = (i, j, k) and of A' = A - e_D0 (D0 = y for x-edges, x for y-edges) and
// forms KIND 0 at A and KIND 2 at A'; KIND 3 and 1 (at C = A - e_z and C' = A' - e_z) and Bf_D0
// at C are those formed at the plane before, carried in registers.  No cross-lane reads: every
// lane holds an edge.  (With the next plane's reads in flight the kernel spilled 77 VGPRs.)
// Same arithmetic as edge_at, the same states bit for bit.
template <typename T, int DIM, int BS, int AX, bool HK>
__device__ __forceinline__ void
edge_column(const FusedSlot<T> & S, const View<T, DIM, BS> & V, int o, int c[3], T dtdx)
{
  static_assert(DIM == 3 && AX != 2, "edge_column: 3D x- and y-edges");
  using G = Geo<DIM, BS>;
  constexpr int D0 = AX == 0 ? 1 : 0, D1 = 2, NH = AX == 0 ? 7 : 6;
  // the reference's corner order (RT, RB, LT, LB) per axis: KIND 2 is RB for x-edges, LT for
  // y-edges
  constexpr int slot3 = AX == 1 ? kRB : kLT;
  constexpr int slot2 = AX == 1 ? kLT : kRB;
  int           cp[3] = { c[0] - (D0 == 0), c[1] - (D0 == 1), -1 };
  T             hC3[NH], hC1[NH], bC0 = T(0), fC0 = T(0);
  for (int m = -1; m < G::NE; ++m)
  {
    EdgeRaw<T, D0, D1> RA, RP;
    c[2] = m;
    cp[2] = m;
    edge_raw_load<T, DIM, BS, D0, D1, HK>(V, c, true, RA);
    edge_raw_load<T, DIM, BS, D0, D1, HK>(V, cp, false, RP);
    T hA0[NH], hA2[NH], h3[NH], h1[NH];
    edge_part_of<T, 0, D0, D1>(RA, S.gt, hA0);
    edge_part_of<T, 3, D0, D1>(RA, S.gt, h3);
    edge_part_of<T, 2, D0, D1>(RP, S.gt, hA2);
    edge_part_of<T, 1, D0, D1>(RP, S.gt, h1);
    const T bA1 = RA.b1, fA1 = RA.f1, bA0 = RA.b0, fA0 = RA.f0, bP1 = RP.b1, fP1 = RP.f1;
    if (m >= 0)
    {
      T qe[4][8];
      edge_corner<T, 0, D0, D1>(hA0, bA0, fA0, bA1, fA1, qe[kLB]);
      edge_corner<T, 3, D0, D1>(hC3, bC0, fC0, bA1, fA1, qe[slot3]);
      edge_corner<T, 2, D0, D1>(hA2, bA0, fA0, bP1, fP1, qe[slot2]);
      edge_corner<T, 1, D0, D1>(hC1, bC0, fC0, bP1, fP1, qe[kRT]);
      const T E = edge_emf_fast<AX>(qe, S.gt) * dtdx;
      S.emf[(long(o) * G::NCE + AX) * G::NEC + IdxC<DIM, BS>::edge(c[0], c[1], c[2])] = E;
    }
    for (int v = 0; v < NH; ++v)
    {
      hC3[v] = h3[v];
      hC1[v] = h1[v];
    }
    bC0 = bA0;
    fC0 = fA0;
  }
}

// The z-edges of an (i, j) position, marching along z beside the x- and y-edges (which march
// along z too), so that the three directions read the records of the same planes close together
// in time and share them in the L2.  The z-edge at (i, j, k) has its four cells in plane k: A =
// (i, j, k) (KIND 0), C = A - e_y (KIND 3), A' = A - e_x (KIND 2) and C' = A' - e_y (KIND 1), with
// Bf_x at A and C and Bf_y at A and A'; each iteration reads the four records, no cross-lane
// reads.  Same arithmetic as edge_at, the same states bit for bit.
template <typename T, int DIM, int BS, bool HK>
__device__ __forceinline__ void
edge_column_z(const FusedSlot<T> & S, const View<T, DIM, BS> & V, int o, int c[3], T dtdx)
{
  static_assert(DIM == 3, "edge_column_z: 3D edges");
  using G = Geo<DIM, BS>;
  constexpr int D0 = 0, D1 = 1, NH = 6;
  for (int m = 0; m < BS; ++m)
  {
    EdgeRaw<T, D0, D1> RA, RC, RAp, RCp;
    int                cC[3] = { c[0], c[1] - 1, m }, cAp[3] = { c[0] - 1, c[1], m };
    int                cCp[3] = { c[0] - 1, c[1] - 1, m };
    c[2] = m;
    edge_raw_load<T, DIM, BS, D0, D1, HK>(V, c, true, RA);
    edge_raw_load<T, DIM, BS, D0, D1, HK, false>(V, cC, true, RC);
    edge_raw_load<T, DIM, BS, D0, D1, HK>(V, cAp, false, RAp);
    edge_raw_load<T, DIM, BS, D0, D1, HK, false>(V, cCp, false, RCp);
    T hA0[NH], hAp[NH], hC3[NH], hCp[NH];
    edge_part_of<T, 0, D0, D1>(RA, S.gt, hA0);
    edge_part_of<T, 2, D0, D1>(RAp, S.gt, hAp);
    edge_part_of<T, 3, D0, D1>(RC, S.gt, hC3);
    edge_part_of<T, 1, D0, D1>(RCp, S.gt, hCp);
    T qe[4][8];
    edge_corner<T, 0, D0, D1>(hA0, RA.b0, RA.f0, RA.b1, RA.f1, qe[kLB]);
    edge_corner<T, 3, D0, D1>(hC3, RC.b0, RC.f0, RA.b1, RA.f1, qe[kLT]);
    edge_corner<T, 2, D0, D1>(hAp, RA.b0, RA.f0, RAp.b1, RAp.f1, qe[kRB]);
    edge_corner<T, 1, D0, D1>(hCp, RC.b0, RC.f0, RAp.b1, RAp.f1, qe[kRT]);
    const T E = edge_emf_fast<2>(qe, S.gt) * dtdx;
    S.emf[(long(o) * G::NCE + 2) * G::NEC + IdxC<DIM, BS>::edge(c[0], c[1], c[2])] = E;
  }
}

// ---------------------------------------------------------------------------------------------
// k_edge: every edge of a block once.  In 3D every wavefront marches along z, one (i, j) position
// per lane, i fastest: the x- and y-edges read two cells per plane and carry the plane before's
// corner parts (edge_column), the z-edges read their four cells in each plane (edge_column_z).
// The three directions march along z together and share each plane's records in the L2:
// marching the z-edges along y instead (one cell read per edge, the previous row carried) read
// 4.0 GB from HBM on the 16^3 class of MIX against 3.1 along z (BL3D k_edge 1.76 -> 1.73 ms; the
// 32^3 class, whose marches drift apart by more planes than the L2 holds, was unchanged).
// Lanes sharing the cells of the row before by cross-lane reads (8 x 8 tiles for the x-edges,
// rows of b + 2 with a halo lane for the y- and z-edges) were slower: MIX k_edge 3.92 (x), 3.79
// (y) and 3.76 (z) ms against 3.78, 3.74 and 3.67 with one edge per lane.  G::EWB wavefronts per
// block: x-edges first, then y, then z, each by plane wavefront.  A lane holds up to 256
// registers, so the 3D kernels run 2 wavefronts per SIMD (at 3, with 30 VGPRs spilled: MIX
// k_edge 3.67 -> 3.83 ms).  In 2D a plane of lanes would be one
// row of b + 1 positions, too short for a wavefront: the edges run one per lane (edge_at), 4
// wavefronts per SIMD.  A wavefront's direction is uniform and compiled in.
// ---------------------------------------------------------------------------------------------
template <typename T, int DIM, int BS>
__device__ __forceinline__ void
edge_item(const FusedSlot<T> & S, long r0)
{
  using G = Geo<DIM, BS>;
  constexpr int per = G::EWB * 64;
  const int     o = __builtin_amdgcn_readfirstlane(int(r0) / per);
  const int     r = int(r0) - o * per;
  const int     w = __builtin_amdgcn_readfirstlane(r / 64);
  const int     lane = r - w * 64;
  const View<T, DIM, BS> V(S, o);
  const T                dtdx = T(S.dt / dxd<BS>(S.m, S.scaling, o));
  // the integrator is uniform per slot, so per wavefront; each variant is compiled on its own
  if constexpr (DIM == 2)
  {
    const int re = w * 64 + lane;
    if (re >= G::EPD)
      return;
    if (S.rk2)
      edge_at<T, DIM, BS, 2, false>(S, V, o, re, dtdx);
    else
      edge_at<T, DIM, BS, 2, true>(S, V, o, re, dtdx);
  }
  else
  {
    // x-edges, then y, then z, each by plane wavefront
    const int     a = w < G::NWX ? 0 : (w < G::NWX + G::NWP ? 1 : 2);
    const int     pw = a == 0 ? w : (a == 1 ? w - G::NWX : w - G::NWX - G::NWP);
    if (a == 2)
    {
      // z-edges: (i, j), b + 1 by b + 1, one edge per lane
      const int p = pw * 64 + lane;
      if (p >= G::NE * G::NE)
        return;
      int c[3] = { p % G::NE, p / G::NE, 0 };
      if (S.rk2)
        edge_column_z<T, DIM, BS, false>(S, V, o, c, dtdx);
      else
        edge_column_z<T, DIM, BS, true>(S, V, o, c, dtdx);
      return;
    }
    // x-edges: (i, j), b along x by b + 1 along y; y-edges: b + 1 along x by b along y; one edge
    // per lane
    const int p = pw * 64 + lane;
    if (p >= (a == 0 ? G::NPLX : G::NPLY))
      return;
    const int nx = a == 0 ? BS : BS + 1;
    int       c[3] = { p % nx, p / nx, 0 };
    if (S.rk2)
    {
      if (a == 0)
        edge_column<T, DIM, BS, 0, false>(S, V, o, c, dtdx);
      else
        edge_column<T, DIM, BS, 1, false>(S, V, o, c, dtdx);
    }
    else if (a == 0)
      edge_column<T, DIM, BS, 0, true>(S, V, o, c, dtdx);
    else
      edge_column<T, DIM, BS, 1, true>(S, V, o, c, dtdx);
  }
}

template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(DIM == 3 ? 2 : 4)))
k_edge(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[4])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (edge_item<T, DIM, BS>(S, r0)));
  }
}


// CorrectEmfAtBlockBorder from each slot's gather list: dst = (sum of sources) / cnt / ratio,
// summed in list order; the values are staged in fixval and scattered by k_scatter.
template <typename T>
__global__ void __launch_bounds__(256)
k_fix(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.m.nfix)
      continue;
    const EmfFix f = S.m.fix[r0];
    T            v = 0;
    for (int n = 0; n < f.nsrc; ++n)
      v += S.emf[S.m.fix_src[f.first + n]];
    S.fixval[r0] = v / T(f.cnt) / T(f.ratio);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_scatter(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.m.nfix)
      continue;
    S.emf[S.m.fix[r0].dst] = S.fixval[r0];
  }
}

// ComputeDtMHD over every slot's owned cells (ComputeDtMHDFunctor.cpp:86-151): per workgroup a
// max of (c_f + |v|)/dx summed over directions, one atomic into the slot's reduction word;
// a NaN candidate raises flag bit 2.  Every workgroup lies in one slot (slot ranges are
// padded to 512 items).
template <typename T, int DIM, int BS>
__device__ __forceinline__ double
inv_dt_item(const FusedSlot<T> & S, long r0)
{
  using IX = IdxC<DIM, BS>;
  constexpr int ncell = IX::cells();
  const int     o = int(r0) / ncell;
  const int     ci = int(r0) - o * ncell;
  const int     i = ci % BS, j = (ci / BS) % BS, k = DIM == 3 ? ci / (BS * BS) : 0;
  const double * Bo = S.B + long(o) * IX::fper();
  const Gas<double> & g = S.gd;
  double        u[8], q[8];
  for (int v = 0; v < 5; ++v)
    u[v] = S.U[(long(o) * 5 + v) * ncell + ci];
  if (DIM == 2)
  {
    u[kA] = 0.5 * (Bo[IX::face(0, i, j, 0)] + Bo[IX::face(0, i + 1, j, 0)]);
    u[kB] = 0.5 * (Bo[IX::face(1, i, j, 0)] + Bo[IX::face(1, i, j + 1, 0)]);
    u[kC] = 0.5 * (Bo[IX::face(2, i, j, 0)] + Bo[IX::face(2, i, j, 0)]);
  }
  else
  {
    u[kA] = 0.5 * (Bo[IX::face(0, i, j, k)] + Bo[IX::face(0, i + 1, j, k)]);
    u[kB] = 0.5 * (Bo[IX::face(1, i, j, k)] + Bo[IX::face(1, i, j + 1, k)]);
    u[kC] = 0.5 * (Bo[IX::face(2, i, j, k)] + Bo[IX::face(2, i, j, k + 1)]);
  }
  to_primitive(u, q, g);
  return cell_inv_dt<DIM, BS>(q, S, o);
}

template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256)
k_inv_dt(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  // a bounded grid: each workgroup folds many groups of 256 cells into a running maximum and
  // flushes it (block reduction + one atomic) whenever its next group lies in another slot
  __shared__ double wmax[4];
  double cand = 0.0;
  bool   nan = false;
  int    cur = -1;
  auto   flush = [&](int s) {
    for (int o = 32; o > 0; o >>= 1)
      cand = fmax(cand, __shfl_xor(cand, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
      wmax[wave] = cand;
    if (__any(nan) && lane == 0)
      atomicOr(slots[s].red + 1, 2ull);
    __syncthreads();
    if (threadIdx.x == 0)
    {
      double mx = wmax[0];
      for (int w = 1; w < 4; ++w)
        mx = fmax(mx, wmax[w]);
      atomicMax(slots[s].red, (unsigned long long)__double_as_longlong(mx));
    }
    __syncthreads();
    cand = 0.0;
    nan = false;
  };
  for (long g0 = xcd_group() * blockDim.x; g0 < total; g0 += long(gridDim.x) * blockDim.x)
  {
    const int s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, g0));
    if (s != cur)
    {
      if (cur >= 0)
        flush(cur);
      cur = s;
    }
    const FusedSlot<T> & S = slots[s];
    const long           r0 = g0 + threadIdx.x - off[s];
    if (r0 < S.cnt[kStageInvDt])
    {
      double c = 0.0;
      MHDCT_BS_DISPATCH(S.m.b, (c = inv_dt_item<T, DIM, BS>(S, r0)));
      if (c != c)
        nan = true;
      else
        cand = fmax(cand, c);
    }
  }
  if (cur >= 0)
    flush(cur);
}

// Per slot: clear the reduction words before k_inv_dt (a kernel, so that the step can be
// captured in a graph), then dt = cfl / inv_dt into both parities of the slot array, and the
// slot's clock and step count advanced.
template <typename T>
__global__ void __launch_bounds__(64)
k_dt_clear(FusedSlot<T> * __restrict__ slots, int nslots)
{
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += gridDim.x * blockDim.x)
  {
    slots[s].red[0] = 0;
  }
}

template <typename T>
__global__ void __launch_bounds__(64)
k_dt_set(FusedSlot<T> * even, FusedSlot<T> * odd, FusedSlot<T> * even2, FusedSlot<T> * odd2, int nslots,
         double * __restrict__ clock)
{
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += gridDim.x * blockDim.x)
  {
    const double inv = __longlong_as_double((long long)even[s].red[0]);
    const double dt = even[s].cfl / inv;
    even[s].dt = dt;
    odd[s].dt = dt;
    if (even2)
      even2[s].dt = dt; // the second-stage arrays of an RK2_SSP batch
    if (odd2)
      odd2[s].dt = dt;
    clock[2 * s] += dt;
    clock[2 * s + 1] += 1.0;
  }
}

// face-B update (ReadEmfAndUpdateFunctor.cpp:58-163), FP64 accumulation
template <typename T, int DIM, int BS>
__device__ __forceinline__ void
bupdate_item(const FusedSlot<T> & S, long r0)
{

  using IX = IdxC<DIM, BS>;
    constexpr int        per = IX::fper();
    const int o = int(r0) / per;
    const int fi = int(r0) - o * per;
    int       r = fi;
    int       c = 0;
    if (r >= IX::fsize(0))
    {
      r -= IX::fsize(0);
      c = 1;
      if (r >= IX::fsize(1))
      {
        r -= IX::fsize(1);
        c = 2;
      }
    }
    const int     nx = c == 0 ? BS + 1 : BS;
    const int     ny = c == 1 ? BS + 1 : BS;
    const int     i = r % nx, j = (r / nx) % ny, k = DIM == 3 ? r / (nx * ny) : 0;
    constexpr int ne = IX::ecells();
    const double  Bv = S.B[long(o) * per + fi];
    if (DIM == 2)
    {
      const T * e = S.emf + long(o) * ne;
      if (c == 0)
        S.B2[long(o) * per + fi] = Bv + (double(e[IX::edge(i, j + 1, 0)]) - double(e[IX::edge(i, j, 0)]));
      else if (c == 1)
        S.B2[long(o) * per + fi] = Bv - (double(e[IX::edge(i + 1, j, 0)]) - double(e[IX::edge(i, j, 0)]));
      return;
    }
    const T * ex = S.emf + (long(o) * 3 + 0) * ne;
    const T * ey = S.emf + (long(o) * 3 + 1) * ne;
    const T * ez = S.emf + (long(o) * 3 + 2) * ne;
    double    dB = 0;
    if (c == 0)
    {
      dB += (double(ez[IX::edge(i, j + 1, k)]) - double(ez[IX::edge(i, j, k)]));
      dB -= (double(ey[IX::edge(i, j, k + 1)]) - double(ey[IX::edge(i, j, k)]));
    }
    else if (c == 1)
    {
      dB += (double(ex[IX::edge(i, j, k + 1)]) - double(ex[IX::edge(i, j, k)]));
      dB -= (double(ez[IX::edge(i + 1, j, k)]) - double(ez[IX::edge(i, j, k)]));
    }
    else
    {
      dB += (double(ey[IX::edge(i + 1, j, k)]) - double(ey[IX::edge(i, j, k)]));
      dB -= (double(ex[IX::edge(i, j + 1, k)]) - double(ex[IX::edge(i, j, k)]));
    }
    S.B2[long(o) * per + fi] = Bv + dB;
}

template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256)
k_bupdate(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (r0 >= S.cnt[5])
      continue;
    MHDCT_BS_DISPATCH(S.m.b, (bupdate_item<T, DIM, BS>(S, r0)));
  }
}


// a multiple of 8 workgroups (see xcd_group), at most `cap` (itself a multiple of 8)
inline unsigned
grid_for(long n, int threads, int cap)
{
  long g = (n + threads - 1) / threads;
  g = (g + kXcd - 1) / kXcd * kXcd;
  if (g > cap)
    g = cap;
  if (g < kXcd)
    g = kXcd;
  return unsigned(g);
}

// Launch the fused stages of one step (which < 0), or only stage `which` of fused_stage_order.
// Uniform gravity source on the owned cells of each slot that has gravity, over the dt stage's
// work list (AddGravitySourceTerm.cpp:42-100); after the conservative update.
template <typename T, int DIM, int BSEL>
__global__ void __launch_bounds__(256)
k_gravity(const FusedSlot<T> * __restrict__ slots, int nslots, const long * __restrict__ off, long total)
{
  for (long t = xcd_group() * blockDim.x + threadIdx.x; t < total; t += long(gridDim.x) * blockDim.x)
  {
    const int            s = __builtin_amdgcn_readfirstlane(slot_of(off, nslots, t));
    const FusedSlot<T> & S = slots[s];
    const long           r0 = t - off[s];
    if (!S.gd.grav || r0 >= S.cnt[kStageInvDt])
      continue;
    const long ncell = DIM == 3 ? long(S.m.b) * S.m.b * S.m.b : long(S.m.b) * S.m.b;
    const long o = r0 / ncell, ci = r0 - o * ncell;
    gravity_cell(DIM, S.gd, S.dt, S.U + o * 5 * ncell + ci, S.U2 + o * 5 * ncell + ci, ncell);
  }
}

// a batch reduces dt in k_prim (k_front, which replaces k_prim, leaves it to k_inv_dt)
template <typename T>
bool
dt_in_prim(const FusedBatch<T> & bt)
{
  return bt.dev_dt != 0 && bt.tune.front == 0;
}

template <typename T>
void launch_dt_set(const FusedBatch<T> & bt, hipStream_t st);

// k_flux_march for the flux stage of the FP64 3D form at tile depth 8 (in FP32 its 60 KB of FP64
// differences per tile cost more than the reads it saves: BL3D 1.01 -> 1.13 ms); false: not
// launched
template <typename T, int DIM, int BSEL>
bool
launch_flux_march(const FusedBatch<T> & bt, const FusedSlot<T> * sl, hipStream_t st, int tcap)
{
  if constexpr (DIM == 3 && sizeof(T) == sizeof(double))
    if (bt.tune.tile_z == 8 && bt.tune.flux_march)
    {
      hipLaunchKernelGGL((k_flux_march<T, BSEL>), grid_for(bt.tot[kStageTile], 1, tcap), 256, 0, st, sl, bt.nslots,
                         bt.d_off[kStageTile], bt.tot[kStageTile]);
      return true;
    }
  return false;
}

template <typename T, int DIM, int BSEL>
int
launch_stages(const FusedBatch<T> & bt, int parity, hipStream_t st, int which)
{
  const int            n = bt.nslots;
  const FusedSlot<T> * sl = bt.d_slots[parity];
  const int            cap = bt.tune.grid_cap / kXcd * kXcd;
  const int            tcap = bt.tune.tile_cap / kXcd * kXcd;
  auto                 grid = [&](int stage) { return grid_for(bt.tot[stage], 256, cap); };
  // workgroup sizes of the cell and edge stages (k_cell, k_edge): a larger workgroup puts more
  // consecutive wavefronts, which read the same cells, on one CU
  const int            ct = bt.tune.cell_threads, et = bt.tune.edge_threads;
  auto                 on = [&](int stage) { return which < 0 || which == stage; };
  const bool           front = bt.tune.front != 0;
  if (on(kStagePrim) && front)
  {
    // one workgroup per sub-tile (k_front has no grid-stride loop)
    const unsigned gf = grid_for(bt.tot[kStageFront], 1, 1 << 30);
    if (bt.rk2)
      hipLaunchKernelGGL((k_front<T, DIM, BSEL, false>), gf, (front_threads<DIM>()), 0, st, sl, n,
                         bt.d_off[kStageFront], bt.tot[kStageFront]);
    else
      hipLaunchKernelGGL((k_front<T, DIM, BSEL, true>), gf, (front_threads<DIM>()), 0, st, sl, n,
                         bt.d_off[kStageFront], bt.tot[kStageFront]);
  }
  if (on(kStagePrim) && !front)
  {
    hipLaunchKernelGGL((k_prim<T, DIM, BSEL>), grid(kStagePrim), 256, 0, st, sl, n, bt.d_off[kStagePrim],
                       bt.tot[kStagePrim]);
    // dt = cfl / inv_dt once the first stage's k_prim has reduced it (the predictor needs it)
    if (dt_in_prim(bt) && parity < 2)
      launch_dt_set(bt, st);
  }
  if (on(kStageEfield) && !front)
    hipLaunchKernelGGL((k_efield<T, DIM, BSEL>), grid(kStageEfield), 256, 0, st, sl, n, bt.d_off[kStageEfield],
                       bt.tot[kStageEfield]);
  if (on(kStageCell) && !front)
  {
    if (bt.rk2)
      hipLaunchKernelGGL((k_cell<T, DIM, BSEL, false, false>), grid_for(bt.tot[kStageCell], ct, cap), ct, 0, st, sl, n,
                         bt.d_off[kStageCell], bt.tot[kStageCell]);
    else if (bt.visc)
      hipLaunchKernelGGL((k_cell<T, DIM, BSEL, true, true>), grid_for(bt.tot[kStageCell], ct, cap), ct, 0, st, sl, n,
                         bt.d_off[kStageCell], bt.tot[kStageCell]);
    else
      hipLaunchKernelGGL((k_cell<T, DIM, BSEL, true, false>), grid_for(bt.tot[kStageCell], ct, cap), ct, 0, st, sl, n,
                         bt.d_off[kStageCell], bt.tot[kStageCell]);
  }
  if (on(kStageBorder) && bt.tot[kStageBorder] > 0)
    hipLaunchKernelGGL((k_border_flux<T, DIM, BSEL>), grid(kStageBorder), 256, 0, st, sl, n,
                       bt.d_off[kStageBorder], bt.tot[kStageBorder]);
  if (on(kStageTile))
  {
    if (launch_flux_march<T, DIM, BSEL>(bt, sl, st, tcap))
    {
    }
    else if (DIM == 3 && bt.tune.tile_z == 4)
      hipLaunchKernelGGL((k_flux_tile<T, DIM, BSEL, 4>), grid_for(bt.tot[kStageTile] * 2, 1, tcap),
                         (tile_threads<DIM, 4>()), 0, st, sl, n, bt.d_off[kStageTile], bt.tot[kStageTile]);
    else
      hipLaunchKernelGGL((k_flux_tile<T, DIM, BSEL, 8>), grid_for(bt.tot[kStageTile], 1, tcap),
                         (tile_threads<DIM, 8>()), 0, st, sl, n, bt.d_off[kStageTile], bt.tot[kStageTile]);
  }
  if (on(kStageEdge))
    hipLaunchKernelGGL((k_edge<T, DIM, BSEL>), grid_for(bt.tot[kStageEdge], et, cap), et, 0, st, sl, n,
                       bt.d_off[kStageEdge],
                       bt.tot[kStageEdge]);
  if (on(kStageFix) && bt.tot[kStageFix] > 0)
  {
    hipLaunchKernelGGL((k_fix<T>), grid(kStageFix), 256, 0, st, sl, n, bt.d_off[kStageFix], bt.tot[kStageFix]);
    hipLaunchKernelGGL((k_scatter<T>), grid(kStageFix), 256, 0, st, sl, n, bt.d_off[kStageFix],
                       bt.tot[kStageFix]);
  }
  if (on(kStageBupdate))
  {
    hipLaunchKernelGGL((k_bupdate<T, DIM, BSEL>), grid(kStageBupdate), 256, 0, st, sl, n, bt.d_off[kStageBupdate],
                       bt.tot[kStageBupdate]);
    if (bt.grav)
      hipLaunchKernelGGL((k_gravity<T, DIM, BSEL>), grid(kStageInvDt), 256, 0, st, sl, n, bt.d_off[kStageInvDt],
                         bt.tot[kStageInvDt]);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, int DIM, int BSEL>
void
launch_inv_dt(const FusedBatch<T> & bt, int parity, hipStream_t st)
{
  hipLaunchKernelGGL((k_inv_dt<T, DIM, BSEL>), grid_for(bt.tot[kStageInvDt], 256, 2048), 256, 0, st,
                     bt.d_slots[parity], bt.nslots, bt.d_off[kStageInvDt], bt.tot[kStageInvDt]);
}

// the batch restricted to block-size class c: its range of slots, its work lists, its block size
template <typename T>
FusedBatch<T>
class_view(const FusedBatch<T> & bt, int c)
{
  const FusedClass & k = bt.cls[c];
  FusedBatch<T>      v = bt;
  v.bsel = k.bs;
  v.nslots = k.n;
  v.ncls = 0;
  for (int p = 0; p < 4; ++p)
    v.d_slots[p] = bt.d_slots[p] ? bt.d_slots[p] + k.s0 : nullptr;
  for (int s = 0; s < kStages; ++s)
  {
    v.d_off[s] = k.d_off[s];
    v.tot[s] = k.tot[s];
  }
  v.clock = bt.clock ? bt.clock + 2 * k.s0 : nullptr;
  return v;
}

template <typename T>
bool
by_class(const FusedBatch<T> & bt)
{
  return bt.bsel == 0 && bt.ncls > 0 && bt.tune.per_class != 0;
}

template <typename T>
void
launch_dt_set(const FusedBatch<T> & bt, hipStream_t st)
{
  const int n = bt.nslots;
  hipLaunchKernelGGL((k_dt_set<T>), dim3((n + 63) / 64), 64, 0, st, bt.d_slots[0], bt.d_slots[1], bt.d_slots[2],
                     bt.d_slots[3], n, bt.clock);
}

// Clear the slots' reduction words; then, unless k_prim reduces dt (dt_in_prim), the reduction
// and dt = cfl / inv_dt here
template <typename T, int DIM, int BSEL>
int
launch_dt(const FusedBatch<T> & bt, int parity, hipStream_t st)
{
  const int n = bt.nslots;
  hipLaunchKernelGGL((k_dt_clear<T>), dim3((n + 63) / 64), 64, 0, st, bt.d_slots[parity], n);
  if (dt_in_prim(bt))
    return hipGetLastError() == hipSuccess ? 0 : -1;
  if (BSEL == 0 && by_class(bt))
    for (int c = 0; c < bt.ncls; ++c)
    {
      const FusedBatch<T> v = class_view(bt, c);
      if (v.bsel == 8)
        launch_inv_dt<T, DIM, 8>(v, parity, st);
      else if (v.bsel == 16)
        launch_inv_dt<T, DIM, 16>(v, parity, st);
      else
        launch_inv_dt<T, DIM, 32>(v, parity, st);
    }
  else
    launch_inv_dt<T, DIM, BSEL>(bt, parity, st);
  launch_dt_set(bt, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace

bool
fused_supports(int dim, int b)
{
  return (dim == 2 || dim == 3) && (b == 8 || b == 16 || b == 32);
}

long
fused_items_per_block(int dim, int b, int stage, bool fp64)
{
  const long g2 = dim == 3 ? long(b + 4) * (b + 4) * (b + 4) : long(b + 4) * (b + 4);
  const long g1 = dim == 3 ? long(b + 2) * (b + 2) * (b + 2) : long(b + 2) * (b + 2);
  const long nt = dim == 3 ? long(b) * b : long(b);
  const long nb1 = b / kTS;
  const long n1 = b + 1;
  const long nc = dim == 3 ? long(b) * b * b : long(b) * b;
  const long np = b + 3;
  const long npe = dim == 3 ? np * np * np : np * np;
  switch (stage)
  {
  case 0: return nc + (g2 - nc + 63) / 64 * 64;
  case 6: return dim == 3 ? (np * np + 63) / 64 * 64 : (npe + 63) / 64 * 64;
  case 8: return nc;
  case 1: return fused_cell_runs(dim, b, fp64) > 0 ? (long(b + 2) * (b + 2) * fused_cell_runs(dim, b, fp64) + 63) / 64 * 64
                                                 : (g1 + 63) / 64 * 64;
  case 2: return dim * 2 * ((nt + 63) / 64 * 64);
  case 3: return dim == 3 ? nb1 * nb1 * nb1 : nb1 * nb1;
  case 4: return long(fused_edge_waves(dim, b)) * 64;
  case 9: return (dim == 3 ? nb1 * nb1 * (b / fused_front_depth(dim, b)) : nb1 * nb1);
  default: return dim == 3 ? 3 * n1 * b * b : 2 * n1 * b + long(b) * b;
  }
}

long
fused_border_items_per_side(int dim, int b)
{
  const long nt = dim == 3 ? long(b) * b : long(b);
  return (nt + 63) / 64 * 64;
}

template <typename T>
int
fused_launch_stage(const FusedBatch<T> & bt, int parity, int which, hipStream_t st)
{
  if (bt.nslots == 0)
    return 0;
  if (by_class(bt))
  {
    // one launch per block-size class, with the kernels specialised to its block size
    for (int c = 0; c < bt.ncls; ++c)
      if (fused_launch_stage<T>(class_view(bt, c), parity, which, st))
        return -1;
    return 0;
  }
  switch (bt.dim * 100 + bt.bsel)
  {
  case 200: return launch_stages<T, 2, 0>(bt, parity, st, which);
  case 208: return launch_stages<T, 2, 8>(bt, parity, st, which);
  case 216: return launch_stages<T, 2, 16>(bt, parity, st, which);
  case 232: return launch_stages<T, 2, 32>(bt, parity, st, which);
  case 300: return launch_stages<T, 3, 0>(bt, parity, st, which);
  case 308: return launch_stages<T, 3, 8>(bt, parity, st, which);
  case 316: return launch_stages<T, 3, 16>(bt, parity, st, which);
  case 332: return launch_stages<T, 3, 32>(bt, parity, st, which);
  default: return -1;
  }
}

template <typename T>
int
fused_launch(const FusedBatch<T> & bt, int parity, hipStream_t st)
{
  return fused_launch_stage<T>(bt, parity, -1, st);
}

template <typename T>
int
fused_launch_dt(const FusedBatch<T> & bt, int parity, hipStream_t st)
{
  if (bt.nslots == 0)
    return 0;
  switch (bt.dim * 100 + bt.bsel)
  {
  case 200: return launch_dt<T, 2, 0>(bt, parity, st);
  case 208: return launch_dt<T, 2, 8>(bt, parity, st);
  case 216: return launch_dt<T, 2, 16>(bt, parity, st);
  case 232: return launch_dt<T, 2, 32>(bt, parity, st);
  case 300: return launch_dt<T, 3, 0>(bt, parity, st);
  case 308: return launch_dt<T, 3, 8>(bt, parity, st);
  case 316: return launch_dt<T, 3, 16>(bt, parity, st);
  case 332: return launch_dt<T, 3, 32>(bt, parity, st);
  default: return -1;
  }
}

template int fused_launch<double>(const FusedBatch<double> &, int, hipStream_t);
template int fused_launch_stage<double>(const FusedBatch<double> &, int, int, hipStream_t);
template int fused_launch_stage<float>(const FusedBatch<float> &, int, int, hipStream_t);
template int fused_launch<float>(const FusedBatch<float> &, int, hipStream_t);
template int fused_launch_dt<double>(const FusedBatch<double> &, int, hipStream_t);
template int fused_launch_dt<float>(const FusedBatch<float> &, int, hipStream_t);

} // namespace mhdct
