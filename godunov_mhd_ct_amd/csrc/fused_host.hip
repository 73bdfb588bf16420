// fused_host.hip -- scratch and work list of the fused form for one context (one mesh slot).
//
// A context of the fused form owns its slot: the scratch arena, the staging of corrected
// border EMFs, the list of coarse-facing block sides and the reduction words.  It advances
// through a one-slot work list with the caller's dt (mhdct_do_time_step), and the same slot can
// be lent to a batch (batch.hip), which advances many slots with per-slot dt on the device.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "fused.hpp"
#include "fused_host.hpp"

namespace mhdct
{

namespace
{

struct FusedState
{
  bool                fp32 = false;
  FusedSlot<double>   hd{};
  FusedSlot<float>    hf{};
  FusedBatch<double>  bd;
  FusedBatch<float>   bf;
  size_t              bytes = 0;   // device bytes held (arena and small block)
  void *              arena = nullptr; // the scratch records: grow-only across meshes
  size_t              arena_bytes = 0;
  void *              small = nullptr; // the small per-mesh buffers in one block, grow-only too
  size_t              small_bytes = 0;
  double *            visc_fl = nullptr; // viscous fluxes (FP64 form with viscosity), grow-only
  size_t              visc_bytes = 0;
};


template <typename T>
int
build(Context & c, const Mesh & mesh, FusedState & F, FusedBatch<T> & bt, FusedSlot<T> & sl)
{
  const int  dim = c.dim, b = c.b;
  const long nq = c.dm.nowned;
  if (!fused_supports(dim, b))
  {
    c.err = "the fused form is built for block sizes 8, 16 and 32";
    return -4;
  }
  const FusedSizes z(dim, b);
  // The arena holds records that every step writes before it reads them, so a new mesh reuses
  // the arena of the previous one when it fits (an AMR cycle then allocates nothing this
  // large).  It is allocated with a margin of 1/2: an adaptive mesh may grow by half before the
  // arena is reallocated (HBM is plentiful on this device; the margin falls back to nothing
  // when it does not fit).
  const size_t need = size_t(nq > 0 ? nq : 1) * z.total() * sizeof(T);
  if (F.arena_bytes < need)
  {
    if (F.arena)
      (void)hipFree(F.arena);
    F.arena = nullptr;
    F.arena_bytes = 0;
    size_t want = need + need / 2;
    if (hipMalloc(&F.arena, want) != hipSuccess)
    {
      (void)hipGetLastError(); // the margin did not fit: try the exact size
      want = need;
      if (hipMalloc(&F.arena, want) != hipSuccess)
      {
        F.arena = nullptr;
        c.err = "hipMalloc failed for fused scratch";
        return -1;
      }
    }
    (void)hipMemset(F.arena, 0, want);
    F.arena_bytes = want;
  }
  F.bytes += F.arena_bytes;
  T * arena = static_cast<T *>(F.arena);
  std::memset(&sl, 0, sizeof(sl));
  sl.m = c.dm;
  sl.gd = gas_of(c.st);
  sl.gt.gamma0 = T(c.st.gamma0);
  sl.gt.smallr = T(c.st.smallr);
  sl.gt.smallc = T(c.st.smallc);
  sl.gt.smallp = T(sl.gd.smallp);
  sl.gt.slope_type = T(c.st.slope_type);
  sl.gt.grav = c.st.gravity_enabled;
  for (int d = 0; d < 3; ++d)
    sl.gt.g[d] = T(c.st.gravity[d]);
  sl.cfl = c.st.cfl;
  sl.scaling = c.st.scaling;
  sl.abort_eint = c.st.abort_when_negative_eint;
  sl.rk2 = c.st.time_integrator == 1;
  sl.visc_mu = (c.st.viscosity_enabled && c.st.viscosity_predictor && !sl.rk2) ? c.st.mu : -1.0;
  sl.red = c.red;
  sl.dt_red = 0; // the per-mesh form takes dt from the host
  // viscous fluxes, 8 variables per flux cell of one direction, as the plain form lays them out
  if (c.st.viscosity_enabled)
  {
    if (sizeof(T) != sizeof(double))
    {
      c.err = "viscosity runs in the FP64 forms only";
      return -5;
    }
    const long   fc = dim == 3 ? long(b + 1) * b * b : long(b + 1) * b;
    const size_t vneed = size_t(nq > 0 ? nq : 1) * 8 * size_t(fc) * sizeof(double);
    if (F.visc_bytes < vneed)
    {
      if (F.visc_fl)
        (void)hipFree(F.visc_fl);
      F.visc_fl = nullptr;
      F.visc_bytes = 0;
      if (hipMalloc(&F.visc_fl, vneed + vneed / 2) != hipSuccess)
      {
        F.visc_fl = nullptr;
        c.err = "hipMalloc failed for the viscous fluxes";
        return -1;
      }
      F.visc_bytes = vneed + vneed / 2;
    }
    F.bytes += F.visc_bytes;
  }
  T * p = arena;
  sl.Qp = p;  p += nq * z.qp;
  sl.Bg = p;  p += nq * z.bg;
  sl.Q2 = p;  p += nq * z.q2;
  sl.S = p;   p += nq * z.s;
  sl.sF = p;  p += nq * z.sf;
  sl.FS = p;  p += nq * z.fs;
  sl.BF = p;  p += nq * z.bf;
  sl.emf = p; p += nq * z.emf;
  sl.Ed = p;  p += nq * z.ed;
  // border fluxes are needed only where a coarser neighbour sums them
  std::vector<int> sides;
  {
    const int nd = dim == 3 ? 27 : 9, ns = 1 << dim;
    for (long o = 0; o < nq; ++o)
      for (int d = 0; d < dim; ++d)
        for (int side = 0; side < 2; ++side)
        {
          int dl[3] = { 0, 0, 0 };
          dl[d] = side ? 1 : -1;
          const int di = (dl[0] + 1) + 3 * (dl[1] + 1) + (dim == 3 ? 9 * (dl[2] + 1) : 0);
          if (mesh.nb[(o * nd + di) * ns].rel < 0)
            sides.push_back(int(o * 2 * dim + 2 * d + side));
        }
  }
  sl.nbside = int(sides.size());
  for (int k = 0; k < kStages; ++k)
  {
    if (k == kStageBorder)
      sl.cnt[k] = long(sl.nbside) * fused_border_items_per_side(dim, b);
    else if (k == kStageFix)
      sl.cnt[k] = c.dm.nfix;
    else
      sl.cnt[k] = nq * fused_items_per_block(dim, b, k, sizeof(T) == sizeof(double));
    if (sl.cnt[k] >= (1L << 31))
    {
      c.err = "mesh too large for one fused work list (2^31 items per stage)";
      return -4;
    }
  }
  // the one-slot work list of this context
  bt.dim = dim;
  bt.bsel = b;
  bt.rk2 = sl.rk2;
  bt.grav = sl.gd.grav;
  bt.visc = sl.visc_mu >= 0.0;
  bt.nslots = 1;
  for (int k = 0; k < kStages; ++k)
    bt.tot[k] = fused_pad_slot(sl.cnt[k]);
  // The small per-mesh buffers share one device block, reused while it fits: the staged
  // corrected border EMFs, the coarse-facing sides, the work-list offsets (one copy from a host
  // staging area), then the slot and the clock (zeroed).
  auto         al = [](size_t n) { return (n + 255) / 256 * 256; };
  const size_t o_fix = 0;
  const size_t o_sides = o_fix + al(size_t(c.dm.nfix > 0 ? c.dm.nfix : 1) * sizeof(T));
  const size_t o_off = o_sides + al((sides.empty() ? 1 : sides.size()) * sizeof(int));
  const size_t o_slot = o_off + al(size_t(kStages) * 2 * sizeof(long));
  const size_t o_clock = o_slot + al(sizeof(FusedSlot<T>));
  const size_t total = o_clock + al(2 * sizeof(double));
  if (F.small_bytes < total)
  {
    if (F.small)
      (void)hipFree(F.small);
    F.small = nullptr;
    F.small_bytes = 0;
    if (hipMalloc(&F.small, total + total / 2) != hipSuccess)
    {
      F.small = nullptr;
      c.err = "hipMalloc failed for the fused work list";
      return -1;
    }
    F.small_bytes = total + total / 2;
  }
  F.bytes += F.small_bytes;
  char *            base = static_cast<char *>(F.small);
  std::vector<char> host(o_slot, 0);
  if (!sides.empty())
    std::memcpy(host.data() + o_sides, sides.data(), sides.size() * sizeof(int));
  for (int k = 0; k < kStages; ++k)
  {
    const long off[2] = { 0, bt.tot[k] };
    std::memcpy(host.data() + o_off + size_t(k) * 2 * sizeof(long), off, sizeof(off));
  }
  if (hipMemcpy(base, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(base + o_slot, 0, total - o_slot) != hipSuccess)
  {
    c.err = "upload of the fused work list failed";
    return -1;
  }
  sl.fixval = reinterpret_cast<T *>(base + o_fix);
  sl.bside = reinterpret_cast<const int *>(base + o_sides);
  for (int k = 0; k < kStages; ++k)
    bt.d_off[k] = reinterpret_cast<long *>(base + o_off) + 2 * k;
  bt.d_slots[0] = reinterpret_cast<FusedSlot<T> *>(base + o_slot);
  bt.d_slots[1] = bt.d_slots[0];
  bt.clock = reinterpret_cast<double *>(base + o_clock);
  return 0;
}

template <typename T>
int
step(Context & c, FusedState & F, FusedBatch<T> & bt, FusedSlot<T> & sl, const double * U, double * U2,
     const double * B, double * B2, double dt, hipStream_t s)
{
  sl.U = U;
  sl.U2 = U2;
  sl.B = B;
  sl.B2 = B2;
  sl.dt = dt;
  if (hipMemcpyAsync(bt.d_slots[0], &sl, sizeof(sl), hipMemcpyHostToDevice, s) != hipSuccess)
  {
    c.err = "slot upload failed";
    return -1;
  }
  int rc = 0;
  if (!c.st.viscosity_enabled)
    rc = fused_launch<T>(bt, 0, s);
  else
  {
    // the stages one by one, with the viscous fluxes and their update after the hydro update
    // (k_flux_tile) and before the face-B update and the gravity source, in the order of the
    // plain form (GodunovImplemV0.cpp:161-192)
    for (int k = 0; k < 8 && rc == 0; ++k)
    {
      rc = fused_launch_stage<T>(bt, 0, fused_stage_order[k], s);
      if (rc == 0 && fused_stage_order[k] == kStageTile)
      {
        launch_viscous(c, reinterpret_cast<const double *>(sl.Qp), 5, F.visc_fl, U2, B2, dt, s);
        rc = hipGetLastError() == hipSuccess ? 0 : -1;
      }
    }
  }
  if (rc)
  {
    c.err = std::string("fused launch failed: ") + hipGetErrorString(hipGetLastError());
    return -1;
  }
  return 0;
}

} // namespace

int
fused_alloc(Context & c, const Mesh & mesh)
{
  auto * F = static_cast<FusedState *>(c.fused);
  const bool fp32 = c.impl == 12;
  if (F && F->fp32 != fp32)
  {
    fused_free(c);
    F = nullptr;
  }
  if (!F)
  {
    F = new FusedState();
    F->fp32 = fp32;
    c.fused = F;
  }
  F->bytes = 0;
  const int rc = F->fp32 ? build<float>(c, mesh, *F, F->bf, F->hf) : build<double>(c, mesh, *F, F->bd, F->hd);
  c.ps.bytes = F->bytes;
  return rc;
}

void
fused_free(Context & c)
{
  auto * F = static_cast<FusedState *>(c.fused);
  if (!F)
    return;
  if (F->arena)
    (void)hipFree(F->arena);
  if (F->small)
    (void)hipFree(F->small);
  if (F->visc_fl)
    (void)hipFree(F->visc_fl);
  delete F;
  c.fused = nullptr;
}

int
fused_step(Context & c, const double * U, double * U2, const double * B, double * B2, double dt, hipStream_t s)
{
  auto * F = static_cast<FusedState *>(c.fused);
  if (!F)
  {
    c.err = "fused form not allocated (set a mesh first)";
    return -1;
  }
  if (c.dm.nowned == 0)
    return 0;
  return F->fp32 ? step<float>(c, *F, F->bf, F->hf, U, U2, B, B2, dt, s)
                 : step<double>(c, *F, F->bd, F->hd, U, U2, B, B2, dt, s);
}

template <>
const FusedSlot<double> *
fused_slot<double>(const Context & c)
{
  const auto * F = static_cast<const FusedState *>(c.fused);
  return (F && !F->fp32) ? &F->hd : nullptr;
}

template <>
const FusedSlot<float> *
fused_slot<float>(const Context & c)
{
  const auto * F = static_cast<const FusedState *>(c.fused);
  return (F && F->fp32) ? &F->hf : nullptr;
}

} // namespace mhdct
