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
  std::vector<void *> allocs;      // per-mesh buffers, rebuilt with every mesh
  size_t              bytes = 0;
  void *              arena = nullptr; // the scratch records: grow-only across meshes
  size_t              arena_bytes = 0;
};

// the per-mesh buffers of F (the arena stays)
void
free_mesh_buffers(FusedState & F)
{
  for (void * p : F.allocs)
    (void)hipFree(p);
  F.allocs.clear();
  F.bytes = 0;
}

template <typename T>
int
build(Context & c, FusedState & F, FusedBatch<T> & bt, FusedSlot<T> & sl)
{
  const int  dim = c.dim, b = c.b;
  const long nq = c.dm.nowned;
  if (!fused_supports(dim, b))
  {
    c.err = "the fused form is built for block sizes 8, 16 and 32";
    return -4;
  }
  auto alloc = [&](size_t bytes, void ** p) -> int {
    if (bytes == 0)
      bytes = 8;
    if (hipMalloc(p, bytes) != hipSuccess)
    {
      c.err = "hipMalloc failed for fused scratch";
      return -1;
    }
    (void)hipMemset(*p, 0, bytes);
    F.allocs.push_back(*p);
    F.bytes += bytes;
    return 0;
  };
  const FusedSizes z(dim, b);
  // The arena holds records that every step writes before it reads them, so a new mesh reuses
  // the arena of the previous one when it fits (an AMR cycle then allocates nothing this
  // large); it grows with a margin of 1/8 when it does not.
  const size_t need = size_t(nq > 0 ? nq : 1) * z.total() * sizeof(T);
  if (F.arena_bytes < need)
  {
    if (F.arena)
      (void)hipFree(F.arena);
    F.arena = nullptr;
    F.arena_bytes = 0;
    size_t want = need + need / 8;
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
  sl.red = c.red;
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
  if (alloc(size_t(c.dm.nfix > 0 ? c.dm.nfix : 1) * sizeof(T), reinterpret_cast<void **>(&sl.fixval)))
    return -1;
  // border fluxes are needed only where a coarser neighbour sums them
  {
    std::vector<int> sides;
    const int        nd = dim == 3 ? 27 : 9, ns = 1 << dim;
    for (long o = 0; o < nq; ++o)
      for (int d = 0; d < dim; ++d)
        for (int side = 0; side < 2; ++side)
        {
          int dl[3] = { 0, 0, 0 };
          dl[d] = side ? 1 : -1;
          const int di = (dl[0] + 1) + 3 * (dl[1] + 1) + (dim == 3 ? 9 * (dl[2] + 1) : 0);
          if (c.mesh.nb[(o * nd + di) * ns].rel < 0)
            sides.push_back(int(o * 2 * dim + 2 * d + side));
        }
    sl.nbside = int(sides.size());
    int * d_sides = nullptr;
    if (alloc(sides.size() * sizeof(int), reinterpret_cast<void **>(&d_sides)))
      return -1;
    if (!sides.empty() &&
        hipMemcpy(d_sides, sides.data(), sides.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return -1;
    sl.bside = d_sides;
  }
  for (int k = 0; k < kStages; ++k)
  {
    if (k == kStageBorder)
      sl.cnt[k] = long(sl.nbside) * fused_border_items_per_side(dim, b);
    else if (k == kStageFix)
      sl.cnt[k] = c.dm.nfix;
    else
      sl.cnt[k] = nq * fused_items_per_block(dim, b, k);
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
  bt.nslots = 1;
  for (int k = 0; k < kStages; ++k)
  {
    bt.tot[k] = fused_pad_slot(sl.cnt[k]);
    const long off[2] = { 0, bt.tot[k] };
    if (alloc(sizeof(off), reinterpret_cast<void **>(&bt.d_off[k])))
      return -1;
    if (hipMemcpy(bt.d_off[k], off, sizeof(off), hipMemcpyHostToDevice) != hipSuccess)
      return -1;
  }
  if (alloc(sizeof(FusedSlot<T>), reinterpret_cast<void **>(&bt.d_slots[0])))
    return -1;
  bt.d_slots[1] = bt.d_slots[0];
  if (alloc(2 * sizeof(double), reinterpret_cast<void **>(&bt.clock)))
    return -1;
  return 0;
}

template <typename T>
int
step(Context & c, FusedBatch<T> & bt, FusedSlot<T> & sl, const double * U, double * U2, const double * B,
     double * B2, double dt, hipStream_t s)
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
  if (fused_launch<T>(bt, 0, s))
  {
    c.err = std::string("fused launch failed: ") + hipGetErrorString(hipGetLastError());
    return -1;
  }
  return 0;
}

} // namespace

int
fused_alloc(Context & c)
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
  else
    free_mesh_buffers(*F);
  const int rc = F->fp32 ? build<float>(c, *F, F->bf, F->hf) : build<double>(c, *F, F->bd, F->hd);
  c.ps.bytes = F->bytes;
  return rc;
}

void
fused_free(Context & c)
{
  auto * F = static_cast<FusedState *>(c.fused);
  if (!F)
    return;
  free_mesh_buffers(*F);
  if (F->arena)
    (void)hipFree(F->arena);
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
  return F->fp32 ? step<float>(c, F->bf, F->hf, U, U2, B, B2, dt, s) : step<double>(c, F->bd, F->hd, U, U2, B, B2, dt, s);
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
