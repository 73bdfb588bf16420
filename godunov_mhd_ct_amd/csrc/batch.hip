// batch.hip -- many meshes advanced together through one work list per stage.
//
// A batch holds "slots": fused-form contexts (each with its own mesh, settings and scratch)
// and the two device buffers of each one's state.  One batch step enqueues, for all slots at
// once: the outside-block fill of the slots that have outside blocks, the dt reduction with
// dt = cfl / inv_dt per slot on the device, and the fused stages, one kernel per stage over
// the concatenated work list of every slot and block size.  The step can be captured in a HIP
// graph (one per parity of the ping-pong).  Slots are added and removed between steps; a
// slot's results do not depend on what else is in the batch.
//
// A batch holds HANCOCK contexts or RK2_SSP contexts, not both (the integrator is compiled into
// the cell stage).  An RK2_SSP step is SolverGodunovMHD::godunov_unsplit's RK2 branch
// (SolverGodunovMHD.cpp:551-647) for every slot at once: the outside fill, dt on the device,
// the first stage into a stage buffer the batch owns per slot, the stage-1 combination, the
// outside fill of the stage buffer, the second stage into the other state buffer, and the
// stage-2 combination.  The fused stages write every owned cell and face of their output, so
// the copies U_RK = U and U2 = U_RK of the per-mesh driver are not needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mhdct.h"
#include "ctx.hpp"
#include "fused.hpp"
#include "fused_host.hpp"
#include "handles.hpp"

using mhdct::Context;
using mhdct::FusedBatch;
using mhdct::FusedSlot;
using mhdct::FusedTune;

struct mhdct_batch
{
  struct Entry
  {
    int         id;
    mhdct_ctx * ctx;
    double *    X[2];  // the two state buffers of U; the initial state is in X[0]
    double *    XB[2]; // the same for B
    double *    R;     // RK2_SSP: the stage buffer of U (owned by the batch), else null
    double *    RB;    // ... and of B
    long long   base;  // batch step count when the slot joined
    double      time;  // simulated time and steps, carried across work-list rebuilds
    double      steps;
  };
  int                   dim = 3;
  int                   impl = MHDCT_IMPL_FUSED;
  int                   integrator = -1; // MHDCT_TIME_* of the slots; -1 while the batch is empty
  std::vector<Entry>    ent;
  int                   next_id = 0;
  long long             gstep = 0; // batch steps enqueued; parity = gstep & 1
  bool                  dirty = true;
  FusedBatch<double>    bd;
  FusedBatch<float>     bf;
  std::vector<void *>   wl; // work-list allocations
  FusedTune             tune;
  bool                  use_graph = false;
  hipGraphExec_t        gx[2] = { nullptr, nullptr };
  hipStream_t           cap = nullptr; // capture stream (graphs cannot be captured on the null stream)
  hipStream_t           last = nullptr; // the stream of the last steps (the clocks are final once it drains)
  bool                  ran = false;
  bool                  fork = false;  // run k_edge on a side stream beside the flux stages
  hipStream_t           side = nullptr;
  hipEvent_t            ev_fork = nullptr, ev_join = nullptr;
  std::string           err;
};

namespace
{

bool
fp32(const mhdct_batch * b)
{
  return b->impl == MHDCT_IMPL_FUSED_FP32;
}

void
drop_graphs(mhdct_batch * b)
{
  for (auto & g : b->gx)
    if (g)
    {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void
free_worklist(mhdct_batch * b)
{
  for (void * p : b->wl)
    (void)hipFree(p);
  b->wl.clear();
}

void
free_stage_buffers(mhdct_batch::Entry & e)
{
  if (e.R)
    (void)hipFree(e.R);
  if (e.RB)
    (void)hipFree(e.RB);
  e.R = nullptr;
  e.RB = nullptr;
}

// the stage coefficients (c0, c1) of y = c0 x + c1 y after each stage of the two-stage SSP
// Runge-Kutta scheme (Shu & Osher 1988; SolverGodunovMHD.cpp:583-588, 617-622)
constexpr double kRk2Stage1[2] = { 0.0, 1.0 };
constexpr double kRk2Stage2[2] = { 0.5, 0.5 };

// the slot's current state buffers at batch parity p (slots step together from their base)
int
cur_index(const mhdct_batch::Entry & e, int p)
{
  return p ^ int(e.base & 1);
}

template <typename T>
int
rebuild(mhdct_batch * b, FusedBatch<T> & bt, hipStream_t st)
{
  // the old work list may still be read by enqueued kernels
  if (hipStreamSynchronize(st) != hipSuccess)
  {
    b->err = "stream synchronisation failed";
    return -1;
  }
  // (the entries hold the clocks: save_clocks ran before the entry list changed)
  drop_graphs(b);
  free_worklist(b);
  // slots ordered by block size, so that each block-size class is a range of slots
  std::stable_sort(b->ent.begin(), b->ent.end(), [](const mhdct_batch::Entry & x, const mhdct_batch::Entry & y) {
    return x.ctx->c.b < y.ctx->c.b;
  });
  bt = FusedBatch<T>();
  bt.dim = b->dim;
  bt.nslots = int(b->ent.size());
  bt.bsel = b->ent.empty() ? 0 : b->ent[0].ctx->c.b;
  for (const auto & e : b->ent)
  {
    if (e.ctx->c.b != bt.bsel)
      bt.bsel = 0;
    if (e.ctx->c.st.gravity_enabled)
      bt.grav = 1;
  }
  bt.tune = b->tune;
  bt.rk2 = b->integrator == MHDCT_TIME_RK2_SSP ? 1 : 0;
  bt.dev_dt = 1;
  if (bt.nslots == 0)
  {
    b->dirty = false;
    return 0;
  }
  auto alloc = [&](size_t bytes, void ** p) -> int {
    if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess)
    {
      b->err = "hipMalloc failed for the batch work list";
      return -1;
    }
    b->wl.push_back(*p);
    return 0;
  };
  const int narr = bt.rk2 ? 4 : 2;
  std::vector<FusedSlot<T>> hs[4];
  for (int p = 0; p < narr; ++p)
    hs[p].resize(b->ent.size());
  std::vector<long> off[mhdct::kStages];
  for (auto & o : off)
    o.assign(b->ent.size() + 1, 0);
  for (size_t i = 0; i < b->ent.size(); ++i)
  {
    const auto &         e = b->ent[i];
    const FusedSlot<T> * s = mhdct::fused_slot<T>(e.ctx->c);
    for (int p = 0; p < 2; ++p)
    {
      FusedSlot<T> v = *s;
      const int    c = cur_index(e, p);
      v.U = e.X[c];
      v.U2 = bt.rk2 ? e.R : e.X[1 - c];
      v.B = e.XB[c];
      v.B2 = bt.rk2 ? e.RB : e.XB[1 - c];
      v.dt = 0.0;
      v.dt_red = 1; // the first (or only) stage reduces dt in its primitive stage
      hs[p][i] = v;
      if (bt.rk2)
      {
        // the second stage: from the stage buffers into the other state buffers, with the dt
        // of the first
        v.dt_red = 0;
        v.U = e.R;
        v.U2 = e.X[1 - c];
        v.B = e.RB;
        v.B2 = e.XB[1 - c];
        hs[2 + p][i] = v;
      }
    }
    for (int k = 0; k < mhdct::kStages; ++k)
      off[k][i + 1] = off[k][i] + mhdct::fused_pad_slot(s->cnt[k]);
  }
  for (int p = 0; p < narr; ++p)
  {
    if (alloc(hs[p].size() * sizeof(FusedSlot<T>), reinterpret_cast<void **>(&bt.d_slots[p])) ||
        hipMemcpy(bt.d_slots[p], hs[p].data(), hs[p].size() * sizeof(FusedSlot<T>), hipMemcpyHostToDevice) !=
          hipSuccess)
      return -1;
  }
  for (int k = 0; k < mhdct::kStages; ++k)
  {
    bt.tot[k] = off[k].back();
    if (alloc(off[k].size() * sizeof(long), reinterpret_cast<void **>(&bt.d_off[k])) ||
        hipMemcpy(bt.d_off[k], off[k].data(), off[k].size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
      return -1;
  }
  // the block-size classes and their work lists (each class's offsets from its first slot)
  if (bt.bsel == 0)
  {
    for (int i = 0; i < bt.nslots;)
    {
      int j = i;
      while (j < bt.nslots && b->ent[j].ctx->c.b == b->ent[i].ctx->c.b)
        ++j;
      if (bt.ncls == 3)
      {
        b->err = "more than three block sizes in one batch";
        return -1;
      }
      mhdct::FusedClass & c = bt.cls[bt.ncls++];
      c.bs = b->ent[i].ctx->c.b;
      c.s0 = i;
      c.n = j - i;
      for (int k = 0; k < mhdct::kStages; ++k)
      {
        std::vector<long> o(size_t(c.n) + 1);
        for (int q = 0; q <= c.n; ++q)
          o[q] = off[k][i + q] - off[k][i];
        c.tot[k] = o.back();
        if (alloc(o.size() * sizeof(long), reinterpret_cast<void **>(&c.d_off[k])) ||
            hipMemcpy(c.d_off[k], o.data(), o.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
          return -1;
      }
      i = j;
    }
  }
  std::vector<double> clk(size_t(2) * bt.nslots);
  for (size_t i = 0; i < b->ent.size(); ++i)
  {
    clk[2 * i] = b->ent[i].time;
    clk[2 * i + 1] = b->ent[i].steps;
  }
  if (alloc(clk.size() * sizeof(double), reinterpret_cast<void **>(&bt.clock)) ||
      hipMemcpy(bt.clock, clk.data(), clk.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return -1;
  b->dirty = false;
  return 0;
}

// read the device clocks into the entries (before the entry list changes)
template <typename T>
int
save_clocks(mhdct_batch * b, FusedBatch<T> & bt)
{
  if (!bt.clock || bt.nslots == 0 || b->dirty)
    return 0;
  std::vector<double> clk(size_t(2) * bt.nslots);
  // only this batch's steps write its clocks: wait for its stream(s), not for the whole device
  // (another batch may be running beside it)
  const bool drained = !b->ran || (hipStreamSynchronize(b->last) == hipSuccess &&
                                   (!b->side || hipStreamSynchronize(b->side) == hipSuccess));
  if (!drained || hipMemcpy(clk.data(), bt.clock, clk.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
  {
    b->err = "clock readback failed";
    return -1;
  }
  for (int i = 0; i < bt.nslots && i < int(b->ent.size()); ++i)
  {
    b->ent[i].time = clk[2 * i];
    b->ent[i].steps = clk[2 * i + 1];
  }
  return 0;
}

template <typename T>
int
enqueue(mhdct_batch * b, FusedBatch<T> & bt, int p, hipStream_t st)
{
  for (auto & e : b->ent)
    if (e.ctx->c.dm.noutside > 0)
    {
      const int c = cur_index(e, p);
      if (mhdct::launch_fill_outside(e.ctx->c, e.X[c], e.XB[c], st))
      {
        b->err = "outside fill failed: " + e.ctx->c.err;
        return -1;
      }
    }
  int rc = mhdct::fused_launch_dt<T>(bt, p, st);
  if (rc == 0 && bt.rk2)
  {
    rc = mhdct::fused_launch<T>(bt, p, st); // stage 1: X[c] -> R
    for (auto & e : b->ent)
    {
      const int c = cur_index(e, p);
      rc = rc ? rc
              : mhdct::launch_linear_combination(e.ctx->c, e.X[c], e.R, e.XB[c], e.RB, kRk2Stage1[0], kRk2Stage1[1],
                                                 st);
      if (rc == 0 && e.ctx->c.dm.noutside > 0)
        rc = mhdct::launch_fill_outside(e.ctx->c, e.R, e.RB, st);
    }
    rc = rc ? rc : mhdct::fused_launch<T>(bt, 2 + p, st); // stage 2: R -> X[1-c]
    for (auto & e : b->ent)
    {
      const int c = cur_index(e, p);
      rc = rc ? rc
              : mhdct::launch_linear_combination(e.ctx->c, e.X[c], e.X[1 - c], e.XB[c], e.XB[1 - c], kRk2Stage2[0],
                                                 kRk2Stage2[1], st);
    }
  }
  else if (rc == 0 && b->fork)
  {
    // k_edge needs only k_cell's records: it runs on the side stream beside k_border_flux and
    // k_flux_tile, and the EMF correction waits for both branches
    using namespace mhdct;
    for (int k : { kStagePrim, kStageEfield, kStageCell })
      rc = rc ? rc : fused_launch_stage<T>(bt, p, k, st);
    rc = rc ? rc : (hipEventRecord(b->ev_fork, st) != hipSuccess || hipStreamWaitEvent(b->side, b->ev_fork, 0) != hipSuccess);
    rc = rc ? rc : fused_launch_stage<T>(bt, p, kStageEdge, b->side);
    rc = rc ? rc : (hipEventRecord(b->ev_join, b->side) != hipSuccess);
    for (int k : { kStageBorder, kStageTile })
      rc = rc ? rc : fused_launch_stage<T>(bt, p, k, st);
    rc = rc ? rc : (hipStreamWaitEvent(st, b->ev_join, 0) != hipSuccess);
    for (int k : { kStageFix, kStageBupdate })
      rc = rc ? rc : fused_launch_stage<T>(bt, p, k, st);
  }
  else if (rc == 0)
    rc = mhdct::fused_launch<T>(bt, p, st);
  if (rc)
  {
    b->err = std::string("batch launch failed: ") + hipGetErrorString(hipGetLastError());
    return -1;
  }
  return 0;
}

template <typename T>
int
capture(mhdct_batch * b, FusedBatch<T> & bt, int p)
{
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
  {
    b->err = "graph capture could not start";
    return -1;
  }
  const int rc = enqueue(b, bt, p, b->cap);
  if (hipStreamEndCapture(b->cap, &g) != hipSuccess || rc)
  {
    if (g)
      (void)hipGraphDestroy(g);
    if (b->err.empty())
      b->err = "graph capture failed";
    return -1;
  }
  const hipError_t e = hipGraphInstantiate(&b->gx[p], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
  {
    b->gx[p] = nullptr;
    b->err = std::string("graph instantiation failed: ") + hipGetErrorString(e);
    return -1;
  }
  return 0;
}

template <typename T>
int
step(mhdct_batch * b, FusedBatch<T> & bt, int nsteps, hipStream_t st)
{
  if (b->dirty && rebuild<T>(b, bt, st))
    return -1;
  if (bt.nslots == 0)
    return 0;
  b->last = st;
  b->ran = true;
  for (int n = 0; n < nsteps; ++n)
  {
    const int p = int(b->gstep & 1);
    if (b->use_graph)
    {
      if (!b->gx[p] && capture<T>(b, bt, p))
        return -1;
      if (hipGraphLaunch(b->gx[p], st) != hipSuccess)
      {
        b->err = "graph launch failed";
        return -1;
      }
    }
    else if (enqueue<T>(b, bt, p, st))
      return -1;
    ++b->gstep;
  }
  return 0;
}

// Steps with HIP events recorded between the stage launches on `st`: the average milliseconds
// per step of each timed stage go to ms[MHDCT_NUM_STAGES] (outside fill, dt, then the fused
// stages in launch order).
template <typename T>
int
step_timed(mhdct_batch * b, FusedBatch<T> & bt, int nsteps, hipStream_t st, double * ms)
{
  if (b->dirty && rebuild<T>(b, bt, st))
    return -1;
  constexpr int ns = MHDCT_NUM_STAGES;
  for (int k = 0; k < ns; ++k)
    ms[k] = 0.0;
  if (bt.nslots == 0 || nsteps < 1)
    return 0;
  if (bt.rk2)
  {
    b->err = "timed steps are for HANCOCK batches (an RK2_SSP step runs every stage twice)";
    return -1;
  }
  b->last = st;
  b->ran = true;
  // per step: 2 events around each of the 10 stages, on the stream the stage runs on
  std::vector<hipEvent_t> ev(size_t(nsteps) * 2 * ns);
  for (auto & e : ev)
    if (hipEventCreate(&e) != hipSuccess)
    {
      b->err = "event creation failed";
      return -1;
    }
  using namespace mhdct;
  int rc = 0;
  for (int n = 0; n < nsteps && rc == 0; ++n)
  {
    const int    p = int(b->gstep & 1);
    hipEvent_t * E = &ev[size_t(n) * 2 * ns];
    auto         timed = [&](int slot, hipStream_t s, auto && launch) {
      (void)hipEventRecord(E[2 * slot], s);
      rc = rc ? rc : launch(s);
      (void)hipEventRecord(E[2 * slot + 1], s);
    };
    timed(0, st, [&](hipStream_t s) {
      int r = 0;
      for (auto & e : b->ent)
        if (r == 0 && e.ctx->c.dm.noutside > 0)
        {
          const int c = cur_index(e, p);
          r = launch_fill_outside(e.ctx->c, e.X[c], e.XB[c], s);
        }
      return r;
    });
    timed(1, st, [&](hipStream_t s) { return fused_launch_dt<T>(bt, p, s); });
    timed(2, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStagePrim, s); });
    timed(3, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageEfield, s); });
    timed(4, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageCell, s); });
    hipStream_t es = st;
    if (b->fork)
    {
      rc = rc ? rc : (hipEventRecord(b->ev_fork, st) != hipSuccess || hipStreamWaitEvent(b->side, b->ev_fork, 0) != hipSuccess);
      es = b->side;
    }
    timed(7, es, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageEdge, s); });
    if (b->fork)
      rc = rc ? rc : (hipEventRecord(b->ev_join, b->side) != hipSuccess);
    timed(5, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageBorder, s); });
    timed(6, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageTile, s); });
    if (b->fork)
      rc = rc ? rc : (hipStreamWaitEvent(st, b->ev_join, 0) != hipSuccess);
    timed(8, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageFix, s); });
    timed(9, st, [&](hipStream_t s) { return fused_launch_stage<T>(bt, p, kStageBupdate, s); });
    ++b->gstep;
  }
  if (rc == 0 && hipStreamSynchronize(st) == hipSuccess && hipStreamSynchronize(b->side) == hipSuccess)
  {
    for (int n = 0; n < nsteps; ++n)
      for (int k = 0; k < ns; ++k)
      {
        float f = 0.0f;
        (void)hipEventElapsedTime(&f, ev[size_t(n) * 2 * ns + 2 * k], ev[size_t(n) * 2 * ns + 2 * k + 1]);
        ms[k] += double(f) / nsteps;
      }
  }
  else if (rc == 0)
    rc = -1;
  for (auto & e : ev)
    (void)hipEventDestroy(e);
  if (rc)
    b->err = "timed step failed";
  return rc;
}

// time `reps` steps at the current parity without advancing it (the current state is only
// read; the other buffer and the slot dt are overwritten, the clocks are not touched)
template <typename T>
double
trial(mhdct_batch * b, FusedBatch<T> & bt, int reps, hipStream_t st)
{
  double * keep = bt.clock;
  double * scratch = nullptr;
  if (hipMalloc(&scratch, size_t(2) * bt.nslots * sizeof(double)) != hipSuccess)
    return -1.0;
  (void)hipMemset(scratch, 0, size_t(2) * bt.nslots * sizeof(double));
  bt.clock = scratch;
  const int  p = int(b->gstep & 1);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double ms = -1.0;
  if (enqueue<T>(b, bt, p, st) == 0) // warm-up
  {
    (void)hipEventRecord(e0, st);
    int rc = 0;
    for (int r = 0; r < reps && rc == 0; ++r)
      rc = enqueue<T>(b, bt, p, st);
    (void)hipEventRecord(e1, st);
    float f = 0.0f;
    if (rc == 0 && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&f, e0, e1) == hipSuccess)
      ms = double(f) / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  bt.clock = keep;
  (void)hipFree(scratch);
  return ms;
}

template <typename T>
int
autotune(mhdct_batch * b, FusedBatch<T> & bt, int reps, hipStream_t st, int * grid_cap, int * tile_cap, int * tile_z)
{
  if (b->dirty && rebuild<T>(b, bt, st))
    return -1;
  if (bt.nslots == 0)
    return 0;
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int         full = 1 << 20;
  const std::vector<int> grids = { ncu * 4, ncu * 8, ncu * 16, ncu * 32, full };
  const std::vector<int> tiles = { ncu, ncu * 2, ncu * 4, full };
  double            best = -1.0;
  FusedTune         pick = bt.tune;
  const std::vector<int> depths = b->dim == 3 ? std::vector<int>{ 8, 4 } : std::vector<int>{ 8 };
  for (int z : depths)
    for (int g : grids)
      for (int t : tiles)
      {
        bt.tune.grid_cap = g;
        bt.tune.tile_cap = t;
        bt.tune.tile_z = z;
        const double ms = trial<T>(b, bt, reps, st);
        if (ms > 0 && (best < 0 || ms < best))
        {
          best = ms;
          pick = bt.tune;
        }
      }
  if (best < 0)
  {
    b->err = "autotune: no configuration ran";
    bt.tune = b->tune;
    return -1;
  }
  b->tune = pick;
  bt.tune = pick;
  drop_graphs(b);
  if (grid_cap)
    *grid_cap = pick.grid_cap;
  if (tile_cap)
    *tile_cap = pick.tile_cap;
  if (tile_z)
    *tile_z = pick.tile_z;
  return 0;
}

} // namespace

extern "C"
{
  int
  mhdct_batch_create(int dim, int impl_version, mhdct_batch ** out)
  {
    if (!out)
      return -1;
    *out = nullptr;
    if (dim != 2 && dim != 3)
      return -1;
    if (impl_version != MHDCT_IMPL_FUSED && impl_version != MHDCT_IMPL_FUSED_FP32)
      return -1;
    auto * b = new mhdct_batch();
    b->dim = dim;
    b->impl = impl_version;
    if (hipStreamCreate(&b->cap) != hipSuccess || hipStreamCreate(&b->side) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
    {
      delete b;
      return -1;
    }
    *out = b;
    return 0;
  }

  int
  mhdct_batch_add(mhdct_batch * b, mhdct_ctx * ctx, double * U, double * U2, double * B, double * B2, int * slot_id)
  {
    if (!b)
      return -1;
    b->err.clear();
    if (!ctx || !ctx->c.has_mesh)
    {
      b->err = "the context has no mesh";
      return -1;
    }
    if (ctx->c.impl != b->impl || ctx->c.dim != b->dim)
    {
      b->err = "the context's dimension or implementation differs from the batch";
      return -1;
    }
    const int integ = ctx->c.st.time_integrator;
    if (integ != MHDCT_TIME_HANCOCK && integ != MHDCT_TIME_RK2_SSP)
    {
      b->err = "unknown time integrator";
      return -1;
    }
    if (!b->ent.empty() && integ != b->integrator)
    {
      b->err = "a batch holds contexts of one time integrator (HANCOCK or RK2_SSP)";
      return -1;
    }
    if (ctx->c.st.viscosity_enabled)
    {
      // the viscous stage takes dt from the host, a batch keeps dt on the device
      b->err = "a batch does not run viscosity: step the context with mhdct_do_time_step";
      return -1;
    }
    if ((fp32(b) ? static_cast<const void *>(mhdct::fused_slot<float>(ctx->c))
                 : static_cast<const void *>(mhdct::fused_slot<double>(ctx->c))) == nullptr)
    {
      b->err = "the context has no fused slot";
      return -1;
    }
    if (ctx->c.dm.nowned > 0 && (!U || !U2 || !B || !B2))
    {
      b->err = "null state pointer";
      return -1;
    }
    for (const auto & e : b->ent)
      if (e.ctx == ctx)
      {
        b->err = "the context is already in the batch";
        return -1;
      }
    if (fp32(b) ? save_clocks<float>(b, b->bf) : save_clocks<double>(b, b->bd))
      return -1;
    mhdct_batch::Entry e{};
    if (integ == MHDCT_TIME_RK2_SSP)
    {
      // the stage buffers, zeroed so that the outside blocks (refilled before each use) hold
      // finite values
      const mhdct::Shapes S{ ctx->c.dim, ctx->c.b };
      const size_t        nu = size_t(ctx->c.dm.ntotal > 0 ? ctx->c.dm.ntotal : 1) * 5 * S.cells() * sizeof(double);
      const size_t        nb = size_t(ctx->c.dm.ntotal > 0 ? ctx->c.dm.ntotal : 1) * S.fper() * sizeof(double);
      if (hipMalloc(&e.R, nu) != hipSuccess || hipMalloc(&e.RB, nb) != hipSuccess ||
          hipMemset(e.R, 0, nu) != hipSuccess || hipMemset(e.RB, 0, nb) != hipSuccess)
      {
        free_stage_buffers(e);
        b->err = "hipMalloc failed for the RK2_SSP stage buffers";
        return -1;
      }
    }
    e.id = b->next_id++;
    e.ctx = ctx;
    e.X[0] = U;
    e.X[1] = U2;
    e.XB[0] = B;
    e.XB[1] = B2;
    e.base = b->gstep;
    e.time = 0.0;
    e.steps = 0.0;
    b->ent.push_back(e);
    b->integrator = integ;
    b->dirty = true;
    if (slot_id)
      *slot_id = e.id;
    return 0;
  }

  int
  mhdct_batch_remove(mhdct_batch * b, int slot_id)
  {
    if (!b)
      return -1;
    b->err.clear();
    if (fp32(b) ? save_clocks<float>(b, b->bf) : save_clocks<double>(b, b->bd))
      return -1;
    auto it = std::find_if(b->ent.begin(), b->ent.end(), [&](const mhdct_batch::Entry & e) { return e.id == slot_id; });
    if (it == b->ent.end())
    {
      b->err = "no such slot";
      return -1;
    }
    if (it->R && hipDeviceSynchronize() != hipSuccess)
    {
      b->err = "device synchronisation failed";
      return -1;
    }
    free_stage_buffers(*it);
    b->ent.erase(it);
    if (b->ent.empty())
      b->integrator = -1;
    b->dirty = true;
    return 0;
  }

  int
  mhdct_batch_step(mhdct_batch * b, int nsteps, void * stream)
  {
    if (!b)
      return -1;
    b->err.clear();
    hipStream_t st = static_cast<hipStream_t>(stream);
    return fp32(b) ? step<float>(b, b->bf, nsteps, st) : step<double>(b, b->bd, nsteps, st);
  }

  int
  mhdct_batch_step_timed(mhdct_batch * b, int nsteps, void * stream, double * stage_ms)
  {
    if (!b || !stage_ms)
      return -1;
    b->err.clear();
    hipStream_t st = static_cast<hipStream_t>(stream);
    return fp32(b) ? step_timed<float>(b, b->bf, nsteps, st, stage_ms)
                   : step_timed<double>(b, b->bd, nsteps, st, stage_ms);
  }

  const char *
  mhdct_stage_name(int stage)
  {
    static const char * names[MHDCT_NUM_STAGES] = { "fill_outside", "dt",  "k_prim", "k_efield", "k_cell",
                                                    "k_border_flux", "k_flux_tile", "k_edge", "k_fix+k_scatter",
                                                    "k_bupdate" };
    return (stage >= 0 && stage < MHDCT_NUM_STAGES) ? names[stage] : "";
  }

  int
  mhdct_batch_slot_state(mhdct_batch * b, int slot_id, int * current, double * time, int64_t * steps)
  {
    if (!b)
      return -1;
    b->err.clear();
    if (fp32(b) ? save_clocks<float>(b, b->bf) : save_clocks<double>(b, b->bd))
      return -1;
    for (const auto & e : b->ent)
      if (e.id == slot_id)
      {
        const long long n = (long long)e.steps;
        if (current)
          *current = int(n & 1);
        if (time)
          *time = e.time;
        if (steps)
          *steps = n;
        return 0;
      }
    b->err = "no such slot";
    return -1;
  }

  int
  mhdct_batch_check(mhdct_batch * b, int * bad_slot)
  {
    if (!b)
      return -1;
    b->err.clear();
    if (bad_slot)
      *bad_slot = -1;
    if (hipDeviceSynchronize() != hipSuccess)
    {
      b->err = "device synchronisation failed";
      return -1;
    }
    for (const auto & e : b->ent)
    {
      unsigned long long f = 0;
      if (hipMemcpy(&f, e.ctx->c.red + 1, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess)
      {
        b->err = "flag readback failed";
        return -1;
      }
      if (f)
      {
        if (bad_slot)
          *bad_slot = e.id;
        b->err = (f & 2ull) ? "NaN state met in the dt reduction" : "negative internal energy detected";
        return (f & 2ull) ? -3 : -2;
      }
    }
    return 0;
  }

  int
  mhdct_batch_set_graph(mhdct_batch * b, int enable)
  {
    if (!b)
      return -1;
    b->use_graph = enable != 0;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_set_overlap(mhdct_batch * b, int enable)
  {
    if (!b)
      return -1;
    b->fork = enable != 0;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_set_tune(mhdct_batch * b, int grid_cap, int tile_cap, int tile_z)
  {
    if (!b || grid_cap < 8 || tile_cap < 8 || (tile_z != 8 && tile_z != 4))
      return -1;
    b->tune.grid_cap = grid_cap;
    b->tune.tile_cap = tile_cap;
    b->tune.tile_z = tile_z;
    b->bd.tune = b->tune;
    b->bf.tune = b->tune;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_set_class_launch(mhdct_batch * b, int enable)
  {
    if (!b)
      return -1;
    b->tune.per_class = enable != 0;
    b->bd.tune = b->tune;
    b->bf.tune = b->tune;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_set_stage_threads(mhdct_batch * b, int cell_threads, int edge_threads)
  {
    if (!b)
      return -1;
    if (!(cell_threads == 256 || cell_threads == 512) ||
        !(edge_threads == 256 || edge_threads == 512))
    {
      b->err = "mhdct_batch_set_stage_threads: k_cell and k_edge 256 or 512 threads";
      return -1;
    }
    b->tune.cell_threads = cell_threads;
    b->tune.edge_threads = edge_threads;
    b->bd.tune = b->tune;
    b->bf.tune = b->tune;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_set_front(mhdct_batch * b, int enable)
  {
    if (!b)
      return -1;
    b->tune.front = enable != 0;
    b->bd.tune = b->tune;
    b->bf.tune = b->tune;
    drop_graphs(b);
    return 0;
  }

  int
  mhdct_batch_autotune(mhdct_batch * b, int reps, void * stream, int * grid_cap, int * tile_cap, int * tile_z)
  {
    if (!b || reps < 1)
      return -1;
    b->err.clear();
    hipStream_t st = static_cast<hipStream_t>(stream);
    return fp32(b) ? autotune<float>(b, b->bf, reps, st, grid_cap, tile_cap, tile_z)
                   : autotune<double>(b, b->bd, reps, st, grid_cap, tile_cap, tile_z);
  }

  int
  mhdct_batch_num_slots(const mhdct_batch * b)
  {
    return b ? int(b->ent.size()) : 0;
  }

  const char *
  mhdct_batch_last_error(const mhdct_batch * b)
  {
    return b ? b->err.c_str() : "null batch";
  }

  void
  mhdct_batch_destroy(mhdct_batch * b)
  {
    if (!b)
      return;
    (void)hipDeviceSynchronize();
    drop_graphs(b);
    free_worklist(b);
    for (auto & e : b->ent)
      free_stage_buffers(e);
    if (b->cap)
      (void)hipStreamDestroy(b->cap);
    if (b->side)
      (void)hipStreamDestroy(b->side);
    if (b->ev_fork)
      (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join)
      (void)hipEventDestroy(b->ev_join);
    delete b;
  }
}
